"""Sleep partition on the GPU against the oracle (solver/particles.rs:227-312,
step.rs:107-131,299-315), particles matched by their muscle_group_id key:
the same particles sleep and wake, sleeping particles stay frozen, the
adaptive substep schedule (CFL over the awake particles) matches, impulses
and sleep_tag / wake_tag move the same particles, and the diagnostics count
the sleepers."""
import numpy as np
import pytest

import emerge_hip as eh
import scenes as S


def _flags(p):
    return p["sleeping"].astype(np.int64)


@pytest.mark.gpu
@pytest.mark.parametrize("reorder", [None, 8], ids=["upload_reorder", "reorder8"])
def test_sleep_matches_oracle(reorder, oracle):
    g = S.sleep_collide("gpu")
    if reorder is not None:
        g.set_reorder_interval(reorder)  # sleeping particles move slots with the rest
    o = S.sleep_collide("oracle32")
    n = g.num_particles()
    done = 0
    prev = None
    for steps in (1, 2, 10, 25, 45):
        g.step_n(steps - done)
        o.step_n(steps - done)
        done = steps
        pg, po = S.by_key(g.particles()), S.by_key(o.particles())
        mism = int(np.sum(_flags(pg) != _flags(po)))
        e = S.rel_errors(S.state_matrix(pg), S.state_matrix(po))
        sleeping = int(pg["sleeping"].sum())
        print(f"after {steps}: sleeping gpu {sleeping} oracle {int(po['sleeping'].sum())}, flag mismatches {mism}, "
              f"x {e['x']:.2e} v {e['v']:.2e}, substeps {g.stats().total_substeps}/{o.stats().total_substeps}")
        assert mism <= max(2, n // 100), mism
        assert e["x"] < 1e-4, e
        assert g.stats().total_substeps == o.stats().total_substeps
        assert g.diagnostics_snapshot()["sleeping_count"] == sleeping
        if steps == 1:
            assert sleeping == int((pg["user_tag"] == 1).sum())  # the resting body, all of it
        full = int((pg["user_tag"] == 1).sum())
        if prev is not None and int(prev["sleeping"].sum()) == full and sleeping == full:
            # nothing woke in the interval: the sleeping body has not moved (a particle
            # can otherwise wake, move and fall asleep again between two checkpoints)
            still = pg["sleeping"] == 1
            assert np.array_equal(prev["x"][still], pg["x"][still])
        prev = pg.copy()
    assert int(prev["sleeping"].sum()) < int((prev["user_tag"] == 1).sum())  # the arrival woke some
    g.close()


@pytest.mark.gpu
def test_impulse_and_tags_match_oracle(oracle):
    g = S.sleep_collide("gpu")
    o = S.sleep_collide("oracle32")
    for sim in (g, o):
        sim.step_n(1)
        sim.apply_impulse((31.0, 24.0), 2.0, (1.0, 0.0))
    pg, po = S.by_key(g.particles()), S.by_key(o.particles())
    near = np.hypot(po["x"][:, 0] - 31.0, po["x"][:, 1] - 24.0) <= 2.0
    assert np.array_equal(_flags(pg), _flags(po)) and near.sum() > 4 and np.all(pg["sleeping"][near] == 0)
    assert np.array_equal(pg["v"][near], po["v"][near])  # at rest before: one f32 expression each
    for sim in (g, o):
        sim.sleep_tag(2)
        sim.wake_tag(1)
        sim.step_n(2)
    pg, po = S.by_key(g.particles()), S.by_key(o.particles())
    assert np.array_equal(_flags(pg), _flags(po))
    assert np.all(pg["sleeping"][pg["user_tag"] == 2] == 1)  # the thrown body, frozen in flight
    assert g.stats().total_substeps == o.stats().total_substeps
    g.close()


@pytest.mark.gpu
def test_sleep_excludes_other_modes():
    sim = S.sleep_collide("gpu")
    with pytest.raises(eh.EmergeError) as ei:
        eh.SceneBatch([sim, S.corotated_drop("gpu")])
    assert ei.value.code == -2
    sim.close()
