"""Batched scenes: many independent simulations stepped together through one
launch per stage.  Each scene must take exactly the substep schedule it takes
alone (its own CFL dt) and reach the same state up to f32 roundoff of the
grid sums; edits between batch steps, and a scene stepped on its own, bring
the batch back in step."""
import numpy as np
import pytest

import emerge_hip as eh
import scenes as S

# well-conditioned scenes of different grid sizes, models and boundaries;
# quickstart and multi_material carry a Newtonian fluid (density pass)
BATCH_SCENES = ["quickstart", "sand_column", "snow_drop", "corotated_drop", "multi_material", "terrain_slide",
                "viscoelastic_drop"]


def _close(sims):
    for s in sims:
        s.close()


def _x_err(a, b):
    return float(np.abs(a["x"].astype(np.float64) - b["x"]).max())


@pytest.mark.gpu
def test_batch_matches_scenes_stepped_alone():
    alone = [S.SCENES[n].build("gpu") for n in BATCH_SCENES]
    batched = [S.SCENES[n].build("gpu") for n in BATCH_SCENES]
    batch = eh.SceneBatch(batched)
    for steps in (1, 9):
        for s in alone:
            s.step_n(steps)
        batch.step_n(steps)
        for name, a, b in zip(BATCH_SCENES, alone, batched):
            sa, sb = a.stats(), b.stats()
            assert sb.total_substeps == sa.total_substeps and sb.last_substeps == sa.last_substeps, (name, sa, sb)
            assert sb.frame_index == sa.frame_index
            pa, pb = a.particles(), b.particles()
            assert np.array_equal(pa["mass"], pb["mass"]) and np.array_equal(pa["user_tag"], pb["user_tag"])
            e = S.rel_errors(S.state_matrix(pb), S.state_matrix(pa))
            print(f"{name}: x {e['x']:.2e} v {e['v']:.2e} F {e['F']:.2e}")
            assert e["x"] < 2e-5 and e["F"] < 1e-3, (name, e)
    batch.close()
    _close(alone + batched)


@pytest.mark.gpu
def test_batch_resyncs_after_edits():
    """An impulse on one scene and an individual step of another between batch
    steps: the next batch step restarts every loop and still matches."""
    names = ["corotated_drop", "snow_drop", "sand_column"]
    alone = [S.SCENES[n].build("gpu") for n in names]
    batched = [S.SCENES[n].build("gpu") for n in names]
    batch = eh.SceneBatch(batched)
    batch.step_n(2)
    for s in alone:
        s.step_n(2)
    for group in (alone, batched):
        group[0].apply_impulse((24.0, 9.0), 5.0, (20.0, 40.0))
        group[1].step_n(1)
    batch.step_n(3)
    alone[0].step_n(3)
    alone[1].step_n(3)
    alone[2].step_n(3)
    for name, a, b in zip(names, alone, batched):
        assert b.stats().total_substeps == a.stats().total_substeps, name
        e = S.rel_errors(S.state_matrix(b.particles()), S.state_matrix(a.particles()))
        assert e["x"] < 2e-5, (name, e)
    batch.close()
    _close(alone + batched)


@pytest.mark.gpu
def test_batch_follows_gravity_change():
    """set_gravity on a batched scene between batch steps applies from the next batch
    step, as it does on the scene stepped alone."""
    names = ["corotated_drop", "sand_column"]
    alone = [S.SCENES[n].build("gpu") for n in names]
    batched = [S.SCENES[n].build("gpu") for n in names]
    batch = eh.SceneBatch(batched)
    batch.step_n(2)
    for s in alone:
        s.step_n(2)
    for group in (alone, batched):
        group[0].set_gravity((3.0, -30.0))
    batch.step_n(4)
    for s in alone:
        s.step_n(4)
    for name, a, b in zip(names, alone, batched):
        assert b.stats().total_substeps == a.stats().total_substeps, name
        e = S.rel_errors(S.state_matrix(b.particles()), S.state_matrix(a.particles()))
        assert e["x"] < 2e-5 and e["F"] < 1e-3, (name, e)
    # the changed scene really moved differently from the unchanged one
    ref = S.SCENES[names[0]].build("gpu")
    ref.step_n(6)
    assert float(np.abs(ref.particles()["x"] - batched[0].particles()["x"]).max()) > 1e-2
    batch.close()
    _close(alone + batched + [ref])


@pytest.mark.gpu
def test_batch_errors():
    sims = [S.corotated_drop("gpu"), S.snow_drop("gpu")]
    with pytest.raises(eh.EmergeError) as ei:
        eh.SceneBatch([sims[0], sims[0]])
    assert ei.value.code == -1
    empty = eh.Simulation.empty(eh.SimConfig.standard(64, 0.02, (0.0, -9.81)))
    with pytest.raises(eh.EmergeError) as ei:
        eh.SceneBatch([sims[0], empty])
    assert ei.value.code == -1
    sg = eh.NaccSurrogate(0)
    nacc = S.nacc_block("gpu")
    nacc.set_surrogate_mode(sg, "surrogate")
    with pytest.raises(eh.EmergeError) as ei:
        eh.SceneBatch([sims[1], nacc])
    assert ei.value.code == -2
    nacc.set_surrogate_mode(None, "plain")
    # a scene switched to a surrogate mode after batching fails the next batch step
    batch = eh.SceneBatch([sims[1], nacc])
    batch.step_n(1)
    nacc.set_surrogate_mode(sg, "surrogate")
    with pytest.raises(eh.EmergeError) as ei:
        batch.step_n(1)
    assert ei.value.code == -2
    nacc.set_surrogate_mode(None, "plain")
    batch.step_n(1)
    batch.close()
    _close(sims + [empty, nacc])
    sg.close()


@pytest.mark.gpu
def test_batch_profiling_counts_stages():
    """Per-stage timing of a batch: one begin_step per batch step, one sort on the
    first step (every scene's upload reorder), substep launch slots at least the
    longest scene's substeps, a control gather per drain; the timed steps take the
    same schedules as an untimed batch of the same scenes; the autotuner refuses to
    run while profiling is on."""
    names = ["quickstart", "sand_column", "multi_material"]
    a = [S.SCENES[k].build("gpu") for k in names]
    b = [S.SCENES[k].build("gpu") for k in names]
    ba, bb = eh.SceneBatch(a), eh.SceneBatch(b)
    bb.set_profiling(True)
    ba.step_n(3)
    bb.step_n(3)
    t = bb.kernel_times()
    print(t)
    assert set(t) == {"kb_begin_step", "kb_sort", "kb_substep", "kb_gather_ctrl"}
    assert t["kb_begin_step"][0] == 3
    assert t["kb_sort"][0] >= 1
    assert t["kb_substep"][0] >= max(s.stats().total_substeps for s in b)
    assert t["kb_gather_ctrl"][0] >= 1
    assert all(ms > 0.0 for _, ms in t.values())
    for x, y in zip(a, b):
        assert x.stats().total_substeps == y.stats().total_substeps
    with pytest.raises(eh.EmergeError):
        bb.autotune(1)
    bb.set_profiling(False)
    assert all(v == (0, 0.0) for v in bb.kernel_times().values())
    ba.close()
    bb.close()
    for s in a + b:
        s.close()
