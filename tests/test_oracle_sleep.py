"""The oracle's sleep partition against the reference's rules
(solver/particles.rs:227-312, step.rs:107-131,299-315): the resting body
falls asleep after one substep and the sleeping particles form the tail of
the arrays; sleeping particles do not move; the thrown body's arrival wakes
its neighbours; add_body inserts in front of the sleeping zone; impulses,
sleep_tag and wake_tag move particles across the boundary."""
import numpy as np

import emerge_hip as eh
import scenes as S


def _partition_ok(sim):
    p = sim.particles()
    k = sim.active_count()
    return bool(np.all(p["sleeping"][:k] == 0) and np.all(p["sleeping"][k:] == 1)), k, p


def test_resting_body_sleeps_and_forms_the_tail(oracle):
    o = S.sleep_collide("oracle32")
    o.step_n(1)
    ok, k, p = _partition_ok(o)
    assert ok
    rest = p["user_tag"] == 1
    assert np.all(p["sleeping"][rest] == 1) and np.all(p["sleeping"][~rest] == 0)
    assert k == int((~rest).sum())
    x0 = S.by_key(p)["x"].copy()
    o.step_n(3)
    q = S.by_key(o.particles())
    asleep = q["sleeping"] == 1
    assert asleep.sum() > 0 and np.array_equal(q["x"][asleep], x0[asleep])  # frozen while asleep


def test_arrival_wakes_neighbours(oracle):
    o = S.sleep_collide("oracle32")
    o.step_n(1)
    n_rest = int((o.particles()["user_tag"] == 1).sum())
    woke = False
    for _ in range(60):
        o.step_n(1)
        ok, k, p = _partition_ok(o)
        assert ok
        if int(p["sleeping"].sum()) < n_rest:
            woke = True
            break
    assert woke


def test_add_body_goes_in_front_of_the_sleepers(oracle):
    o = S.sleep_collide("oracle32")
    o.step_n(1)
    k0 = o.active_count()
    tag = o.add_body(eh.SpawnRegion(spacing=0.5, box_size=(4, 4), box_center=(50.0, 50.0), rng_seed=9))
    ok, k, p = _partition_ok(o)
    assert ok and k == k0 + int((p["user_tag"] == tag).sum())
    assert np.all(np.flatnonzero(p["user_tag"] == tag) < k)


def test_impulse_and_tags(oracle):
    o = S.sleep_collide("oracle32")
    o.step_n(1)
    o.apply_impulse((31.0, 24.0), 2.0, (1.0, 0.0))
    ok, k, p = _partition_ok(o)
    assert ok
    near = np.hypot(p["x"][:, 0] - 31.0, p["x"][:, 1] - 24.0) <= 2.0
    assert near.sum() > 4 and np.all(p["sleeping"][near] == 0)
    o.sleep_tag(2)
    ok, k, p = _partition_ok(o)
    assert ok and np.all(p["sleeping"][p["user_tag"] == 2] == 1)
    o.wake_tag(1)
    ok, k, p = _partition_ok(o)
    assert ok and np.all(p["sleeping"][p["user_tag"] == 1] == 0)
