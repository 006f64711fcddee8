"""GPU substep vs the CPU oracle (f32 restatement of Simulation::step) and the
FP64 oracle, per particle, after 1 step and after the scene's N steps.

Criterion: the GPU is as accurate as the reference's own f32 arithmetic.  The
f32 error of the reference algorithm is measured as the largest error against
the FP64 oracle of thirteen f32 oracle runs: particles in the scene's order,
and in twelve seeded random orders (the grid sums then add in other orders,
as the GPU's atomics do).  Scenes with plasticity thresholds are
ill-conditioned at f32 roundoff: one NACC step spreads C by 7e-4 .. 1.4e-2
over particle orders alone, ten Von Mises steps spread v by 1e-5 .. 4e-5, so
a few orders underestimate the reference's spread.  For each field, the GPU's error against the FP64 oracle
must stay within RATIO x that f32 error (plus an absolute floor of a few f32
ulp), and after one step the GPU must also be within the stated relative
tolerance of the f32 oracle directly.  Errors are relative to
the field's own magnitude (scenes.rel_errors).  Scenes whose dynamics are
chaotic over N steps (quickstart_impact: fluid splash) are checked after one
step and, after N steps, only through invariants.
"""
import numpy as np
import pytest

import scenes as S

RATIO = 8.0
FLOOR = 2e-6
PERM_SEEDS = tuple(range(1, 13))
# One step, GPU vs f32 oracle.  Set from the measured spread of the f32 oracle
# itself against the FP64 oracle on these scenes (the reference algorithm
# amplifies f32 roundoff most in the Drucker-Prager / NACC log-strain terms).
ONE_STEP_TOL = {"x": 1e-6, "v": 2e-3, "C": 5e-3, "F": 1e-4, "Jp": 1e-5, "q": 1e-4, "lvs": 2e-2, "density": 2e-5}
CHAOTIC = {"quickstart_impact"}
# Fields whose law is discontinuous, compared by counting particles instead of
# by the largest error.  mu(I) stores mu(I) in friction_hardening as a
# diagnostic that nothing reads back; it jumps from mu_static (p_trial <= 0,
# sand_mui.rs:158-167) to nearly mu_dynamic as p_trial -> 0+ (sand_mui.rs:
# 176-200), so free-surface particles at p_trial ~ 0 flip on summation order
# alone.  The GPU may flip at most a few times as many particles as the f32
# oracle does in any of its particle orders, plus 2 % of the particles.
FLIP_FIELDS = {"mui_column": ("q", 1e-3)}
# Scenes whose plastic branch is a threshold test (Von Mises: yield when the
# trial deviatoric stress exceeds the current yield stress, von_mises.rs:
# 108-141).  A particle sitting on the threshold takes either branch
# depending on summation order; once it flips, its v, C, F and q leave the
# f32 spread for good (one flip in about a hundred GPU runs of ductile_drop
# moved v by 1.8e-3 against a spread of 4e-5; 60 oracle orders and 40 GPU
# runs without a flip stay below 1e-4).  After N steps, at most
# max(3, 2 %) particles may exceed the bound, each staying finite and within
# 100 x the bound; every other particle must meet the bound.
BRANCH_SCENES = {"ductile_drop"}


def _per_particle(a, ref, field):
    """Per-particle errors on the field's own scale (scenes.rel_errors' norm)."""
    floors = {"x": 1.0, "v": 1e-3, "C": 1e-3, "F": 1.0, "Jp": 1.0, "q": 1e-3, "lvs": 1e-3, "density": 1e-6}
    d = np.abs(np.asarray(a, np.float64) - ref).reshape(len(ref), -1).max(axis=1)
    return d / max(float(np.max(np.abs(ref))) if ref.size else 0.0, floors[field])


def _flips(a, ref, tol):
    return int(np.sum(np.abs(np.asarray(a, np.float64) - ref) > tol))


def _run(scene, backend, steps, perm_seed=None, reorder=None):
    sim = scene.build(backend)
    if reorder is not None and backend == "gpu":
        sim.set_reorder_interval(reorder)  # physical reorders mid-run, not just at the upload
    inv = None
    if perm_seed is not None:  # same particles, another order
        p = sim.particles()
        perm = np.random.default_rng(perm_seed).permutation(len(p))
        sim.set_particles(p[perm])
        inv = np.argsort(perm)
    out = []
    done = 0
    for s in steps:
        sim.step_n(s - done)
        done = s
        if backend == "oracle64":
            st = S.f64_state(sim.particles_f64())
        else:
            st = S.state_matrix(sim.particles())
        if inv is not None:
            st = {k: v[inv] for k, v in st.items()}
        out.append(st)
    return out, sim


# every scene as configured (the physical reorder at the upload, then every
# 4096 substeps), plus two scenes reordering every 8 substeps
PARITY_CASES = [(n, None) for n in S.SCENES] + [("multi_material", 8), ("sand_column", 8)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,reorder", PARITY_CASES,
                         ids=[n if r is None else f"{n}-reorder{r}" for n, r in PARITY_CASES])
def test_scene_parity(name, reorder, oracle, capsys):
    scene = S.SCENES[name]
    steps = [1, scene.steps]
    g, gsim = _run(scene, "gpu", steps, reorder=reorder)
    o32, osim = _run(scene, "oracle32", steps)
    o32p = [_run(scene, "oracle32", steps, perm_seed=seed)[0] for seed in PERM_SEEDS]
    o64, _ = _run(scene, "oracle64", steps)
    for k, s in enumerate(steps):
        e_g32 = S.rel_errors(g[k], o32[k])
        e_g64 = S.rel_errors(g[k], o64[k])
        e_a = S.rel_errors(o32[k], o64[k])
        e_p = [S.rel_errors(x[k], o64[k]) for x in o32p]
        e_b = {f: max(e[f] for e in e_p) for f in e_a}
        e_3264 = {f: max(e_a[f], e_b[f]) for f in e_a}
        with capsys.disabled():
            print(f"\n[{name} after {s} steps] n={len(g[k]['x'])}")
            for f in ("x", "v", "C", "F", "Jp", "q", "lvs", "density"):
                print(f"  {f:8s} gpu-vs-o32 {e_g32[f]:.3e}  gpu-vs-o64 {e_g64[f]:.3e}  o32-vs-o64 {e_a[f]:.3e}"
                      f"  o32(permuted, max of {len(PERM_SEEDS)})-vs-o64 {e_b[f]:.3e}")
        flip = FLIP_FIELDS.get(name)
        if flip is not None:
            f, tol = flip
            n = len(g[k][f])
            fg = _flips(g[k][f], o64[k][f], tol)
            fo = max(_flips(x[f], o64[k][f], tol) for x in [o32[k]] + [y[k] for y in o32p])
            with capsys.disabled():
                print(f"  {f} flips beyond {tol}: gpu {fg}, f32 oracle (max over orders) {fo} of {n}")
            assert fg <= 4 * fo + 0.02 * n, (name, s, f, fg, fo)
        if k == 0:
            # absolute bound, widened where the reference algorithm itself
            # spreads f32 roundoff further (|gpu - o32| <= |gpu - o64| + |o32 - o64|)
            for f, tol in ONE_STEP_TOL.items():
                if flip is not None and f == flip[0]:
                    continue
                bound = max(tol, (RATIO + 1.0) * e_3264[f] + FLOOR)
                assert e_g32[f] <= bound, (name, s, f, e_g32[f], bound)
        if k == 0 or name not in CHAOTIC:
            for f in ONE_STEP_TOL:
                if flip is not None and f == flip[0]:
                    continue
                bound = RATIO * e_3264[f] + FLOOR
                if k > 0 and name in BRANCH_SCENES:
                    e = _per_particle(g[k][f], o64[k][f], f)
                    allowed = max(3, int(np.ceil(0.02 * len(e))))
                    over = np.sort(e)[::-1]
                    with capsys.disabled():
                        if over[0] > bound:
                            print(f"  {f}: {int(np.sum(e > bound))} particle(s) beyond {bound:.3e}, largest {over[0]:.3e}")
                    assert np.all(np.isfinite(e)) and over[0] <= 100.0 * bound, (name, s, f, over[0], bound)
                    assert over[min(allowed, len(over) - 1)] <= bound, (name, s, f, over[:allowed + 1], bound)
                    continue
                assert e_g64[f] <= bound, (name, s, f, e_g64[f], e_3264[f])
    # invariants that hold for every scene after N steps
    p = gsim.particles()
    assert np.all(np.isfinite(p["x"])) and np.all(np.isfinite(p["deformation_gradient"]))
    assert np.array_equal(p["mass"], osim.particles()["mass"])
    if name == "mui_column":  # mu(I) stays between mu_static and mu_dynamic (sand_mui.rs:190-196)
        mp = S.Mats("gpu").mu_i(1e5, 0.3).params
        q = p["friction_hardening"]
        assert q.min() >= mp.dp_h0 and q.max() <= mp.dp_h1
