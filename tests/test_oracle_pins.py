"""Independent pins for three oracle functions that whole-step parity rests on:

- choose_substep_dt / cfl_bound / affine_cfl_speed_contribution
  (src/spacetime/solver/step.rs:434-488) and the step loop around it (:23-60);
- project_particle_state_to_admissible (step.rs:347-429), with the reference's own
  project_invalid_state_recovers_every_guarded_field (tests/physics_correctness.rs:2237-2300)
  restated on the oracle;
- NaccMaterial::project's hardening update and its Case A / B / C returns
  (src/matter/materials/nacc.rs:129-223).

Every expected value comes from a closed form the mathematics fixes (a speed bound,
an ellipse intersection computed another way, p(J) at the caps, trace and
direction of the deviatoric return), not from the oracle's own formula.  Each
group also plants a plausible mistake in the oracle (oracle_py.mutation) and
shows that its check fails on it.
"""
import math

import numpy as np
import pytest

import emerge_hip as eh


def r32(x):
    """The f32 value a float argument or config field takes when it crosses the C-ABI."""
    return float(np.float32(x))


# --------------------------------------------------------------------- CFL


def _one_particle_sim(oracle, v=(0.0, 0.0), c=(0.0, 0.0, 0.0, 0.0), density=1.0, mat=None, cfg=None, n=1,
                      x=(32.0, 32.0), precision=64):
    cfg = cfg or eh.SimConfig.standard(64, 1.0, (0.0, 0.0))
    sim = oracle.OracleSim.empty(cfg, precision)
    sim.with_default_material(mat or oracle.material("neohookean", 1e-6, 1e-6))
    p = eh.zeroed_particles(n)
    p["x"] = x
    p["v"] = v
    p["velocity_gradient"] = c
    p["mass"] = 1.0
    p["initial_volume"] = 1.0
    p["volume"] = 1.0
    p["density"] = density
    sim.set_particles(p)
    return sim


def test_cfl_speed_bound(oracle):
    """|v| = 5 with C = 0: dt = cfl·dx/|v| = 0.9/5 (material bound far away)."""
    sim = _one_particle_sim(oracle, v=(3.0, 4.0))
    assert sim.choose_substep_dt(1.0) == pytest.approx(0.9 / 5.0, rel=1e-7)  # f32 coefficients
    # remaining time below the bound caps it
    assert sim.choose_substep_dt(0.1) == pytest.approx(r32(0.1), rel=1e-12)


def test_cfl_affine_term(oracle):
    """affine_cfl_speed_contribution: speed += ||C||_F · 1.5·sqrt(2) · dx."""
    c = (0.1, 0.2, 0.3, 0.4)
    speed = 5.0 + math.sqrt(0.01 + 0.04 + 0.09 + 0.16) * 1.5 * math.sqrt(2.0)
    sim = _one_particle_sim(oracle, v=(3.0, 4.0), c=c)
    got = sim.choose_substep_dt(1.0)
    assert got == pytest.approx(0.9 / speed, rel=1e-7)
    # cfl_include_affine_speed = false drops it
    cfg = eh.SimConfig.standard(64, 1.0, (0.0, 0.0)).replace(cfl_include_affine_speed=0)
    assert _one_particle_sim(oracle, v=(3.0, 4.0), c=c, cfg=cfg).choose_substep_dt(1.0) == pytest.approx(0.18)
    # a planted mistake (affine term dropped) fails the check
    with oracle.mutation(oracle.MUT_CFL_NO_AFFINE):
        bad = sim.choose_substep_dt(1.0)
    assert abs(bad - 0.9 / speed) > 1e-3 * (0.9 / speed)


def test_cfl_material_bound_is_elastic_wave_dt(oracle):
    """elastic_wave_dt: c_P = sqrt((λ+2µ)/ρ); NH(400, 200) at ρ = 4 -> dt = 0.5 / sqrt(200)."""
    sim = _one_particle_sim(oracle, v=(0.01, 0.0), density=4.0, mat=oracle.material("neohookean", 400.0, 200.0))
    assert sim.choose_substep_dt(1.0) == pytest.approx(0.5 / math.sqrt(200.0), rel=1e-7)


def test_cfl_fastest_particle_and_ignored_bounds(oracle):
    """max over particles of the speed, min of finite positive material bounds;
    a zero-stiffness material (infinite bound) is ignored, as is rest (|v| <= eps)."""
    sim = _one_particle_sim(oracle, n=3)
    p = sim.particles()
    p["v"] = [(0.0, 1.0), (0.0, -6.0), (2.0, 0.0)]
    p["x"] = [(20.0, 20.0), (30.0, 30.0), (40.0, 40.0)]
    sim.set_particles(p)
    assert sim.choose_substep_dt(1.0) == pytest.approx(0.15, rel=1e-7)
    rest = _one_particle_sim(oracle)
    assert rest.choose_substep_dt(0.7) == pytest.approx(r32(0.7), rel=1e-12)


def test_cfl_clamps(oracle):
    """cfl_bound: dt clamped to [min(min_dt, max_dt), max_dt]; min_dt = 1e-3."""
    fast = _one_particle_sim(oracle, v=(1.0e4, 0.0))
    assert fast.choose_substep_dt(1.0) == pytest.approx(r32(1.0e-3), rel=1e-12)       # 9e-5 raised to min_dt
    assert fast.choose_substep_dt(5.0e-4) == pytest.approx(r32(5.0e-4), rel=1e-12)    # min_dt above the time left
    with oracle.mutation(oracle.MUT_CFL_NO_MIN_CLAMP):
        bad = fast.choose_substep_dt(1.0)
    assert bad == pytest.approx(9.0e-5, rel=1e-6)


def test_cfl_fixed_dt(oracle):
    """adaptive_timestep = false: min(max_dt, config.dt), whatever the state."""
    cfg = eh.SimConfig.standard(64, 0.05, (0.0, 0.0)).replace(adaptive_timestep=0)
    sim = _one_particle_sim(oracle, v=(100.0, 0.0), cfg=cfg)
    assert sim.choose_substep_dt(1.0) == pytest.approx(0.05, rel=1e-7)
    assert sim.choose_substep_dt(0.01) == pytest.approx(0.01, rel=1e-7)


def test_step_loop_schedule(oracle):
    """step.rs:23-60: a step of dt = 1 with a particle at |v| = 5 (no gravity, C = 0,
    so the free particle keeps v) takes five substeps of 0.18 and one of the rest."""
    sim = _one_particle_sim(oracle, v=(3.0, 4.0), x=(20.5, 20.5))
    sim.step()
    dts = sim.substep_dts()
    assert len(dts) == 6 and sim.stats().last_substeps == 6
    np.testing.assert_allclose(dts[:5], 0.18, rtol=1e-6)
    assert dts[5] == pytest.approx(1.0 - 5 * 0.18, rel=1e-5)
    assert sim.stats().last_sim_time_dropped <= 1.2e-7
    # max_substeps_per_step caps the loop and reports the time left
    cfg = eh.SimConfig.standard(64, 1.0, (0.0, 0.0)).replace(max_substeps_per_step=4)
    sim = _one_particle_sim(oracle, v=(3.0, 4.0), x=(20.5, 20.5), cfg=cfg)
    sim.step()
    assert sim.stats().last_substeps == 4
    assert sim.stats().last_sim_time_dropped == pytest.approx(1.0 - 4 * 0.18, rel=1e-5)


# -------------------------------------------------------------- projection

def _guarded_scene(oracle, precision=32):
    """physics_correctness.rs:2237-2253: 32^2, dt 0.02, no gravity, projection on,
    spacing 0.5, a 4x4-cell box centred at (16, 16), NeoHookean(100, 200)."""
    cfg = eh.SimConfig.standard(32, 0.02, (0.0, 0.0)).replace(project_invalid_state=1)
    spawn = eh.SpawnRegion(spacing=0.5, box_size=(4, 4), box_center=(16.0, 16.0), initial_velocity_scale=0.0)
    sim = oracle.OracleSim.new(cfg, spawn, precision)
    sim.with_default_material(oracle.material("neohookean", 100.0, 200.0))
    return sim, cfg


def _corrupt(p):
    """The fourteen corruptions of physics_correctness.rs:2257-2279."""
    nan = np.float32(np.nan)
    p["x"][0] = (nan, nan)
    p["x"][1] = (1.0e9, 1.0e9)
    p["v"][2] = (nan, nan)
    p["velocity_gradient"][3] = (nan, nan, 0.0, 0.0)
    p["deformation_gradient"][4] = (nan, nan, 0.0, 0.0)
    p["deformation_gradient"][5] = (0.0, 0.0, 0.0, 0.0)
    p["deformation_gradient"][6] = (1.0e6, 0.0, 0.0, 1.0e6)
    p["plastic_volume_ratio"][7] = -1.0
    p["hardening_scale"][8] = nan
    p["friction_hardening"][9] = nan
    p["log_volume_strain"][9] = nan
    p["mass"][10] = -1.0
    p["volume"][11] = 0.0
    p["initial_volume"][12] = nan
    p["density"][13] = -1.0
    return p


def _check_repairs(p, before, flags, cfg):
    R, bt = cfg.grid_res, cfg.boundary_thickness
    centre = ((bt - 1) + (R - bt)) * 0.5
    assert flags[0] and tuple(p["x"][0]) == (centre, centre)
    assert not flags[1] and tuple(p["x"][1]) == (R - bt, R - bt)  # clamped, not counted
    assert flags[2] and tuple(p["v"][2]) == (0.0, 0.0)
    assert flags[3] and tuple(p["velocity_gradient"][3]) == (0.0, 0.0, 0.0, 0.0)
    assert flags[4] and tuple(p["deformation_gradient"][4]) == (1.0, 0.0, 0.0, 1.0)
    assert flags[5] and tuple(p["deformation_gradient"][5]) == (1.0, 0.0, 0.0, 1.0)
    f6 = p["deformation_gradient"][6].astype(np.float64)
    assert flags[6]
    assert f6[1] == 0.0 and f6[2] == 0.0 and f6[0] == f6[3]
    assert f6[0] * f6[3] == pytest.approx(cfg.j_max, rel=1e-5)  # det scaled exactly to j_max
    assert flags[7] and p["plastic_volume_ratio"][7] == 1.0
    assert flags[8] and p["hardening_scale"][8] == 1.0
    assert flags[9] and p["friction_hardening"][9] == 0.0 and p["log_volume_strain"][9] == 0.0
    assert flags[10] and p["mass"][10] == cfg.particle_mass
    assert flags[11] and p["volume"][11] == max(before["initial_volume"][11], cfg.projection_min_volume)
    assert flags[12] and p["initial_volume"][12] == max(cfg.default_initial_volume, cfg.projection_min_volume)
    assert flags[13]
    assert p["density"][13] == pytest.approx(max(p["mass"][13] / p["volume"][13], cfg.projection_min_density),
                                             rel=1e-7)
    # every untouched particle comes back bit-identical and unflagged
    for i in range(14, len(p)):
        assert not flags[i]
        assert p[i].tobytes() == before[i].tobytes()


def test_projection_repairs_every_guarded_field(oracle):
    """project_invalid_state_recovers_every_guarded_field on the oracle, one field at a
    time, each repaired value stated exactly (step.rs:347-429)."""
    sim, cfg = _guarded_scene(oracle)
    p0 = sim.particles()
    assert len(p0) >= 14
    before = _corrupt(p0.copy())
    sim.set_particles(before)
    flags = [sim.project_particle(i) for i in range(len(before))]
    _check_repairs(sim.particles(), before, flags, cfg)


def test_projection_density_floor_and_rescale_are_pinned(oracle):
    """A valid density below projection_min_density is floored without being counted;
    F with J > j_max is scaled by sqrt(j_max/J).  Planted mistakes fail."""
    sim, cfg = _guarded_scene(oracle)
    p = sim.particles()
    p["density"][0] = 1e-9
    p["deformation_gradient"][1] = (1.0e6, 0.0, 0.0, 1.0e6)
    sim.set_particles(p)
    assert not sim.project_particle(0)
    assert sim.particles()["density"][0] == np.float32(cfg.projection_min_density)
    assert sim.project_particle(1)
    f = sim.particles()["deformation_gradient"][1].astype(np.float64)
    assert f[0] * f[3] == pytest.approx(cfg.j_max, rel=1e-5)
    for mut, idx in ((oracle.MUT_PROJ_NO_DENSITY_FLOOR, 0), (oracle.MUT_PROJ_JSCALE_LINEAR, 1)):
        sim.set_particles(p)
        with oracle.mutation(mut):
            sim.project_particle(idx)
        q = sim.particles()
        if idx == 0:
            assert q["density"][0] != np.float32(cfg.projection_min_density)
        else:
            g = q["deformation_gradient"][1].astype(np.float64)
            assert abs(g[0] * g[3] - cfg.j_max) > 1.0


def test_projected_scene_keeps_running(oracle):
    """physics_correctness.rs:2281-2372 through step(): after one step every particle is
    admissible, and twenty more steps do not diverge again."""
    sim, cfg = _guarded_scene(oracle)
    sim.set_particles(_corrupt(sim.particles()))
    sim.step()
    p = sim.particles()
    f = p["deformation_gradient"].astype(np.float64)
    det = f[:, 0] * f[:, 3] - f[:, 2] * f[:, 1]
    for name in ("x", "v", "velocity_gradient", "deformation_gradient", "plastic_volume_ratio", "hardening_scale",
                 "friction_hardening", "log_volume_strain", "mass", "initial_volume", "volume", "density"):
        assert np.all(np.isfinite(p[name])), name
    assert np.all(det > 0) and np.all(det <= cfg.j_max * 1.01)
    for name in ("plastic_volume_ratio", "hardening_scale", "mass", "initial_volume", "volume", "density"):
        assert np.all(p[name] > 0), name
    assert sim.stats().last_j_projection_count >= 13
    sim.step_n(20)
    p = sim.particles()
    f = p["deformation_gradient"].astype(np.float64)
    assert np.all(np.isfinite(p["x"])) and np.all(np.isfinite(p["v"]))
    assert np.all(f[:, 0] * f[:, 3] - f[:, 2] * f[:, 1] > 0)


# -------------------------------------------------------------------- NACC
MU, KAPPA, M, BETA, XI = 3000.0, 2000.0, r32(1.2), r32(0.3), 2.0  # the f32 material fields


def _nacc(oracle, xi=XI):
    return oracle.material("nacc", MU, KAPPA, M, BETA, xi)


def _p_of_j(j):  # nacc.rs: p = -kappa/2 (J^2 - 1)
    return -0.5 * KAPPA * (j * j - 1.0)


def _p0(alpha, xi=XI):
    return KAPPA * (r32(1e-5) + math.sinh(xi * max(-alpha, 0.0)))


def _sv(f):
    F = np.array([[f[0], f[2]], [f[1], f[3]]], np.float64)
    return np.linalg.svd(F, compute_uv=False)


def _trial(s1, s2, alpha):
    j = s1 * s2
    sq = np.array([s1 * s1, s2 * s2])
    s = MU / j * (sq - sq.mean())
    p = _p_of_j(j)
    p0 = _p0(alpha)
    return j, sq, s, p, p0


def _hardened_alpha(s1, s2, alpha):
    """alpha after the hardening update, from the geometry: the ray from the ellipse
    centre (p_c, 0) through the trial point (p_tr, q_tr) meets the yield ellipse
    M^2 (p + beta p0)(p - p0) + (1 + 2 beta) q^2 = 0 at t = M A / sqrt(M^2 dp^2 +
    (1 + 2 beta) q^2), A = (1 + beta) p0 / 2 (the ellipse's half-axis in p)."""
    j, _, s, p, p0 = _trial(s1, s2, alpha)
    pc = (1.0 - BETA) * p0 * 0.5
    q = math.sqrt(2.0) * float(np.linalg.norm(s))
    dp = p - pc
    a_half = (1.0 + BETA) * p0 * 0.5
    t = M * a_half / math.sqrt(M * M * dp * dp + (1.0 + 2.0 * BETA) * q * q)
    px = pc + t * dp
    jx = math.sqrt(abs(1.0 - 2.0 * px / KAPPA))
    return alpha + math.log(j / jx), (px, t * q)


# trial singular values and alpha (f32 values) spread over the hardening branch: p_tr from
# near the tensile tip to near the cap
HARDEN_CASES = [tuple(r32(v) for v in c) for c in
                [(1.03, 0.975, -0.1), (1.05, 0.96, -0.02), (0.99, 0.96, -0.05), (1.01, 0.97, -0.1),
                 (1.06, 0.95, -0.15), (1.05, 0.98, -0.05)]]


@pytest.mark.parametrize("s1,s2,alpha", HARDEN_CASES)
def test_nacc_hardening_alpha_is_the_ellipse_intersection(s1, s2, alpha, oracle):
    j, _, s, p, p0 = _trial(s1, s2, alpha)
    y = (1 + 2 * BETA) * 2 * float(s @ s) + M * M * (p + BETA * p0) * (p - p0)
    assert y > 1e-4 and -BETA * p0 + 1e-4 < p < p0 - 1e-4 and p0 > 1e-4  # the hardening branch runs
    want, (px, qx) = _hardened_alpha(s1, s2, alpha)
    # the intersection is on the ellipse
    assert abs(M * M * (px + BETA * p0) * (px - p0) + (1 + 2 * BETA) * qx * qx) < 1e-9 * p0 * p0
    m = _nacc(oracle)
    f = np.array([[s1, 0.0, 0.0, s2]])
    _, a64 = oracle.nacc_project_batch(m, f, np.array([alpha]), precision=64)
    assert a64[0] == pytest.approx(want, rel=1e-9, abs=1e-12)
    _, a32 = oracle.nacc_project(m, [s1, 0.0, 0.0, s2], alpha)
    assert a32 == pytest.approx(want, rel=2e-4, abs=2e-6)
    # a rotated trial state (F = R1 diag R2) hardens the same way
    r1, r2 = 0.4, -1.1
    Q1 = np.array([[math.cos(r1), -math.sin(r1)], [math.sin(r1), math.cos(r1)]])
    Q2 = np.array([[math.cos(r2), -math.sin(r2)], [math.sin(r2), math.cos(r2)]])
    Fr = Q1 @ np.diag([s1, s2]) @ Q2
    _, ar = oracle.nacc_project_batch(m, np.array([[Fr[0, 0], Fr[1, 0], Fr[0, 1], Fr[1, 1]]]), np.array([alpha]),
                                      precision=64)
    assert ar[0] == pytest.approx(want, rel=1e-5, abs=1e-6)  # F rounded to f32 after the rotation


@pytest.mark.parametrize("mut", ["MUT_NACC_OTHER_ROOT", "MUT_NACC_HARDEN_SIGN"])
def test_nacc_hardening_planted_mistakes_fail(mut, oracle):
    s1, s2, alpha = HARDEN_CASES[0]
    want, _ = _hardened_alpha(s1, s2, alpha)
    with oracle.mutation(getattr(oracle, mut)):
        _, a = oracle.nacc_project_batch(_nacc(oracle), np.array([[s1, 0.0, 0.0, s2]]), np.array([alpha]),
                                         precision=64)
    assert abs(a[0] - want) > 1e-3 * abs(want - alpha)


def test_nacc_no_hardening_when_xi_zero(oracle):
    """hardening_enabled = xi > 0 (nacc.rs:63-71): with xi = 0 alpha never moves."""
    m = _nacc(oracle, xi=0.0)
    for s1, s2, _ in HARDEN_CASES:
        _, a = oracle.nacc_project_batch(m, np.array([[s1, 0.0, 0.0, s2]]), np.array([-0.1]), precision=64)
        assert a[0] == r32(-0.1)


@pytest.mark.parametrize("s1,s2,alpha", HARDEN_CASES)
def test_nacc_case_c_returns_to_the_trial_pressure_ellipse(s1, s2, alpha, oracle):
    """Case C (nacc.rs:214-222): b' = sigma'^2 keeps the trial trace; its deviator is a
    positive multiple of the trial deviator, scaled so that (with the pre-hardening
    p0 and the trial pressure) y0 |s'|^2 + y1 = 0."""
    j, sq, s, p, p0 = _trial(s1, s2, alpha)
    y0 = (1 + 2 * BETA) * 2
    y1 = M * M * (p + BETA * p0) * (p - p0)
    fo, _ = oracle.nacc_project_batch(_nacc(oracle), np.array([[s1, 0.0, 0.0, s2]]), np.array([alpha]), precision=64)
    F = np.array([[fo[0, 0], fo[0, 2]], [fo[0, 1], fo[0, 3]]])
    assert abs(F[0, 1]) < 1e-12 and abs(F[1, 0]) < 1e-12  # diagonal in, diagonal out
    b = np.array([F[0, 0] ** 2, F[1, 1] ** 2])
    assert b.sum() == pytest.approx(sq.sum(), rel=1e-12)
    s_new = MU / j * (b - b.mean())
    assert np.dot(s_new, s) > 0 and abs(s_new[0] * s[1] - s_new[1] * s[0]) < 1e-9 * np.dot(s, s)
    assert y0 * float(s_new @ s_new) + y1 == pytest.approx(0.0, abs=1e-9 * abs(y1))


def test_nacc_case_a_cap(oracle):
    """Case A (p_tr > p0): isotropic F' with p(det F') = p0, alpha += ln(J_tr / det F')."""
    alpha = 0.0
    s = r32(0.98)
    j = s * s
    p0 = _p0(alpha)
    assert _p_of_j(j) > p0
    fo, a = oracle.nacc_project_batch(_nacc(oracle), np.array([[s, 0.0, 0.0, s]]), np.array([alpha]), precision=64)
    jn = fo[0, 0] * fo[0, 3] - fo[0, 2] * fo[0, 1]
    assert fo[0, 0] == pytest.approx(fo[0, 3], rel=1e-14)
    assert _p_of_j(jn) == pytest.approx(p0, rel=1e-9)
    assert a[0] == pytest.approx(alpha + math.log(j / jn), rel=1e-12)
    with oracle.mutation(oracle.MUT_NACC_CASE_A_NO_ALPHA):
        _, bad = oracle.nacc_project_batch(_nacc(oracle), np.array([[s, 0.0, 0.0, s]]), np.array([alpha]),
                                           precision=64)
    assert abs(bad[0] - (alpha + math.log(j / jn))) > 1e-3


def test_nacc_case_b_tip(oracle):
    """Case B (p_tr < -beta p0): isotropic F' with p(det F') = -beta p0, alpha += ln(J_tr / det F')."""
    alpha = r32(-0.05)
    s = r32(1.02)
    j = s * s
    p0 = _p0(alpha)
    assert _p_of_j(j) < -BETA * p0
    fo, a = oracle.nacc_project_batch(_nacc(oracle), np.array([[s, 0.0, 0.0, s]]), np.array([alpha]), precision=64)
    jn = fo[0, 0] * fo[0, 3] - fo[0, 2] * fo[0, 1]
    assert fo[0, 0] == pytest.approx(fo[0, 3], rel=1e-14)
    assert _p_of_j(jn) == pytest.approx(-BETA * p0, rel=1e-9)
    assert a[0] == pytest.approx(alpha + math.log(j / jn), rel=1e-12)


# ------------------------------------------------- threaded G2P (CPU baseline)
def test_parallel_g2p_phase1_is_bit_identical(oracle):
    """G2P phase 1 on several threads (the reference's rayon loop, g2p.rs:189-196) gives
    the serial result bit for bit: each particle reads the grid and writes only itself."""
    import scenes as S
    runs = []
    for threads in (1, 8):
        oracle.set_threads(threads)
        try:
            sim = S.multi_material("oracle32")
            sim.step_n(5)
            runs.append(sim.particles())
            sim.close()
        finally:
            oracle.set_threads(1)
    assert runs[0].tobytes() == runs[1].tobytes()
