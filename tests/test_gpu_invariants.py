"""GPU path: bit-exact cases, conservation, full-size invariants, error paths."""
import numpy as np
import pytest

import emerge_hip as eh
import scenes as S


def _isolated_particles(n_side=8, spacing=5.0, origin=12.0, seed=0):
    """Particles 5 cells apart: no two 3x3 stencils share a cell, so every grid
    sum has one term and the GPU result must equal the CPU one bit for bit."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij"), -1).reshape(-1, 2)
    n = len(g)
    p = eh.zeroed_particles(n)
    p["x"] = (origin + spacing * g + rng.uniform(0, 1, (n, 2))).astype(np.float32)
    p["v"] = rng.normal(0, 2.0, (n, 2)).astype(np.float32)
    p["velocity_gradient"] = rng.normal(0, 0.5, (n, 4)).astype(np.float32)
    F = np.eye(2)[None] + rng.normal(0, 0.05, (n, 2, 2))
    p["deformation_gradient"] = np.stack([F[:, 0, 0], F[:, 1, 0], F[:, 0, 1], F[:, 1, 1]], -1).astype(np.float32)
    p["mass"] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    p["initial_volume"] = rng.uniform(0.5, 1.5, n).astype(np.float32)
    p["volume"] = p["initial_volume"]
    p["density"] = p["mass"] / p["volume"]
    p["hardening_scale"] = rng.uniform(0.8, 1.5, n).astype(np.float32)
    return p


def _ulp_close(got, ref, ulps=16):
    """max |got - ref| <= `ulps` f32 ulps of the field's largest magnitude."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = float(np.max(np.abs(ref))) if ref.size else 0.0
    tol = ulps * float(np.spacing(np.float32(max(scale, 1e-30))))
    err = float(np.max(np.abs(got - ref))) if ref.size else 0.0
    return err <= tol, err, tol


def _rel_ulp_close(got, ref, ulps=64):
    """Per element: |got - ref| <= `ulps` ulps of |ref|.  For density and
    volume, m / (V0 det F): the determinant's cancellation roughly doubles the
    relative error F carries, so they get four times the 16-ulp field bound."""
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    tol = ulps * np.spacing(np.abs(ref)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - ref.astype(np.float64))
    return bool(np.all(err <= tol)), float(np.max(err / np.maximum(tol, 1e-45)) * ulps if err.size else 0.0), ulps


def _pair(material_fn, cfg, particles, n_steps=1):
    out = []
    for backend in ("gpu", "oracle32"):
        m = material_fn(S.Mats(backend))
        sim = S.new_sim(backend, cfg).with_default_material(m)
        sim.set_particles(particles)
        sim.step_n(n_steps)
        out.append(sim)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("adaptive", [0, 1])
def test_isolated_corotated_ulp_close(adaptive, oracle):
    """Corotated has no transcendental: with no shared grid cells every grid sum
    has one term, so the only difference from the oracle's f32 arithmetic is
    where dt enters the stress term (the device accumulates m v + m C d and
    -4 V0 tau d apart and adds dt x the latter in the grid update, because P2G
    runs before the CFL reduction picks dt).  That is an ulp per grid value;
    after 3 steps every field must stay within 16 ulps of the oracle, with the
    same number of substeps."""
    cfg = eh.SimConfig.standard(64, 0.01, (0.0, -9.81)).replace(adaptive_timestep=adaptive)
    pts = _isolated_particles()
    g, o = _pair(lambda m: m.corotated(400.0, 200.0), cfg, pts, n_steps=3)
    pg, po = g.particles(), o.particles()
    assert g.stats().total_substeps == o.stats().total_substeps
    for f in ("x", "v", "velocity_gradient", "deformation_gradient"):
        ok, err, tol = _ulp_close(pg[f], po[f])
        assert ok, (f, err, tol)
    for f in ("volume", "density"):
        ok, err, tol = _rel_ulp_close(pg[f], po[f])
        assert ok, (f, err, tol)
    assert abs(g.stats().last_step_dt - o.stats().last_step_dt) <= 1e-6 * o.stats().last_step_dt


@pytest.mark.gpu
def test_isolated_other_models_close(oracle):
    """Same isolated layout for every model: only device transcendentals
    (log/exp/sin/pow, <= a few ulp) separate GPU from oracle."""
    cfg = eh.SimConfig.standard(64, 0.01, (0.0, -9.81))
    pts = _isolated_particles(seed=3)
    pts["deformation_gradient"] = (1.0, 0.0, 0.0, 1.0)
    makers = {
        "neohookean": lambda m: m.neohookean(400.0, 200.0),
        "drucker_prager": lambda m: m.drucker_prager(3.5e5, 0.3, friction_deg=30.0),
        "snow": lambda m: m.snow(1.4e5, 0.2),
        "fluid": lambda m: m.fluid(1.0, 0.1, 100.0, 7.0),
        "nacc": lambda m: m.nacc(2e5, 0.3, 1.85, 0.5, 0.8, advance_f=True),
        "von_mises": lambda m: m.von_mises(5e4, 0.35, 250.0, 2e3),
        "viscoelastic": lambda m: m.viscoelastic(2e4, 0.45, 20.0),
        "rankine": lambda m: m.rankine(1e5, 0.25, 40.0, 2.0),
        "mu_i": lambda m: m.mu_i(1e5, 0.3),
        "granular_fluid": lambda m: m.saturated_loam(1e5, 0.2),
        "bingham": lambda m: m.bingham(1.0, 0.1, 100.0, 7.0, 2.0),
    }
    # start the solids strained so that their return maps run: Von Mises and
    # mu(I) beyond yield (mu(I) compressed), Rankine in tension, the
    # granular fluid past its compression limit, Kelvin-Voigt sheared
    start_f = {"von_mises": (1.02, 0.0, 0.0, 0.98), "viscoelastic": (1.05, 0.01, 0.02, 0.97),
               "rankine": (1.01, 0.002, 0.0, 1.0), "mu_i": (0.96, 0.01, 0.0, 1.01),
               "granular_fluid": (0.55, 0.02, 0.0, 1.0)}
    for name, fn in makers.items():
        pts["deformation_gradient"] = start_f.get(name, (1.0, 0.0, 0.0, 1.0))
        g, o = _pair(fn, cfg, pts, n_steps=2)
        o64 = S.new_sim("oracle64", cfg).with_default_material(fn(S.Mats("oracle64")))
        o64.set_particles(pts)
        o64.step_n(2)
        ref = S.f64_state(o64.particles_f64())
        e_g = S.rel_errors(S.state_matrix(g.particles()), ref)
        e_o = S.rel_errors(S.state_matrix(o.particles()), ref)
        assert g.stats().total_substeps == o.stats().total_substeps, name
        # the GPU is as close to the FP64 oracle as the f32 oracle is (plus a few ulp)
        for f in ("x", "v", "C", "F", "q", "lvs", "Jp", "density"):
            assert e_g[f] <= 4.0 * e_o[f] + 1e-6, (name, f, e_g[f], e_o[f])


@pytest.mark.gpu
def test_p2g_grid_matches_oracle_and_conserves(oracle):
    """One fixed-dt substep on identical input: the GPU grid (mass, momentum
    before normalisation) equals the oracle's cell by cell to f32 summation
    roundoff, Σ mass = Σ m_p and Σ momentum = Σ m_p v_p."""
    cfg = eh.SimConfig.unsafe_defaults().replace(grid_res=64, dt=2e-3, min_dt=1e-3, gravity=(0.0, -9.81))
    src = S.quickstart_impact("oracle32")
    src.step_n(2)
    pts = src.particles()
    pts["material_id"] = 0
    g, o = _pair(lambda m: m.neohookean(400.0, 200.0), cfg, pts, n_steps=1)
    gm, gmom = g.grid()
    o64 = oracle.OracleSim.empty(cfg, 64).with_default_material(oracle.material("neohookean", 400.0, 200.0))
    o64.set_particles(pts)
    om, omom = o64.p2g_grid(2e-3)
    assert np.max(np.abs(gm - om)) <= 1e-6 * np.max(om)
    assert np.max(np.abs(gmom - omom)) <= 1e-5 * np.max(np.abs(omom))
    msum = pts["mass"].astype(np.float64).sum()
    assert abs(gm.astype(np.float64).sum() - msum) <= 1e-6 * msum
    pm = (pts["mass"][:, None].astype(np.float64) * pts["v"]).sum(0)
    scale = np.abs(pts["mass"][:, None] * pts["v"]).sum()
    assert np.max(np.abs(gmom.astype(np.float64).sum(0) - pm)) <= 1e-5 * scale


@pytest.mark.gpu
def test_substep_schedule_matches_oracle(oracle):
    """The device-side adaptive loop takes the same substeps as step.rs:34-49."""
    g, o = S.quickstart("gpu"), S.quickstart("oracle32")
    for _ in range(5):
        g.step()
        o.step()
        sg, so = g.stats(), o.stats()
        assert sg.last_substeps == so.last_substeps
        assert sg.last_step_dt == so.last_step_dt
        assert sg.last_sim_time_dropped == so.last_sim_time_dropped


@pytest.mark.gpu
def test_full_size_invariants():
    """The bench workload (1,048,576 particles, 512^2): mass exactly conserved,
    every particle returned once, state finite and inside the domain, J >= 0.6
    floor of Drucker-Prager, grid mass = particle mass."""
    import bench
    sim = bench.build_workload("gpu")
    p0 = sim.particles()
    sim.step_n(3)
    p = sim.particles()
    assert len(p) == len(p0) == 1 << 20
    assert np.array_equal(p["mass"], p0["mass"])
    assert np.array_equal(p["user_tag"], p0["user_tag"])
    assert np.all(np.isfinite(p["x"])) and np.all(np.isfinite(p["deformation_gradient"]))
    assert p["x"].min() >= 1.0 and p["x"].max() <= 510.0
    F = p["deformation_gradient"].astype(np.float64)
    J = F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]
    assert J.min() > 0.6 - 1e-4
    gm, _ = sim.grid()
    assert abs(gm.astype(np.float64).sum() - p["mass"].astype(np.float64).sum()) < 1e-5 * len(p)
    # the particles moved (gravity acted) and stayed ordered by id on download
    assert np.any(p["x"] != p0["x"])


@pytest.mark.gpu
def test_multi_material_full_size_invariants():
    """The multi-material scene at full size (1024^2, ~4M particles of four models, density
    pass for the water): every particle returned once with its mass, id and
    material, state finite and inside the domain, grid mass = particle mass,
    every model's particles moved."""
    sim = S.multi_material_full("gpu")
    p0 = sim.particles()
    n = len(p0)
    assert 3_500_000 < n < 4_600_000, n
    sim.step_n(2)
    assert sim.stats().last_substeps >= 1
    p = sim.particles()
    assert len(p) == n
    for f in ("mass", "user_tag", "material_id"):
        assert np.array_equal(p[f], p0[f]), f
    assert np.all(np.isfinite(p["x"])) and np.all(np.isfinite(p["v"])) and np.all(np.isfinite(p["deformation_gradient"]))
    assert p["x"].min() >= 1.0 and p["x"].max() <= 1022.0
    gm, _ = sim.grid()
    total = p["mass"].astype(np.float64).sum()
    assert abs(gm.astype(np.float64).sum() - total) <= 1e-5 * total
    for mid in range(4):
        sel = p["material_id"] == mid
        assert sel.sum() > 0 and np.any(p["x"][sel] != p0["x"][sel]), mid


@pytest.mark.gpu
def test_error_paths():
    cfg = eh.SimConfig.standard(64, 0.05, (0.0, -1.0))
    sim = eh.Simulation.empty(cfg)
    bad = eh.NeoHookeanMaterial.new(1.0, 1.0)
    bad.params.model = 12  # no such ConstitutiveModel (materials/mod.rs:53-66)
    with pytest.raises(eh.EmergeError) as ei:
        sim.set_material(0, bad)
    assert ei.value.code == -2
    active = eh.ViscoelasticMaterial.new(1.0, 1.0, 1.0)
    active.params.active_stress_coeff = 0.5  # the particle activation field is not carried
    with pytest.raises(eh.EmergeError) as ei:
        sim.set_material(0, active)
    assert ei.value.code == -2
    with pytest.raises(eh.EmergeError) as ei:  # friction.rs:29-32
        sim.with_boundary(eh.FrictionBoundary.new(2, 1.5))
    assert ei.value.code == -1
    with pytest.raises(eh.EmergeError) as ei:
        sim.set_material(3, eh.NeoHookeanMaterial.new(1.0, 1.0))  # not contiguous (registry.rs:39-46)
    assert ei.value.code == -1
    with pytest.raises(eh.EmergeError) as ei:  # outside the domain (config.rs:485-505)
        sim.add_body(eh.SpawnRegion(box_size=(6, 6), box_center=(59.5, 19.8), spacing=0.5))
    assert ei.value.code == -1
    p = eh.zeroed_particles(4)
    p["x"] = 20.0
    p["mass"] = 1.0
    p["contact_group"] = 1
    sim.set_particles(p)
    with pytest.raises(eh.EmergeError) as ei:
        sim.step()
    assert ei.value.code == -2
    sim.close()


@pytest.mark.gpu
def test_zero_particles_and_pinned(oracle):
    """An empty simulation still advances time; pinned particles keep x and get v = C = 0."""
    cfg = eh.SimConfig.standard(32, 0.05, (0.0, -1.0))
    sim = eh.Simulation.empty(cfg)
    sim.step_n(2)
    assert sim.stats().frame_index == 2
    pts = _isolated_particles(n_side=3, spacing=6.0, origin=6.0)
    pts["pinned"][::2] = 1
    g, o = _pair(lambda m: m.corotated(100.0, 50.0), cfg, pts, n_steps=2)
    pg, po = g.particles(), o.particles()
    assert np.array_equal(pg["x"][::2], pts["x"][::2])
    assert np.all(pg["v"][::2] == 0)
    for f in ("x", "deformation_gradient"):
        ok, err, tol = _ulp_close(pg[f], po[f])
        assert ok, (f, err, tol)


@pytest.mark.gpu
def test_projection_repairs_invalid_state(oracle):
    """project_particle_state_to_admissible (step.rs:347-429) on injected bad
    fields, as physics_correctness.rs:2237-2372 does: NaN velocity, inverted F,
    zero mass, out-of-domain position -- same repair and count as the oracle."""
    cfg = eh.SimConfig.standard(64, 0.01, (0.0, -9.81))
    pts = _isolated_particles(n_side=4, seed=5)
    pts["v"][0] = (np.nan, 1.0)
    pts["deformation_gradient"][1] = (-1.0, 0.0, 0.0, 1.0)
    pts["mass"][2] = 0.0
    pts["x"][3] = (200.0, -5.0)
    pts["density"][4] = np.inf
    pts["deformation_gradient"][5] = (10.0, 0.0, 0.0, 10.0)  # J = 100 > j_max = 50
    g, o = _pair(lambda m: m.corotated(400.0, 200.0), cfg, pts, n_steps=1)
    assert g.stats().last_j_projection_count == o.stats().last_j_projection_count >= 5
    pg, po = g.particles(), o.particles()
    assert np.array_equal(pg["mass"], po["mass"])
    for f in ("x", "v", "deformation_gradient"):
        ok, err, tol = _ulp_close(pg[f], po[f])
        assert ok, (f, err, tol)
    ok, err, tol = _rel_ulp_close(pg["density"], po["density"])
    assert ok, ("density", err, tol)


@pytest.mark.gpu
def test_profiling_counts_substeps():
    """Step-level timing counts every step and every substep taken."""
    sim = S.quickstart("gpu")
    sim.step_n(2)
    sim.set_profiling(True)
    s0 = sim.stats().total_substeps
    sim.step_n(3)
    t = sim.kernel_times()
    sim.set_profiling(False)
    assert t["step_setup"][0] == 3
    assert t["k_density_p2g+k_substep"][0] == sim.stats().total_substeps - s0
    assert all(ms > 0.0 for _, ms in t.values())


@pytest.mark.gpu
def test_diagnostics_snapshot_matches_state():
    """diagnostics_snapshot (snapshot.rs:227-372) against the same sums taken
    in float64 over the downloaded particles, and against the grid."""
    sim = S.multi_material("gpu")
    sim.step_n(3)
    d = sim.diagnostics_snapshot()
    p = sim.particles()
    n = len(p)
    m = p["mass"].astype(np.float64)
    v = p["v"].astype(np.float64)
    F = p["deformation_gradient"].astype(np.float64)
    J = F[:, 0] * F[:, 3] - F[:, 2] * F[:, 1]
    assert d["particle_count"] == n and d["valid_particle_count"] == n and d["non_finite_particle_values"] == 0
    assert abs(d["total_particle_mass"] - m.sum()) <= 1e-6 * m.sum()
    mom = (m[:, None] * v).sum(0)
    scale = (m[:, None] * np.abs(v)).sum()
    assert np.max(np.abs(np.array(d["total_particle_momentum"]) - mom)) <= 1e-6 * scale
    ke = (0.5 * m * (v ** 2).sum(1)).sum()
    assert abs(d["total_kinetic_energy"] - ke) <= 1e-5 * ke + 1e-12
    assert d["max_particle_speed"] == pytest.approx(np.sqrt((v ** 2).sum(1)).max(), rel=1e-6)
    assert d["min_deformation_j"] == pytest.approx(J.min(), rel=1e-5)
    assert d["max_deformation_j"] == pytest.approx(J.max(), rel=1e-5)
    gm, _ = sim.grid()
    assert d["total_grid_mass"] == pytest.approx(gm.astype(np.float64).sum(), rel=1e-6)
    assert d["active_grid_cells"] == int((gm > 0).sum())
    assert d["relative_mass_error"] < 1e-5
    # materials per cell of floor(x)
    cells = np.floor(p["x"]).astype(np.int64)
    r = sim.config.grid_res
    key = cells[:, 0] * r + cells[:, 1]
    lo, hi = {}, {}
    for k, mid in zip(key, p["material_id"]):
        lo[k] = min(lo.get(k, mid), mid)
        hi[k] = max(hi.get(k, mid), mid)
    mixed = {k for k in lo if lo[k] != hi[k]}
    assert d["mixed_material_cell_ratio"] == pytest.approx(len(mixed) / len(lo), rel=1e-6)
    assert d["mixed_material_particle_ratio"] == pytest.approx(np.isin(key, list(mixed)).sum() / n, rel=1e-6)
    st = sim.stats()
    assert d["substeps_last_step"] == st.last_substeps and d["frame_index"] == st.frame_index == 3
