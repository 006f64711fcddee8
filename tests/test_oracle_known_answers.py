"""Pins the CPU oracle before it is trusted as the checker.

The reference ships no golden vectors (no fixtures exist anywhere in the
crate); what it does ship are known-answer and invariant unit tests next
to the math.  Each test below restates one of them against the oracle (file:
line of the reference test in the docstring), plus checks the mathematics
fixes on its own (closed forms, textbook SVD, conservation), chosen so that a
dropped term, a sign error or a transposed operand fails at least one.
"""
import math

import numpy as np
import pytest

import emerge_hip as eh
import scenes as S

# ------------------------------------------------------------- B-spline

@pytest.mark.parametrize("d", [0.0, 0.1, 0.35, -0.2, 0.49, -0.5])
def test_weights_sum_to_one_and_reproduce_linear(d, oracle):
    """grid/kernel.rs:59-66 (sum to 1) + first moment Σ w_i·(cell_i - x + 1/2) = 0,
    the property the APIC terms rely on (a wrong sign of d fails it)."""
    x = np.float32(10.0) + np.float32(0.5) + np.float32(d)
    (bx, by), wx, wy = oracle.quadratic_weights(float(x), 3.75)
    assert bx == math.floor(x)
    assert abs(float(wx.sum()) - 1.0) < 1e-6
    dist = np.array([bx - 1, bx, bx + 1], np.float64) - float(x) + 0.5
    assert abs(float(np.dot(wx.astype(np.float64), dist))) < 1e-6
    # closed form of axis_weights (kernel.rs:20-25)
    dd = float(x) - bx - 0.5
    np.testing.assert_allclose(wx, [0.5 * (0.5 - dd) ** 2, 0.75 - dd ** 2, 0.5 * (0.5 + dd) ** 2], atol=1e-7)


# ------------------------------------------------------------------ svd2
SVD_CASES = {
    "identity": [1, 0, 0, 1],
    "diagonal": [3.0, 0.0, 0.0, 1.5],
    "shear": [1.0, 0.5, 0.5, 1.0],
    "generic": [1.2, 0.3, -0.1, 0.9],
    "rank_deficient": [1.0, 2.0, 2.0, 4.0],
    "negative_det": [1.0, 0.0, 0.0, -1.0],
    "near_identity": [0.99, 0.01, -0.01, 1.01],
    "pure_rotation": [math.cos(0.7), math.sin(0.7), -math.sin(0.7), math.cos(0.7)],
}


def _m(v):  # column-major [a, b, c, d] -> row-major numpy matrix
    return np.array([[v[0], v[2]], [v[1], v[3]]], np.float64)


@pytest.mark.parametrize("name", list(SVD_CASES))
def test_svd2_reference_cases(name, oracle):
    """materials/svd.rs:91-183: reconstruction, det U = det V = +1, σ1 ≥ |σ2|, U orthogonal."""
    f = SVD_CASES[name]
    u, s, vt = oracle.svd2(f)
    U, Vt, F = _m(u), _m(vt), _m(f)
    R = U @ np.diag(s.astype(np.float64)) @ Vt
    assert np.max(np.abs(R - F)) < 1e-5
    assert abs(np.linalg.det(U) - 1) < 1e-5 and abs(np.linalg.det(Vt) - 1) < 1e-5
    assert s[0] >= 0 and s[0] >= abs(s[1]) - 1e-6
    assert np.max(np.abs(U.T @ U - np.eye(2))) < 1e-5
    if name == "negative_det":  # svd.rs:159-175: inversion is carried by a signed σ2
        assert s[1] <= 0.0


def test_svd2_matches_textbook_svd(oracle):
    """|σ| against LAPACK's SVD on random matrices (a textbook routine), and the
    sign of σ1σ2 equals sign(det F)."""
    rng = np.random.default_rng(0)
    for _ in range(500):
        f = rng.normal(0, 1, 4).astype(np.float32)
        u, s, vt = oracle.svd2(f.tolist(), precision=64)
        ref = np.linalg.svd(_m(f), compute_uv=False)
        np.testing.assert_allclose(np.sort(np.abs(s))[::-1], ref, rtol=1e-9, atol=1e-9)
        assert np.sign(s[0] * s[1]) == np.sign(np.linalg.det(_m(f))) or abs(np.linalg.det(_m(f))) < 1e-12


def test_polar_symmetric_is_identity(oracle):
    """corotated.rs:238-248: symmetric F gives R = I exactly."""
    r = oracle.polar([1.3, -0.15, -0.15, 1.6])
    assert np.array_equal(r, np.array([1, 0, 0, 1], np.float32))
    # and R is a rotation for a general F, with R^T F symmetric (the defining property)
    f = [1.2, 0.4, -0.3, 0.8]
    R, F = _m(oracle.polar(f)), _m(f)
    assert abs(np.linalg.det(R) - 1) < 1e-6
    S_ = R.T @ F
    assert abs(S_[0, 1] - S_[1, 0]) < 1e-6


# -------------------------------------------------------- elastic stress
def _particle(f, h=1.0, jp=1.0, q=0.0, c=(0, 0, 0, 0), density=1.0, volume=1.0):
    p = eh.zeroed_particles(1)[0]
    p["deformation_gradient"] = f
    p["velocity_gradient"] = c
    p["mass"] = 1.0
    p["initial_volume"] = 1.0
    p["volume"] = volume
    p["density"] = density
    p["hardening_scale"] = h
    p["plastic_volume_ratio"] = jp
    p["friction_hardening"] = q
    return p


def _hooke(lam, mu, eps):
    tr = eps[0, 0] + eps[1, 1]
    return lam * tr * np.eye(2) + 2 * mu * eps


@pytest.mark.parametrize("kind,lam,mu,e", [
    ("neohookean", 1000.0, 800.0, [[1, 0], [0, 0]]),    # elastic_tests.rs:66-94
    ("neohookean", 500.0, 1200.0, [[0, 1], [1, 0]]),    # elastic_tests.rs:96-122
    ("corotated", 1000.0, 800.0, [[1, 0], [0, 0]]),     # corotated.rs:182-205
    ("corotated", 500.0, 1200.0, [[0, 1], [1, 0]]),     # corotated.rs:207-230
    ("snow", 1000.0, 800.0, [[1, 0], [0, -0.4]]),       # snow.rs:235-261
])
def test_small_strain_matches_hookes_law(kind, lam, mu, e, oracle):
    delta = 1e-4
    E = np.array(e, np.float64)
    F = np.eye(2) + delta * E
    f = [F[0, 0], F[1, 0], F[0, 1], F[1, 1]]
    if kind == "snow":
        m = oracle.material("snow", lam, mu, 10.0, 0.025, 0.0075, 0.6, 20.0)
    else:
        m = oracle.material(kind, lam, mu)
    tau = _m(oracle.kirchhoff(m, _particle(f)))
    pred = _hooke(lam, mu, delta * E)
    assert np.linalg.norm(tau - pred) / np.linalg.norm(pred) < 1e-3


def test_neohookean_hooke_error_is_first_order(oracle):
    """elastic_tests.rs:124-160: relative error halves when the strain halves."""
    m = oracle.material("neohookean", 1000.0, 800.0)
    E = np.diag([1.0, -0.3])

    def err(delta):
        F = np.eye(2) + delta * E
        tau = _m(oracle.kirchhoff(m, _particle([F[0, 0], F[1, 0], F[0, 1], F[1, 1]])))
        pred = _hooke(1000.0, 800.0, delta * E)
        return np.linalg.norm(tau - pred) / np.linalg.norm(pred)

    big, small = err(1e-2), err(5e-3)
    assert 0.3 * big < small < 0.7 * big


def test_neohookean_closed_form(oracle):
    """elastic.rs:235-254: τ = (µ/J)·dev(FFᵀ) + (λ+µ)·ln J·I, at a large strain."""
    lam, mu = 400.0, 200.0
    f = [1.3, 0.2, -0.1, 0.8]
    F = _m(f)
    J = np.linalg.det(F)
    B = F @ F.T
    ref = (mu / J) * (B - 0.5 * np.trace(B) * np.eye(2)) + (lam + mu) * np.log(J) * np.eye(2)
    tau = _m(oracle.kirchhoff(oracle.material("neohookean", lam, mu), _particle(f)))
    np.testing.assert_allclose(tau, ref, rtol=2e-6, atol=1e-3)


def test_corotated_closed_form(oracle):
    """corotated.rs:62-78: τ = 2µh(F−R)Fᵀ + λh(J−1)J·I with h = hardening_scale."""
    lam, mu, h = 400.0, 200.0, 1.7
    f = [1.1, 0.3, -0.2, 0.9]
    F = _m(f)
    Ufull, _, Vt = np.linalg.svd(F)
    R = Ufull @ Vt
    J = np.linalg.det(F)
    ref = 2 * mu * h * (F - R) @ F.T + lam * h * (J - 1) * J * np.eye(2)
    tau = _m(oracle.kirchhoff(oracle.material("corotated", lam, mu), _particle(f, h=h)))
    np.testing.assert_allclose(tau, ref, rtol=1e-5, atol=1e-3)


def test_fluid_tait_pressure_and_viscosity(oracle):
    """fluid.rs:157-176: p = max(k((ρ/ρ0)^γ − 1), floor); τ = −p·I + µ·dev(C + Cᵀ)."""
    m = oracle.material("fluid", 4.0, 0.5, 10.0, 7.0)
    c = [0.2, -0.1, 0.3, 0.05]
    tau = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=c, density=4.4)))
    p = 10.0 * ((4.4 / 4.0) ** 7 - 1.0)
    C = _m(c)
    sym = C + C.T
    ref = -p * np.eye(2) + 0.5 * (sym - 0.5 * np.trace(sym) * np.eye(2))
    np.testing.assert_allclose(tau, ref, rtol=1e-5)
    # density clamp at 2ρ0 and the tension floor -0.1 (fluid.rs:157-162)
    t2 = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], density=100.0)))
    assert abs(-t2[0, 0] - 10.0 * (2.0 ** 7 - 1.0)) < 1e-2
    t3 = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], density=0.5)))
    assert abs(t3[0, 0] - 0.1) < 1e-7


# ---------------------------------------------------------- Drucker-Prager
def test_dp_alpha_known_answer(oracle):
    """sand.rs:375-381,420: α(q_init) = 0.386019 for the default φ0 = 35°
    (q_init = h3/h1), i.e. φ_eff = asin(√2 α) = 33.087°."""
    lam, mu = oracle.lame_from_young(1e5, 0.2)
    m = oracle.material("drucker_prager", 2000.0, 3000.0)
    q_init = m.params.dp_h3 / m.params.dp_h1
    a = oracle.dp_alpha(m, q_init)
    assert abs(a - 0.386019) < 2e-6
    assert abs(math.degrees(math.asin(math.sqrt(2) * a)) - 33.087) < 2e-3


def _dp_marginal(ratio, trace, phi_deg):  # sand.rs:391-398
    phi = math.radians(phi_deg)
    dev_norm = -trace * ratio * math.sin(phi) / math.sqrt(2)
    diff = dev_norm * math.sqrt(2)
    e1, e2 = (trace + diff) * 0.5, (trace - diff) * 0.5
    return math.exp(e1), math.exp(e2)


def test_dp_marginal_yield(oracle):
    """sand.rs:415-449: a 25° state is elastic, a 40° state yields and q grows."""
    m = oracle.material("drucker_prager", 2000.0, 3000.0)
    q_init = m.params.dp_h3 / m.params.dp_h1
    ratio = (2000.0 + 3000.0) / 3000.0
    for phi, plastic in ((25.0, False), (40.0, True)):
        s1, s2 = _dp_marginal(ratio, -0.01, phi)
        p = _particle([s1, 0, 0, s2], q=q_init)
        out = oracle.update_particle(m, p, 1.0)
        sig = np.array([out["deformation_gradient"][0], out["deformation_gradient"][3]])
        moved = np.linalg.norm(sig - [s1, s2])
        if plastic:
            assert moved > 1e-6 and out["friction_hardening"] > q_init
        else:
            assert moved < 1e-6 and abs(out["friction_hardening"] - q_init) < 1e-6


def test_dp_tension_projects_to_identity_and_floor(oracle):
    """sand.rs:171-173 (trace > 0 → σ = 1) and the J ≥ 0.6 volumetric floor (sand.rs:306-310)."""
    m = oracle.material("drucker_prager", 2000.0, 3000.0)
    out = oracle.update_particle(m, _particle([1.05, 0, 0, 1.02]), 1.0)
    np.testing.assert_allclose(out["deformation_gradient"], [1, 0, 0, 1], atol=1e-6)
    # a purely volumetric state also projects to identity (dev_norm == 0, sand.rs:171)
    out = oracle.update_particle(m, _particle([0.7, 0, 0, 0.7]), 1.0)
    np.testing.assert_allclose(out["deformation_gradient"], [1, 0, 0, 1], atol=1e-6)
    # compressed with shear, inside the cone: elastic, then rescaled up to J = 0.6
    out = oracle.update_particle(m, _particle([0.75, 0, 0, 0.7]), 1.0)
    F = out["deformation_gradient"]
    assert abs(F[0] * F[3] - F[1] * F[2] - 0.6) < 1e-5
    assert abs(F[0] / F[3] - 0.75 / 0.7) < 1e-5


# ------------------------------------------------------------------- snow
def test_snow_clamp_exact(oracle):
    """snow.rs:269-316: inside the limits nothing changes; beyond, σ clamps to
    1 − θc exactly and Jp = σx / (1 − θc); h = clamp(exp(ξ(1 − Jp)), 0.1, 7)."""
    m = oracle.material("snow", 1000.0, 800.0, 10.0, 0.025, 0.0075, 0.6, 20.0)
    sx = 1.0 - 0.99 * 0.025
    out = oracle.update_particle(m, _particle([sx, 0, 0, 1.0]), 1.0)
    assert abs(out["deformation_gradient"][0] - sx) < 1e-5
    assert out["plastic_volume_ratio"] == 1.0
    sx = np.float32(1.0 - 1.5 * 0.025)
    out = oracle.update_particle(m, _particle([float(sx), 0, 0, 1.0]), 1.0)
    assert abs(out["deformation_gradient"][0] - (1 - 0.025)) < 1e-5
    jp = sx / np.float32(1 - 0.025)
    assert abs(out["plastic_volume_ratio"] - jp) < 1e-5
    assert abs(out["hardening_scale"] - min(max(math.exp(10 * (1 - jp)), 0.1), 7.0)) < 1e-4


# -------------------------------------------------------------- Von Mises
def _vm_state(mat_mu, yield_stress, factor, trace):
    """von_mises.rs:207-209,213-220: ε = (tr/2 + d, tr/2 − d) with |dev ε| = factor · σY / (2µ)."""
    d = factor * yield_stress / (2.0 * mat_mu) / math.sqrt(2.0)
    return math.exp(trace * 0.5 + d), math.exp(trace * 0.5 - d)


def _vm_step(m, oracle, s1, s2, kappa=0.0):  # von_mises.rs:187-200
    out = oracle.update_particle(m, _particle([s1, 0, 0, s2], q=kappa), 1.0)
    F = out["deformation_gradient"]
    return np.array([F[0], F[3]], np.float64), float(out["friction_hardening"]), F


def test_von_mises_inside_yield_is_elastic(oracle):
    """von_mises.rs:211-233: 99 % of the yield threshold leaves σ and κ unchanged."""
    m = oracle.material("von_mises", 2000.0, 3000.0, 100.0, 0.0)
    s1, s2 = _vm_state(3000.0, 100.0, 0.99, 0.0)
    sig, kappa, _ = _vm_step(m, oracle, s1, s2)
    assert np.linalg.norm(sig - [s1, s2]) < 1e-4 and kappa == 0.0


def test_von_mises_projects_onto_yield_surface(oracle):
    """von_mises.rs:235-272: 150 % of the threshold lands on |dev ε| = σY/(2µ),
    keeps tr ε = 0.4, and κ grows by γ = (2µ|dev| − σY)/(2µ + H)."""
    m = oracle.material("von_mises", 2000.0, 3000.0, 100.0, 0.0)
    s1, s2 = _vm_state(3000.0, 100.0, 1.5, 0.4)
    sig, kappa, F = _vm_step(m, oracle, s1, s2)
    eps = np.log(np.abs(sig))
    tr = eps.sum()
    dev = np.linalg.norm(eps - tr / 2)
    assert abs(dev - 100.0 / 6000.0) < 1e-4 and abs(tr - 0.4) < 1e-4
    # the closed form of the increment: 2µ|dev_trial| = 1.5 σY
    assert abs(kappa - 0.5 * 100.0 / 6000.0) < 1e-5
    assert F[1] == 0.0 and F[2] == 0.0  # a diagonal trial state stays diagonal


def test_von_mises_hardening_shrinks_increment(oracle):
    """von_mises.rs:274-297: with H > 0 the same trial state yields a smaller γ once κ = 1,
    and γ = (2µ|dev| − σY − Hκ)/(2µ + H) where it yields."""
    H = 500.0
    m = oracle.material("von_mises", 2000.0, 3000.0, 100.0, H)
    s1, s2 = _vm_state(3000.0, 100.0, 1.5, 0.0)
    _, k0, _ = _vm_step(m, oracle, s1, s2, 0.0)
    _, k1, _ = _vm_step(m, oracle, s1, s2, 1.0)
    assert k1 - 1.0 < k0
    el = 1.5 * 100.0
    assert abs(k0 - (el - 100.0) / (6000.0 + H)) < 1e-6
    assert k1 == 1.0  # σY + Hκ = 600 > 150: elastic
    _, kh, _ = _vm_step(m, oracle, s1, s2, 0.05)
    assert abs((kh - 0.05) - (el - 100.0 - H * 0.05) / (6000.0 + H)) < 1e-6


def test_von_mises_stress_is_corotated(oracle):
    """von_mises.rs:94-102: the stress is DP's corotated τ = 2µ(F−R)Fᵀ + λ(J−1)J·I."""
    f = [1.1, 0.3, -0.2, 0.9]
    a = oracle.kirchhoff(oracle.material("von_mises", 400.0, 200.0, 50.0, 0.0), _particle(f))
    b = oracle.kirchhoff(oracle.material("drucker_prager", 400.0, 200.0), _particle(f))
    assert np.array_equal(a, b)


# ----------------------------------------------------------- Viscoelastic
def test_viscoelastic_small_strain_is_hooke(oracle):
    """viscoelastic.rs:219-249: F = I + 1e-4·diag(1, −0.4), C = 0 gives 2µε + λ tr ε I to 1e-3."""
    m = oracle.material("viscoelastic", 1000.0, 800.0, 50.0)
    E = np.diag([1.0, -0.4])
    F = np.eye(2) + 1e-4 * E
    tau = _m(oracle.kirchhoff(m, _particle([F[0, 0], F[1, 0], F[0, 1], F[1, 1]])))
    pred = _hooke(1000.0, 800.0, 1e-4 * E)
    assert np.linalg.norm(tau - pred) / np.linalg.norm(pred) < 1e-3


def test_viscoelastic_viscous_term_exact(oracle):
    """viscoelastic.rs:251-278: F = I isolates τ = η·dev(sym C); pure shear C gives η·C."""
    m = oracle.material("viscoelastic", 1000.0, 800.0, 50.0)
    tau = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=[0.0, 3.0, 3.0, 0.0])))
    assert np.linalg.norm(tau - 50.0 * _m([0.0, 3.0, 3.0, 0.0])) < 1e-4
    # a general C: only the symmetric deviatoric part survives (the spin and the trace drop out)
    c = [0.4, -1.0, 2.0, 0.1]
    tau = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=c)))
    C = _m(c)
    D = 0.5 * (C + C.T)
    np.testing.assert_allclose(tau, 50.0 * (D - 0.5 * np.trace(D) * np.eye(2)), atol=1e-4)


def test_viscoelastic_terms_are_additive(oracle):
    """viscoelastic.rs:280-310: τ(F, C) = τ(F, 0) + τ(I, C) to 1e-3 at a finite strain."""
    m = oracle.material("viscoelastic", 1000.0, 800.0, 50.0)
    F = np.eye(2) + 0.02 * np.diag([1.0, -0.3])
    f = [F[0, 0], F[1, 0], F[0, 1], F[1, 1]]
    c = [0.0, 4.0, 4.0, 0.0]
    both = _m(oracle.kirchhoff(m, _particle(f, c=c)))
    el = _m(oracle.kirchhoff(m, _particle(f)))
    vi = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=c)))
    assert np.linalg.norm(both - (el + vi)) < 1e-3


def test_viscoelastic_closed_form_and_j_clamp(oracle):
    """viscoelastic.rs:103-123: τ = µ(FFᵀ − I) + λ ln(clamp(J, j_min, 1/j_min)) I, the J
    clamp at j_min = 0.01 (viscoelastic.rs:58) applying only inside the logarithm."""
    lam, mu = 400.0, 200.0
    m = oracle.material("viscoelastic", lam, mu, 0.0)
    f = [1.3, 0.2, -0.1, 0.8]
    F = _m(f)
    ref = mu * (F @ F.T - np.eye(2)) + lam * np.log(np.linalg.det(F)) * np.eye(2)
    np.testing.assert_allclose(_m(oracle.kirchhoff(m, _particle(f))), ref, rtol=2e-6, atol=1e-3)
    f = [0.05, 0.0, 0.0, 0.05]  # J = 0.0025 < j_min
    ref = mu * (0.0025 - 1.0) * np.eye(2) + lam * math.log(0.01) * np.eye(2)
    np.testing.assert_allclose(_m(oracle.kirchhoff(m, _particle(f))), ref, rtol=1e-5)


def test_viscoelastic_update_and_timestep(oracle):
    """viscoelastic.rs:132-139: F ← (I + dt C) F with no plastic return;
    viscoelastic.rs:161-190: dt = min(elastic wave dt, cfl_v h² ρ/η)."""
    m = oracle.material("viscoelastic", 400.0, 200.0, 30.0)
    c = [0.1, 0.2, -0.3, 0.05]
    out = oracle.update_particle(m, _particle([1.2, 0.1, 0.0, 0.9], c=c), 0.5)
    ref = (np.eye(2) + 0.5 * _m(c)) @ _m([1.2, 0.1, 0.0, 0.9])
    np.testing.assert_allclose(_m(out["deformation_gradient"]), ref, rtol=1e-6)
    J = np.linalg.det(ref)
    assert abs(out["volume"] - J) < 1e-6 and abs(out["density"] - 1.0 / J) < 1e-5
    # utils.rs:145-164 elastic wave dt = cfl·dx / sqrt((λ+2µ)/ρ); viscous = cfl_v·dx²·ρ/η
    for rho in (2.0, 2000.0):
        el = 0.5 / math.sqrt(800.0 / rho)
        assert oracle.timestep_bound(m, rho, 1.0) == pytest.approx(min(el, 0.5 * rho / 30.0), rel=1e-6)
    assert oracle.timestep_bound(oracle.material("viscoelastic", 400.0, 200.0, 0.0), 2.0, 1.0) == \
        pytest.approx(0.5 / math.sqrt(400.0), rel=1e-6)


# ----------------------------------------------------------------- Rankine
def _rk_step(m, oracle, s1, s2, damage=0.0):  # rankine.rs:247-261
    out = oracle.update_particle(m, _particle([s1, 0, 0, s2], q=damage), 1.0)
    F = out["deformation_gradient"]
    return np.array([F[0], F[3]], np.float64), float(out["friction_hardening"])


def _rk_tau(lam, mu, sig):  # principal Kirchhoff stress of a Hencky strain (rankine.rs:174-178)
    e = np.log(np.abs(sig))
    a = 2 * mu + lam
    return np.array([a * e[0] + lam * e[1], lam * e[0] + a * e[1]])


def test_rankine_inside_cutoff_is_elastic(oracle):
    """rankine.rs:271-287: τ1 = 0.99 σt leaves σ and the damage unchanged."""
    m = oracle.material("rankine", 2000.0, 3000.0, 100.0, 1.0)
    ex = 0.99 * 100.0 / 8000.0
    sig, dmg = _rk_step(m, oracle, math.exp(ex), 1.0)
    assert np.linalg.norm(sig - [math.exp(ex), 1.0]) < 1e-4 and dmg == 0.0


def test_rankine_projects_to_cutoff(oracle):
    """rankine.rs:289-334: τ1 = 1.5 σt lands on τ1 = σt with τ2 unchanged, and the
    damage is |ε_trial − ε_proj| (rankine.rs:189-191), 0.5 σt / (2µ + λ) · ... in closed form."""
    lam, mu = 2000.0, 3000.0
    m = oracle.material("rankine", lam, mu, 100.0, 1.0)
    ex = 1.5 * 100.0 / 8000.0
    sig, dmg = _rk_step(m, oracle, math.exp(ex), 1.0)
    tau = _rk_tau(lam, mu, sig)
    assert abs(tau[0] - 100.0) < 1e-3 and abs(tau[1] - lam * ex) < 1e-3
    # ε_trial − ε_proj = S (τ_trial − τ_proj) with the compliance S of the 2D law, Δτ = (50, 0)
    det = 4 * mu * (mu + lam)
    de = np.array([(2 * mu + lam) * 50.0 / det, -lam * 50.0 / det])
    assert abs(dmg - np.linalg.norm(de)) < 1e-6


def test_rankine_biaxial_corner_and_softening(oracle):
    """rankine.rs:336-361: equal biaxial tension returns both components to the corner;
    rankine.rs:113-117 + utils.rs:43-49: the cutoff softens as σt e^{-s d}, floored at 5 %,
    and the damage saturates at −ln(0.05)/s."""
    lam, mu = 2000.0, 3000.0
    m = oracle.material("rankine", lam, mu, 100.0, 1.0)
    e = 1.5 * 100.0 / (2 * mu + 2 * lam)
    sig, dmg = _rk_step(m, oracle, math.exp(e), math.exp(e))
    np.testing.assert_allclose(_rk_tau(lam, mu, sig), [100.0, 100.0], atol=1e-3)
    assert dmg > 0.0
    # with damage 0.5 the cutoff is 100 e^{-0.5}
    ex = 1.5 * 100.0 / 8000.0
    sig, _ = _rk_step(m, oracle, math.exp(ex), 1.0, damage=0.5)
    assert abs(_rk_tau(lam, mu, sig)[0] - 100.0 * math.exp(-0.5)) < 1e-3
    # past saturation the cutoff is the 5 % residual and the damage stays at the cap
    sat = -math.log(0.05) / 1.0
    sig, dmg = _rk_step(m, oracle, math.exp(ex), 1.0, damage=10.0)
    assert abs(_rk_tau(lam, mu, sig)[0] - 5.0) < 1e-3
    assert abs(dmg - sat) < 1e-6


def test_rankine_stress_is_corotated(oracle):
    """rankine.rs:157-166: the stress is DP's corotated form."""
    f = [1.1, 0.3, -0.2, 0.9]
    a = oracle.kirchhoff(oracle.material("rankine", 400.0, 200.0, 50.0, 1.0), _particle(f))
    b = oracle.kirchhoff(oracle.material("drucker_prager", 400.0, 200.0), _particle(f))
    assert np.array_equal(a, b)


# -------------------------------------------------------------------- µ(I)
def _mui(oracle, lam=2000.0, mu=3000.0):
    return oracle.material("mu_i_default", lam, mu)


def _mui_step(m, oracle, s1, s2, dt):  # sand_mui.rs:260-273 (init_particle: q = µ1)
    out = oracle.update_particle(m, _particle([s1, 0, 0, s2], q=m.params.dp_h0), dt)
    F = out["deformation_gradient"]
    return np.array([F[0], F[3]], np.float64), out["friction_hardening"]


def test_mui_defaults():
    """sand_mui.rs:62-70: µ1 = tan 20.9°, µ2 = tan 32.8°, Q = 5.58."""
    import oracle_py as o
    m = o.material("mu_i_default", 2000.0, 3000.0)
    assert abs(m.params.dp_h0 - math.tan(math.radians(20.9))) < 1e-6
    assert abs(m.params.dp_h1 - math.tan(math.radians(32.8))) < 1e-6
    assert m.params.dp_h2 == np.float32(5.58) and m.params.model == 8


def test_mui_tension_projects_to_identity(oracle):
    """sand_mui.rs:281-296: p_trial ≤ 0 gives σ = 1 exactly and µ = µ1."""
    m = _mui(oracle)
    sig, mu_after = _mui_step(m, oracle, 1.05, 1.02, 0.1)
    assert np.linalg.norm(sig - 1.0) < 1e-5 and mu_after == m.params.dp_h0


def test_mui_inside_static_cone_is_elastic(oracle):
    """sand_mui.rs:298-339: q = 0.99 µ1 p stays elastic with µ = µ1."""
    m = _mui(oracle)
    mu1 = float(m.params.dp_h0)
    ex = -0.01
    p_trial = -(5000.0) * ex
    dev_needed = 0.99 * mu1 * p_trial / (math.sqrt(2) * 3000.0)
    ey = ex - dev_needed * math.sqrt(2)
    sig, mu_after = _mui_step(m, oracle, math.exp(ex), math.exp(ey), 0.1)
    assert np.linalg.norm(sig - [math.exp(ex), math.exp(ey)]) < 1e-4 and mu_after == m.params.dp_h0


def test_mui_rate_is_a_root_of_its_quadratic(oracle):
    """sand_mui.rs:341-405: beyond yield, µ(I) > µ1 and the γ̇ it implies solves
    a γ̇² + b γ̇ + c = 0 (relative residual < 1e-3); the strain moves along −n̂ by γ̇ dt/√2."""
    m = _mui(oracle)
    dt, ex, ey = 0.1, -0.08, 0.02
    sig, mu_i = _mui_step(m, oracle, math.exp(ex), math.exp(ey), dt)
    mu1, mu2, Q, G = float(m.params.dp_h0), float(m.params.dp_h1), float(m.params.dp_h2), 3000.0
    tr = ex + ey
    p = -5000.0 * tr
    dev = np.array([ex - tr / 2, ey - tr / 2])
    dn = np.linalg.norm(dev)
    dq = math.sqrt(2) * G * dn - mu1 * p
    a, sp = G * dt, math.sqrt(p)
    b = p * (mu2 - mu1) + a * Q * sp - dq
    c = -dq * Q * sp
    assert mu_i > mu1
    g = Q * sp / ((mu2 - mu1) / (mu_i - mu1) - 1.0)
    assert abs(a * g * g + b * g + c) / max(abs(a * g * g), abs(b * g), abs(c)) < 1e-3
    eps_new = np.array([ex, ey]) - dev / dn * (g * dt / math.sqrt(2))
    np.testing.assert_allclose(np.log(sig), eps_new, atol=2e-5)


# ------------------------------------------------------------ GranularFluid
def _loam(oracle):  # GranularFluidMaterial::saturated_loam(1e5, 0.2)  granular_fluid.rs:101-117
    lam, mu = oracle.lame_from_young(1.0e5, 0.2)
    m = oracle.material("granular_fluid", lam, mu, 1.0, 200.0, 5.0, 0.4)
    return m, lam, mu


def test_granular_fluid_stress_closed_form_and_symmetric(oracle):
    """granular_fluid.rs:171-194 at the F of physics_correctness.rs:415-440: τ =
    −p I + dev(2µh(F−R)Fᵀ) + λh(J−1)J I with p = max(k((ρ0/J / ρ0)^7 − 1), 0); τ is symmetric."""
    m, lam, mu = _loam(oracle)
    for f, h in (([1.1, 0.05, 0.1, 0.95], 1.0), ([0.9, 0.02, -0.03, 0.85], 1.6)):
        F = _m(f)
        J = np.linalg.det(F)
        U, _, Vt = np.linalg.svd(F)
        R = U @ Vt
        p = max(200.0 * ((1.0 / J) ** 7 - 1.0), 0.0)
        coro = 2 * mu * h * (F - R) @ F.T
        ref = -p * np.eye(2) + coro - 0.5 * np.trace(coro) * np.eye(2) + lam * h * (J - 1) * J * np.eye(2)
        tau = _m(oracle.kirchhoff(m, _particle(f, h=h)))
        np.testing.assert_allclose(tau, ref, rtol=2e-5, atol=2e-2)
        assert abs(tau[0, 1] - tau[1, 0]) / max(abs(tau[0, 1]), 1.0) < 1e-5
        if h == 1.6:
            assert p > 100.0  # the compressed case exercises the EOS term


def test_granular_fluid_clamp_matches_snow_form(oracle):
    """granular_fluid.rs:200-230: σ clamps to [1 − θc, 1 + θs], Jp = Jp σ1σ2/(σc1σc2)
    clamped to [0.2, 3], h = clamp(exp(ξ(1 − Jp)), 0.1, 7)."""
    m, _, _ = _loam(oracle)
    out = oracle.update_particle(m, _particle([1.05, 0, 0, 0.8], jp=1.0), 1.0)
    F = out["deformation_gradient"]
    assert abs(F[0] - 1.01) < 1e-6 and abs(F[3] - 0.8) < 1e-6
    jp = 1.05 / 1.01
    assert abs(out["plastic_volume_ratio"] - jp) < 1e-6
    assert abs(out["hardening_scale"] - min(max(math.exp(5.0 * (1 - jp)), 0.1), 7.0)) < 1e-5
    assert abs(out["volume"] - 1.01 * 0.8) < 1e-6
    # both limits 0: F is advanced with no projection
    m.params.compression_limit = 0.0
    m.params.stretch_limit = 0.0
    out = oracle.update_particle(m, _particle([1.2, 0, 0, 0.7], c=[0.1, 0.0, 0.0, -0.2]), 0.5)
    np.testing.assert_allclose(out["deformation_gradient"], [1.2 * 1.05, 0, 0, 0.7 * 0.9], rtol=1e-6)


def test_granular_fluid_mass_conserved(oracle):
    """physics_correctness.rs:146-159: saturated loam, 100 zero-gravity steps."""
    m, _, _ = _loam(oracle)
    sim = oracle.OracleSim.new(_zero_gravity(32), _center_spawn(32, 6))
    sim.with_default_material(m)
    m0 = sim.particles()["mass"].sum()
    sim.step_n(100)
    p = sim.particles()
    assert abs(p["mass"].sum() - m0) < 1e-6
    F = p["deformation_gradient"]
    assert np.min(F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]) > 0


# ------------------------------------------------------------------ Bingham
def _bingham(oracle):  # BinghamFluidMaterial::new(1000, 0.5, 5000, 7, 100)  bingham.rs:262-263
    return oracle.material("bingham", 1000.0, 0.5, 5000.0, 7.0, 100.0)


def test_bingham_plug_below_critical_rate(oracle):
    """bingham.rs:261-274: a sub-critical shear rate gives zero deviatoric stress, so
    τ is the Tait pressure alone."""
    m = _bingham(oracle)
    tau = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=[0.0, 1e-6, 1e-6, 0.0], density=1000.0)))
    assert np.array_equal(tau, np.zeros((2, 2)))
    tau = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=[0.0, 1e-6, 1e-6, 0.0], density=1100.0)))
    p = 5000.0 * (1.1 ** 7 - 1.0)
    np.testing.assert_allclose(tau, -p * np.eye(2), rtol=1e-5)


def test_bingham_formula_above_yield(oracle):
    """bingham.rs:276-298: pure shear C = [[0, g], [g, 0]] gives γ̇ = 2g and
    τ = (τ0/γ̇ + η) C; bingham.rs:300-318: the apparent viscosity falls with the rate."""
    m = _bingham(oracle)
    for g in (1.0, 5.0, 10.0):
        C = [0.0, g, g, 0.0]
        tau = _m(oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=C, density=1000.0)))
        pred = (100.0 / (2 * g) + 0.5) * _m(C)
        assert np.linalg.norm(tau - pred) < 1e-3
    slow = oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=[0, 1.0, 1.0, 0], density=1000.0))[1] / 2.0
    fast = oracle.kirchhoff(m, _particle([1, 0, 0, 1], c=[0, 10.0, 10.0, 0], density=1000.0))[1] / 20.0
    assert fast < slow


def test_bingham_record_is_a_fluid(oracle):
    """bingham.rs:205-218: params() reports model Fluid with the yield stress in
    compression_limit; a Newtonian record (compression_limit 0) keeps the Newtonian law,
    whose viscous term is η (C + Cᵀ)_dev, twice the Bingham η D_dev."""
    m = _bingham(oracle)
    assert m.params.model == 1 and m.params.compression_limit == 100.0 and m.params.pressure_floor == 0.0
    newt = oracle.material("fluid", 1000.0, 0.5, 5000.0, 7.0)
    C = [0.0, 3.0, 3.0, 0.0]
    tn = _m(oracle.kirchhoff(newt, _particle([1, 0, 0, 1], c=C, density=1000.0)))
    np.testing.assert_allclose(tn, 0.5 * 2 * _m(C), rtol=1e-6)
    # a zero-yield Bingham record (flag in the extension) is η D_dev above the critical rate
    m0 = oracle.material("bingham", 1000.0, 0.5, 5000.0, 7.0, 0.0)
    tb = _m(oracle.kirchhoff(m0, _particle([1, 0, 0, 1], c=C, density=1000.0)))
    np.testing.assert_allclose(tb, 0.5 * _m(C), rtol=1e-6)


# ------------------------------------------------------------------- NACC
def _pressure(kappa, j):  # nacc.rs:326-329
    return -(kappa * 0.5 * (j - 1.0 / j)) * j


def test_nacc_cap_projects_to_p0(oracle):
    """nacc.rs:337-360: isotropic compression past p0 lands on p = p0 (within 1e-2)."""
    m = oracle.material("nacc", 3000.0, 2000.0, 1.2, 0.0, 0.0)
    p0 = 2000.0 * 1e-5
    j = 0.5
    assert _pressure(2000.0, j) > p0
    f, _ = oracle.nacc_project(m, [math.sqrt(j), 0, 0, math.sqrt(j)], 0.0)
    ja = f[0] * f[3] - f[1] * f[2]
    assert abs(_pressure(2000.0, ja) - p0) < 1e-2


def test_nacc_elastic_interior_untouched(oracle):
    """nacc.rs:383-401: with α = −1 (p0 ≈ 7254 for κ = 2000, ξ = 2) a confined
    small strain is not projected and α is unchanged."""
    m = oracle.material("nacc", 3000.0, 2000.0, 1.2, 0.0, 2.0)
    p0 = 2000.0 * (1e-5 + math.sinh(2.0))
    assert abs(p0 - 7254) < 1.0
    f_in = [0.99895, 0, 0, 0.99905]
    f, a = oracle.nacc_project(m, f_in, -1.0)
    assert np.max(np.abs(f - np.array(f_in, np.float32))) < 1e-6
    assert a == -1.0


def test_nacc_shear_yield_lands_on_ellipse(oracle):
    """Case C (nacc.rs:214-222) puts the trial state on the yield ellipse: the
    projected state has y ≈ 0 (y = (1+2β)·2·|s|² + M²(p + βp0)(p − p0))."""
    mu, kappa, M, beta, xi = 3000.0, 2000.0, 1.2, 0.3, 0.0
    m = oracle.material("nacc", mu, kappa, M, beta, xi)
    f_in = [1.02, 0.0, 0.0, 0.99]   # J < 1 (compression) with shear
    f, _ = oracle.nacc_project(m, f_in, 0.0)
    s = np.linalg.svd(_m(f), compute_uv=False)
    J = s[0] * s[1]
    dev = mu / J * (s ** 2 - 0.5 * np.sum(s ** 2))
    p = _pressure(kappa, J)
    p0 = kappa * 1e-5
    y = (1 + 2 * beta) * 2 * np.sum(dev ** 2) + M * M * (p + beta * p0) * (p - p0)
    y_tr_s = np.array([1.02, 0.99])
    Jt = 1.02 * 0.99
    dev_t = mu / Jt * (y_tr_s ** 2 - 0.5 * np.sum(y_tr_s ** 2))
    y_tr = (1 + 2 * beta) * 2 * np.sum(dev_t ** 2) + M * M * (_pressure(kappa, Jt) + beta * p0) * (_pressure(kappa, Jt) - p0)
    assert y_tr > 1e-4 and abs(y) < 1e-3 * abs(y_tr)


# ------------------------------------------------------- boundaries / misc
def test_slip_wall_laws(oracle):
    """boundary/mod.rs:122-152: tangential velocity passes unchanged, inward
    normal blocked, outward untouched; all four walls."""
    assert oracle.slip_velocity(2, 32, 64, (-3.0, 7.5)) == (0.0, 7.5)
    assert oracle.slip_velocity(2, 32, 64, (4.0, -2.0)) == (4.0, -2.0)
    r = 64
    assert oracle.slip_velocity(2, 62 * r + 30, r, (3.0, 1.0)) == (0.0, 1.0)      # right wall
    assert oracle.slip_velocity(2, 30 * r + 1, r, (1.0, -5.0)) == (1.0, 0.0)      # floor
    assert oracle.slip_velocity(2, 30 * r + 62, r, (1.0, 5.0)) == (1.0, 0.0)      # ceiling
    assert oracle.slip_velocity(2, 30 * r + 30, r, (-9.0, -9.0)) == (-9.0, -9.0)  # interior


def test_coulomb_wall_laws(oracle):
    """boundary/mod.rs:154-222 through FrictionBoundary (friction.rs:40-59) on the
    left wall (x = 0, y = 32 of a 64 grid, thickness 2): µ = 0 is slip with the
    normal zeroed; the tangential speed drops by exactly µ|v_n|; it never reverses;
    outward velocity is untouched."""
    bv = oracle.boundary_velocity
    assert bv("friction", 32, 64, (-3.0, 7.5), mu=0.0) == (0.0, 7.5)
    vx, vy = bv("friction", 32, 64, (-4.0, 10.0), mu=0.3)
    assert vx == 0.0 and abs(vy - (10.0 - 0.3 * 4.0)) < 1e-5
    assert bv("friction", 32, 64, (-10.0, 2.0), mu=0.9) == (0.0, 0.0)
    assert bv("friction", 32, 64, (4.0, -2.0), mu=0.9) == (4.0, -2.0)
    # floor (y < t, normal +Y) and right wall (x > R - t - 1, normal -X)
    r = 64
    vx, vy = bv("friction", 30 * r + 1, r, (6.0, -2.0), mu=0.5)
    assert vy == 0.0 and abs(vx - 5.0) < 1e-6
    vx, vy = bv("friction", 62 * r + 30, r, (2.0, -3.0), mu=0.5)
    assert vx == 0.0 and abs(vy + 2.0) < 1e-6
    # a corner cell is processed wall by wall: left wall then floor
    vx, vy = bv("friction", 0 * r + 0, r, (-1.0, -4.0), mu=0.25)
    assert (vx, vy) == (0.0, 0.0) or (vx == 0.0 and abs(vy) < 1e-6)
    assert bv("friction", 30 * r + 30, r, (-9.0, -9.0), mu=1.0) == (-9.0, -9.0)


def test_heightmap_boundary_laws(oracle):
    """heightmap.rs:66-114: cells at or below the terrain lose downward velocity and
    µ|v_n| of their horizontal speed; slip side and top walls, no bottom wall; heights
    past the end repeat the last one; particles are pushed to terrain + 1."""
    r = 64
    hts = [3.0] * 20 + [3.0 + 0.5 * k for k in range(10)]  # 30 columns, ramp at the end
    bv = lambda idx, v, mu=0.4: oracle.boundary_velocity("heightmap", idx, r, v, mu=mu, heights=hts)
    vx, vy = bv(10 * r + 3, (5.0, -2.0))             # on the floor: y = 3 <= 3
    assert vy == 0.0 and abs(vx - 5.0 * (1 - 0.8 / 5.0)) < 1e-6
    assert bv(10 * r + 4, (5.0, -2.0)) == (5.0, -2.0)  # above the terrain
    assert bv(10 * r + 3, (5.0, 2.0)) == (5.0, 2.0)    # moving up
    assert bv(10 * r + 3, (0.5, -2.0)) == (0.0, 0.0)   # friction stops it
    assert bv(10 * r + 3, (5.0, -2.0), mu=0.0) == (5.0, 0.0)
    assert bv(50 * r + 8, (1.0, -1.0)) == (1.0, -1.0)  # column 50 -> last height 7.5: 8 is above
    vx, vy = bv(50 * r + 7, (1.0, -1.0))                 # 7 <= 7.5: blocked
    assert vy == 0.0 and abs(vx - 0.6) < 1e-6
    assert bv(1 * r + 30, (-2.0, 1.0)) == (0.0, 1.0)    # left wall
    assert bv(30 * r + 62, (1.0, 3.0)) == (1.0, 0.0)    # top wall
    assert bv(30 * r + 1, (1.0, -3.0))[1] == 0.0        # floor comes from the terrain alone
    bc = lambda p: oracle.boundary_clamp("heightmap", p, r, heights=hts)
    assert bc((10.5, 2.0)) == (10.5, 4.0)
    assert bc((10.5, 9.0)) == (10.5, 9.0)
    assert bc((-5.0, 3.5)) == (1.0, 4.0)                # wall clamp, then column 1
    assert bc((100.0, 5.0)) == (62.0, 8.5)              # x clamped to R - t, last height 7.5


def test_predictive_and_friction_clamps(oracle):
    """predictive.rs:31-35: clamp to [wall_min, R - 1 - wall_min]; friction.rs:61-63 and
    mod.rs:103-111: [t - 1, R - t]."""
    assert oracle.boundary_clamp("predictive", (0.5, 70.0), 64, wall_min=3.5) == (3.5, 59.5)
    assert oracle.boundary_clamp("friction", (0.5, 70.0), 64, thickness=3) == (2.0, 61.0)
    assert oracle.boundary_velocity("predictive", 32, 64, (-3.0, 7.5)) == (0.0, 7.5)


def test_lcg_closed_form(oracle):
    """LcgRng (solver/mod.rs:154-175): s <- s·1664525 + 1013904223 mod 2^32; f = s / 2^32 in f32."""
    u, f = oracle.lcg(1, 5)
    s, exp_u = 1, []
    for _ in range(5):
        s = (s * 1664525 + 1013904223) % (1 << 32)
        exp_u.append(s)
    assert list(u) == exp_u
    assert np.array_equal(f, (np.array(exp_u, np.float64).astype(np.float32) / np.float32(4294967296.0)))


def test_lame_and_radians(oracle):
    lam, mu = oracle.lame_from_young(3.5e5, 0.3)  # utils.rs:183-193
    assert abs(lam - 3.5e5 * 0.3 / (1.3 * 0.4)) < 0.05 and abs(mu - 3.5e5 / 2.6) < 0.05
    assert oracle.to_radians(180.0) == np.float32(np.pi)


def test_elastic_wave_dt_closed_form(oracle):
    """utils.rs:145-164: dt = cfl·dx / sqrt((λ+2µ)h/ρ)."""
    m = oracle.material("corotated", 400.0, 200.0)
    assert abs(oracle.timestep_bound(m, 4.0, 2.0) - 0.5 / math.sqrt(800.0 * 2.0 / 4.0)) < 1e-7
    m = oracle.material("drucker_prager", 201923.08, 134615.38)   # h ignored for DP
    assert abs(oracle.timestep_bound(m, 16.0, 5.0) - 0.5 / math.sqrt(471153.84 / 16.0)) < 1e-7


# --------------------------------------------------------- whole-step laws
def _center_spawn(grid_res, side):  # physics_correctness.rs:53-61
    return eh.SpawnRegion(spacing=0.5, box_size=(side, side), box_center=(grid_res * 0.5, grid_res * 0.5))


def _zero_gravity(grid_res):  # physics_correctness.rs:43-51
    return eh.SimConfig.default().replace(grid_res=grid_res, dt=0.05, gravity=(0.0, 0.0), adaptive_timestep=1)


def test_quickstart_lattice(oracle):
    """README.md:32-45: 12x12 and 12x8 boxes at spacing 1 -> 144 + 96 particles at
    integer lattice positions; add_body estimates volumes (particles.rs:414-420)."""
    sim = S.quickstart("oracle32")
    p = sim.particles()
    assert len(p) == 240
    b1, b2 = p[:144], p[144:]
    assert set(np.unique(b1["x"][:, 0])) == set(np.arange(18, 30, dtype=np.float32))
    assert set(np.unique(b1["x"][:, 1])) == set(np.arange(34, 46, dtype=np.float32))
    assert set(np.unique(b2["x"][:, 1])) == set(np.arange(32, 40, dtype=np.float32))
    assert np.all(b2["material_id"] == 1) and np.all(b1["user_tag"] == 1) and np.all(b2["user_tag"] == 2)
    # interior particle: density Σ w·m over the lattice = 1, corner = 0.75^2
    assert np.isclose(b1["density"].max(), 1.0) and np.isclose(b1["density"].min(), 0.5625)
    assert np.allclose(b1["initial_volume"], 1.0 / b1["density"])


@pytest.mark.parametrize("mat", ["neohookean", "fluid", "snow"])
def test_mass_conserved_100_steps(mat, oracle):
    """physics_correctness.rs:91-148: total particle mass unchanged after 100 steps."""
    cfg = _zero_gravity(32)
    if mat == "fluid":
        cfg.recompute_density_each_step = 1
    sim = oracle.OracleSim.new(cfg, _center_spawn(32, 6))
    if mat == "neohookean":
        sim.with_default_material(oracle.material("neohookean", 10.0, 20.0))
    elif mat == "fluid":
        sim.with_default_material(oracle.material("fluid", 4.0, 0.1, 10.0, 4.0))
    else:
        lam, mu = oracle.lame_from_young(1.4e5, 0.2)
        sim.with_default_material(oracle.material("snow", lam, mu, 10.0, 0.025, 0.0075, 0.6, 20.0))
    m0 = sim.particles()["mass"].sum()
    sim.step_n(100)
    assert abs(sim.particles()["mass"].sum() - m0) < 1e-6


def test_zero_velocity_spawn_momentum_drift(oracle):
    """physics_correctness.rs:166-183: |Δp| per particle < 1e-3 after 50 steps."""
    sim = oracle.OracleSim.new(_zero_gravity(32), _center_spawn(32, 8))
    sim.with_default_material(oracle.material("neohookean", 20.0, 40.0))
    p0 = (sim.particles()["mass"][:, None] * sim.particles()["v"]).sum(0)
    sim.step_n(50)
    p = sim.particles()
    p1 = (p["mass"][:, None] * p["v"]).sum(0)
    assert np.linalg.norm(p1 - p0) / len(p) < 1e-3


def test_gravity_impulse(oracle):
    """physics_correctness.rs:185-219: Δp = m·g·t within 5% (fixed dt 0.01, 10 steps)."""
    cfg = eh.SimConfig.default().replace(gravity=(0.0, -9.81), dt=0.01, adaptive_timestep=0)
    sim = oracle.OracleSim.new(cfg, _center_spawn(64, 4))
    sim.with_default_material(oracle.material("neohookean", 100.0, 200.0))
    p = sim.particles()
    m = p["mass"].sum()
    sim.step_n(10)
    p = sim.particles()
    imp = (p["mass"][:, None] * p["v"]).sum(0)
    exp = np.array([0.0, -9.81]) * m * 0.1
    assert np.linalg.norm(imp - exp) / np.linalg.norm(exp) < 0.05


def test_j_stays_positive(oracle):
    """physics_correctness.rs:221-233: J > 0 with projection on, NeoHookean, 200 steps."""
    sim = oracle.OracleSim.new(eh.SimConfig.standard(64, 0.05, (0.0, -9.81)), _center_spawn(64, 8))
    sim.with_default_material(oracle.material("neohookean", 10.0, 20.0))
    sim.step_n(200)
    F = sim.particles()["deformation_gradient"]
    assert np.min(F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]) > 0


def test_p2g_conserves_mass_and_momentum(oracle):
    """Σ grid mass = Σ m_p and Σ grid momentum = Σ m_p v_p + 0: the affine and
    stress terms carry Σ w·(x_i − x_p) = 0 (diagnostics/snapshot.rs:320-360
    defines these errors).  Checked in f64 to roundoff."""
    sim = S.quickstart_impact("oracle64")
    sim.step_n(3)  # non-trivial v, C and stress
    p = sim.particles_f64()
    m, mom = sim.p2g_grid(1e-3)
    assert abs(m.sum() - p[:, 12].sum()) < 1e-9 * p[:, 12].sum()
    pm = (p[:, 12:13] * p[:, 2:4]).sum(0)
    assert np.max(np.abs(mom.sum(0) - pm)) < 1e-8 * max(1.0, np.abs(p[:, 12:13] * p[:, 2:4]).sum())


def test_apic_reproduces_rigid_rotation(oracle):
    """APIC transfers an affine velocity field exactly (Jiang et al. 2015): a
    rigidly rotating, stress-free body keeps v = ω×r and C = [[0, −ω], [ω, 0]]
    through one full P2G/G2P substep (no gravity, Fallback material: no stress)."""
    cfg = eh.SimConfig.unsafe_defaults().replace(grid_res=64, dt=1e-4, min_dt=1e-4, gravity=(0.0, 0.0))
    sim = oracle.OracleSim.empty(cfg, 64)
    pts = eh.zeroed_particles(0)
    xs = np.stack(np.meshgrid(np.arange(24, 40, 0.5), np.arange(24, 40, 0.5), indexing="ij"), -1).reshape(-1, 2)
    pts = eh.zeroed_particles(len(xs))
    w, c = 0.3, np.array([32.0, 32.0])
    r = xs - c
    pts["x"] = xs
    pts["v"] = np.stack([-w * r[:, 1], w * r[:, 0]], -1)
    pts["velocity_gradient"] = (0.0, w, -w, 0.0)  # column-major [[0, -w], [w, 0]]
    pts["mass"] = 1.0
    pts["initial_volume"] = pts["volume"] = 0.25
    pts["density"] = 4.0
    sim.set_particles(pts)
    sim.step_n(1)
    p = sim.particles_f64()
    r_new = p[:, 0:2] - c
    # inputs are f32-rounded samples of the field, so agreement is to ~1e-7;
    # a transposed C or a wrong sign of (x_i - x_p) gives O(w) errors instead
    np.testing.assert_allclose(p[:, 4:8], np.tile([0.0, w, -w, 0.0], (len(p), 1)), atol=1e-6)
    # velocity of the (moved) particle matches the field at its old position
    np.testing.assert_allclose(p[:, 2:4], pts["v"], atol=1e-6)
    assert np.all(np.abs(np.linalg.norm(r_new, axis=1) - np.linalg.norm(r, axis=1)) < 1e-3)


def test_f32_oracle_tracks_f64_oracle(oracle):
    """The f32 restatement and the f64 oracle agree to f32 roundoff on the
    quick-start scene after one step (same substep schedule)."""
    a, b = S.quickstart("oracle32"), S.quickstart("oracle64")
    a.step_n(1)
    b.step_n(1)
    assert a.stats().last_substeps == b.stats().last_substeps
    e = S.rel_errors(S.state_matrix(a.particles()), S.f64_state(b.particles_f64()))
    assert e["x"] < 1e-6 and e["v"] < 1e-4 and e["F"] < 1e-5, e
