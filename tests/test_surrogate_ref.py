"""Pins the CPU reference of the learned NACC return mapping (oracle/surrogate_ref.py)
against what the mathematics fixes: bfloat16 rounding known answers, the
chain-rule gradients against central finite differences, Adam's first step,
and the NACC Kirchhoff stress at a diagonal F (nacc.rs:236-249)."""
import numpy as np
import pytest

import surrogate_ref as R


def _net(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, scale, R.N_PARAMS)
    norm = np.zeros(20)
    norm[R.IN_MU:R.IN_MU + 6] = rng.normal(0.0, 0.5, 6)
    norm[R.IN_SD:R.IN_SD + 6] = rng.uniform(0.5, 2.0, 6)
    norm[R.OUT_MU:R.OUT_MU + 3] = rng.normal(0.0, 0.1, 3)
    norm[R.OUT_SD:R.OUT_SD + 3] = rng.uniform(0.2, 1.0, 3)
    norm[R.Z_MAX] = 8.0
    return w, norm


def test_layout_sizes():
    assert R.N_PARAMS == 8963
    assert (R.B1, R.W2, R.B2, R.W3, R.B3, R.W4, R.B4) == (384, 448, 4544, 4608, 8704, 8768, 8960)


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {
        1.0: 1.0,
        one + 2.0 ** -8: 1.0,                             # tie, even mantissa below
        one + 3 * 2.0 ** -8: one + 2.0 ** -6,              # tie, even mantissa above
        one + 2.0 ** -8 + 2.0 ** -20: one + 2.0 ** -7,     # above the tie
        -(one + 2.0 ** -8 + 2.0 ** -20): -(one + 2.0 ** -7),
        0.1: 0.10009765625,                                # 0x3DCCCCCD -> 0x3DCD
        65280.0: 65280.0,                                  # exactly representable (0x477f)
    }
    x = np.array(list(cases.keys()), dtype=np.float32)
    want = np.array(list(cases.values()), dtype=np.float32)
    got = R.bf16(x)
    assert np.array_equal(got, want), (got, want)
    # idempotent, and at most half a bf16 ulp away
    rng = np.random.default_rng(0)
    v = rng.normal(0.0, 10.0, 10000).astype(np.float32)
    b = R.bf16(v)
    assert np.array_equal(R.bf16(b), b)
    ulp = 2.0 ** (np.floor(np.log2(np.abs(v.astype(np.float64)))) - 7)
    assert np.all(np.abs(b.astype(np.float64) - v) <= 0.5 * ulp + 1e-30)


def test_gradients_match_finite_differences():
    w, norm = _net(1)
    rng = np.random.default_rng(2)
    x = rng.normal(0.0, 1.0, (7, 6))
    y = x[:, :3] + rng.normal(0.0, 0.2, (7, 3))
    g, l0 = R.gradients(w, norm, x, y, exact=True)
    assert l0 == pytest.approx(R.loss(w, norm, x, y, exact=True))
    idx = np.concatenate([rng.choice(R.N_PARAMS, 300, replace=False), np.arange(R.B4, R.N_PARAMS),
                          np.arange(R.B1, R.B1 + 5), [R.W1, R.W4 + 63]])
    h = 1e-6
    fd = []
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd.append((R.loss(wp, norm, x, y) - R.loss(wm, norm, x, y)) / (2 * h))
    fd = np.array(fd)
    err = np.abs(fd - g[idx])
    assert err.max() <= 1e-6 * max(1.0, np.abs(g).max()), err.max()


def test_planted_gradient_mistake_is_caught():
    """A dropped SiLU derivative in the first hidden layer changes the W1 gradient."""
    w, norm = _net(3)
    rng = np.random.default_rng(4)
    x = rng.normal(0.0, 1.0, (5, 6))
    y = x[:, :3] + rng.normal(0.0, 0.2, (5, 3))
    g, _ = R.gradients(w, norm, x, y, exact=True)
    saved = R._silu_grad
    try:
        R._silu_grad = lambda z: np.ones_like(z)
        g_bad, _ = R.gradients(w, norm, x, y, exact=True)
    finally:
        R._silu_grad = saved
    assert np.abs(g_bad[R.W1:R.B1] - g[R.W1:R.B1]).max() > 1e-3 * np.abs(g[R.W1:R.B1]).max()


def test_emulated_gradients_close_to_exact():
    """bf16 operands perturb the gradient by a few parts in a thousand, not more."""
    w, norm = _net(5, scale=0.15)
    rng = np.random.default_rng(6)
    x = rng.normal(0.0, 1.0, (512, 6))
    y = x[:, :3] + rng.normal(0.0, 0.2, (512, 3))
    g, _ = R.gradients(w, norm, x, y, exact=True)
    ge, _ = R.gradients(w, norm, x, y, exact=False)
    assert np.linalg.norm(ge - g) <= 3e-2 * np.linalg.norm(g)


def test_adam_first_step_is_signed_lr():
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    w, m, v = R.adam_step(np.zeros(4), np.zeros(4), np.zeros(4), g, 1, lr=0.01, eps=1e-12)
    assert np.allclose(w[:3], -0.01 * np.sign(g[:3]), rtol=1e-9)
    assert w[3] == 0.0
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_kirchhoff_diagonal_closed_form():
    mu, kappa = 76923.0, 166666.7
    s1, s2 = 1.1, 0.95
    f = np.array([[s1, 0.0, 0.0, s2]])
    j = s1 * s2
    want = np.array([mu / j * (s1 * s1 - 0.5 * (s1 * s1 + s2 * s2)) + 0.5 * kappa * (j * j - 1.0), 0.0, 0.0,
                     mu / j * (s2 * s2 - 0.5 * (s1 * s1 + s2 * s2)) + 0.5 * kappa * (j * j - 1.0)])
    assert np.allclose(R.nacc_kirchhoff(f, mu, kappa)[0], want, rtol=1e-12)
    # rotation covariance: tau(Q F) = Q tau(F) Q^T
    th = 0.3
    q = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    fm = np.array([[1.05, 0.1], [-0.05, 0.9]])
    qf = q @ fm
    col = lambda m: np.array([[m[0, 0], m[1, 0], m[0, 1], m[1, 1]]])
    t0 = R.nacc_kirchhoff(col(fm), mu, kappa)[0]
    t1 = R.nacc_kirchhoff(col(qf), mu, kappa)[0]
    tm = np.array([[t0[0], t0[2]], [t0[1], t0[3]]])
    rot = q @ tm @ q.T
    assert np.allclose(t1, [rot[0, 0], rot[1, 0], rot[0, 1], rot[1, 1]], rtol=1e-10, atol=1e-6)
