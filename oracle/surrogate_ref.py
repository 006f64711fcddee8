"""CPU reference of the learned NACC return-mapping surrogate (numpy).

TEST INFRASTRUCTURE ONLY: tests and bench.py's checks import this as the
checker of the device network; the library never calls it.

The reference crate has no surrogate; its return mapping is
NaccMaterial::project (matter/materials/nacc.rs:129-223), which the oracle
restates in mpm_oracle.hpp (`nacc_project`) and which labels every training
sample.  This module restates the network the library defines on top of it:

    z_in = (x - in_mu) / in_sd                       x = (ln s1, ln s2, alpha, M, beta, xi)
    h1 = silu(W1 z_in + b1), h2 = silu(W2 h1 + b2), h3 = silu(W3 h2 + b3)
    r  = W4 h3 + b4                                  normalised residual
    y  = x[0:3] + r * out_sd + out_mu                (ln s1', ln s2', alpha')
    loss = mean over samples of sum_k (r_k - (y*_k - x_k - out_mu_k) / out_sd_k)^2

Two evaluations of each quantity:
  * `exact=True`: float64 throughout (the mathematical network; the
    finite-difference test of `gradients` runs on this one);
  * `exact=False`: the device's arithmetic -- every MFMA operand (weights,
    normalised inputs, activations, backward signals) rounded to bfloat16
    (round to nearest even), products summed in float64 (an MFMA sums exact
    bf16 x bf16 products in FP32; the order differs, so compare with a
    tolerance), biases and element-wise work in float32.
"""
import numpy as np

N_IN, N_HID, N_OUT = 6, 64, 3
W1, B1 = 0, N_HID * N_IN
W2 = B1 + N_HID
B2 = W2 + N_HID * N_HID
W3 = B2 + N_HID
B3 = W3 + N_HID * N_HID
W4 = B3 + N_HID
B4 = W4 + N_OUT * N_HID
N_PARAMS = B4 + N_OUT  # 8963
IN_MU, IN_SD, OUT_MU, OUT_SD, Z_MAX = 0, 6, 12, 15, 18


def bf16(a):
    """float32 -> nearest-even bfloat16, returned as float32 (finite inputs)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[W1:B1].reshape(N_HID, N_IN), w[B1:W2], w[W2:B2].reshape(N_HID, N_HID), w[B2:W3],
            w[W3:B3].reshape(N_HID, N_HID), w[B3:W4], w[W4:B4].reshape(N_OUT, N_HID), w[B4:N_PARAMS])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _q(a, exact):
    """An MFMA operand: bf16-rounded unless exact."""
    return np.asarray(a, dtype=np.float64) if exact else bf16(np.asarray(a, dtype=np.float32)).astype(np.float64)


def _f32(a, exact):
    return a if exact else np.asarray(a, dtype=np.float32).astype(np.float64)


def normalise(x, norm, exact=True):
    x = np.asarray(x, dtype=np.float64 if exact else np.float32)
    nm = np.asarray(norm, dtype=x.dtype)
    return ((x - nm[IN_MU:IN_MU + 6]) / nm[IN_SD:IN_SD + 6]).astype(np.float64)


def forward(w, norm, x, exact=True):
    """Normalised residual r [n, 3] and the intermediates the backward pass needs."""
    a1, c1, a2, c2, a3, c3, a4, c4 = unpack(w)
    zin = normalise(x, norm, exact)
    h0 = _q(zin, exact)
    z1 = _f32(h0 @ _q(a1, exact).T + _f32(c1, exact), exact)
    h1 = _q(_silu(z1), exact)
    z2 = _f32(h1 @ _q(a2, exact).T + _f32(c2, exact), exact)
    h2 = _q(_silu(z2), exact)
    z3 = _f32(h2 @ _q(a3, exact).T + _f32(c3, exact), exact)
    h3 = _q(_silu(z3), exact)
    r = _f32(h3 @ _q(a4, exact).T + _f32(c4, exact), exact)
    return r, (h0, z1, h1, z2, h2, z3, h3)


def predict(w, norm, x, exact=True):
    """Raw outputs y = x[0:3] + r * out_sd + out_mu."""
    r, _ = forward(w, norm, x, exact)
    nm = np.asarray(norm, dtype=np.float64)
    return np.asarray(x, dtype=np.float64)[:, :3] + (r * nm[OUT_SD:OUT_SD + 3] + nm[OUT_MU:OUT_MU + 3])


def targets(norm, x, y, exact=True):
    dt = np.float64 if exact else np.float32
    x = np.asarray(x, dtype=dt)
    y = np.asarray(y, dtype=dt)
    nm = np.asarray(norm, dtype=dt)
    return (((y - x[:, :3]) - nm[OUT_MU:OUT_MU + 3]) / nm[OUT_SD:OUT_SD + 3]).astype(np.float64)


def loss(w, norm, x, y, exact=True):
    r, _ = forward(w, norm, x, exact)
    e = r - targets(norm, x, y, exact)
    return float(np.mean(np.sum(e * e, axis=1)))


def gradients(w, norm, x, y, exact=True):
    """(d loss / d params [N_PARAMS], loss) by the chain rule."""
    a1, c1, a2, c2, a3, c3, a4, c4 = unpack(w)
    r, (h0, z1, h1, z2, h2, z3, h3) = forward(w, norm, x, exact)
    n = len(r)
    e = r - targets(norm, x, y, exact)
    d_out = _f32(2.0 * e * _f32(1.0 / n, exact), exact)
    d_out_q = _q(d_out, exact)
    dz3 = _f32((d_out_q @ _q(a4, exact)) * _silu_grad(z3), exact)
    dz3_q = _q(dz3, exact)
    dz2 = _f32((dz3_q @ _q(a3, exact)) * _silu_grad(z2), exact)
    dz2_q = _q(dz2, exact)
    dz1 = _f32((dz2_q @ _q(a2, exact)) * _silu_grad(z1), exact)
    dz1_q = _q(dz1, exact)
    g = np.zeros(N_PARAMS)
    g[W1:B1] = (dz1_q.T @ h0).ravel()
    g[B1:W2] = dz1_q.sum(0)
    g[W2:B2] = (dz2_q.T @ h1).ravel()
    g[B2:W3] = dz2_q.sum(0)
    g[W3:B3] = (dz3_q.T @ h2).ravel()
    g[B3:W4] = dz3_q.sum(0)
    g[W4:B4] = (d_out_q.T @ h3).ravel()
    g[B4:N_PARAMS] = d_out_q.sum(0)
    return g, float(np.mean(np.sum(e * e, axis=1)))


def adam_step(w, m, v, g, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1) in float64."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def nacc_kirchhoff(f, mu, kappa):
    """NACC Kirchhoff stress (nacc.rs:236-249): (mu / J) dev(F F^T) + (kappa / 2)(J^2 - 1) I,
    F [n, 4] column-major (a, b, c, d) = [[a, c], [b, d]]; returns [n, 4] column-major."""
    f = np.asarray(f, dtype=np.float64)
    a, b, c, d = f[:, 0], f[:, 1], f[:, 2], f[:, 3]
    j = np.maximum(a * d - c * b, 1e-6)
    b00 = a * a + c * c
    b11 = b * b + d * d
    b01 = a * b + c * d
    hb = 0.5 * (b00 + b11)
    vol = 0.5 * kappa * (j * j - 1.0)
    s = mu / j
    return np.stack([s * (b00 - hb) + vol, s * b01, s * b01, s * (b11 - hb) + vol], axis=1)
