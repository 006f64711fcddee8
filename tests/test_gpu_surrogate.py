"""Learned NACC return mapping on the GPU, through the C-ABI: the network's
forward pass and gradients against the CPU reference (oracle/surrogate_ref.py)
with the device's bfloat16 operand rounding emulated, the plain batch return
mapping against the oracle's NaccMaterial::project (nacc.rs:129-223), the
yield-check gate and fallback, harvesting from the substep, training, and
the substep's surrogate mode."""
import numpy as np
import pytest

import emerge_hip as eh
import scenes as S
import surrogate_ref as R


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _net(seed, scale=0.15):
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, scale, R.N_PARAMS).astype(np.float32)
    norm = np.zeros(20, np.float32)
    norm[R.IN_MU:R.IN_MU + 6] = rng.normal(0.0, 0.5, 6)
    norm[R.IN_SD:R.IN_SD + 6] = rng.uniform(0.5, 2.0, 6)
    norm[R.OUT_MU:R.OUT_MU + 3] = rng.normal(0.0, 0.1, 3)
    norm[R.OUT_SD:R.OUT_SD + 3] = rng.uniform(0.2, 1.0, 3)
    norm[R.Z_MAX] = 8.0
    return w, norm


def _soil(advance_f=True):
    return S.Mats("gpu").nacc(2e5, 0.3, 1.85, 0.5, 0.8, advance_f=advance_f)


def _soil_oracle(oracle, advance_f=True):
    return S.Mats("oracle32").nacc(2e5, 0.3, 1.85, 0.5, 0.8, advance_f=advance_f)


def _trial_states(seed, n, elastic_frac=0.3):
    """Trial F = R(a) diag(s1, s2) R(b) and hardening states: a compacted,
    mildly strained part inside the yield surface and a yielding rest."""
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(-np.pi, np.pi, (2, n))
    ne = int(n * elastic_frac)
    s = np.empty((n, 2))
    s[:ne] = 1.0 + rng.normal(0.0, 2e-4, (ne, 2))
    s[ne:] = rng.uniform(0.8, 1.15, (n - ne, 2))
    alpha = np.empty(n)
    alpha[:ne] = -1.0
    alpha[ne:] = rng.uniform(-0.6, 0.0, n - ne)
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    # F = Ra diag(s) Rb, column-major (F00, F10, F01, F11)
    f00 = ca * s[:, 0] * cb - sa * s[:, 1] * sb
    f10 = sa * s[:, 0] * cb + ca * s[:, 1] * sb
    f01 = -ca * s[:, 0] * sb - sa * s[:, 1] * cb
    f11 = -sa * s[:, 0] * sb + ca * s[:, 1] * cb
    f = np.stack([f00, f10, f01, f11], 1).astype(np.float32)
    return f, alpha.astype(np.float32)


def _run_batch(torch, mat, f, alpha, sg=None, reps=1):
    n = len(f)
    df = torch.from_numpy(f).cuda()
    da = torch.from_numpy(alpha).cuda()
    fo = torch.zeros_like(df)
    ao = torch.zeros_like(da)
    ms = eh.nacc_return_batch(mat, df.data_ptr(), da.data_ptr(), n, fo.data_ptr(), ao.data_ptr(), sg, reps)
    return fo.cpu().numpy(), ao.cpu().numpy(), ms


@pytest.mark.gpu
def test_infer_matches_reference():
    w, norm = _net(10)
    sg = eh.NaccSurrogate(0)
    sg.set_params(w, norm)
    rng = np.random.default_rng(11)
    x = (norm[R.IN_MU:R.IN_MU + 6] + norm[R.IN_SD:R.IN_SD + 6] * rng.normal(0.0, 1.0, (1000, 6))).astype(np.float32)
    y = sg.infer(x)
    ye = R.predict(w, norm, x, exact=False)
    yx = R.predict(w, norm, x, exact=True)
    scale = np.abs(yx - x[:, :3]).max()
    e_emul = np.abs(y - ye).max() / scale
    e_exact = np.abs(y - yx).max() / scale
    print(f"infer: max |gpu - emulated| / residual scale {e_emul:.3e}, vs exact {e_exact:.3e}")
    # bf16 operands: a few parts in a thousand of the residual against the exact network;
    # against the emulation only accumulation order and 1-ulp exp/rcp differences remain
    assert e_emul < 2e-3, e_emul
    assert e_exact < 3e-2, e_exact
    sg.close()


@pytest.mark.gpu
def test_gradient_matches_reference():
    """The device's forward/backward/weight-gradient kernels against the chain
    rule on the CPU; 300 samples exercise the padded final slice."""
    w, norm = _net(12)
    sg = eh.NaccSurrogate(0)
    sg.set_params(w, norm)
    rng = np.random.default_rng(13)
    x = (norm[R.IN_MU:R.IN_MU + 6] + norm[R.IN_SD:R.IN_SD + 6] * rng.normal(0.0, 1.0, (300, 6))).astype(np.float32)
    y = (x[:, :3] + rng.normal(0.0, 0.3, (300, 3))).astype(np.float32)
    g, loss = sg.gradient(x, y)
    ge, le = R.gradients(w, norm, x, y, exact=False)
    gx, lx = R.gradients(w, norm, x, y, exact=True)
    e_emul = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    e_exact = np.linalg.norm(g - gx) / np.linalg.norm(gx)
    print(f"gradient: rel err vs emulated {e_emul:.3e}, vs exact {e_exact:.3e}; loss {loss:.6f} / {le:.6f}")
    assert loss == pytest.approx(le, rel=1e-4)
    assert e_emul < 5e-3, e_emul
    assert e_exact < 5e-2, e_exact
    # per-layer, so a wrong layout in one layer cannot hide behind the others
    for lo, hi in ((R.W1, R.B1), (R.W2, R.B2), (R.W3, R.B3), (R.W4, R.B4), (R.B4, R.N_PARAMS)):
        assert np.linalg.norm(g[lo:hi] - ge[lo:hi]) <= 1e-2 * np.linalg.norm(ge[lo:hi]) + 1e-9, (lo, hi)
    sg.close()


@pytest.mark.gpu
def test_plain_batch_matches_oracle(oracle):
    torch = _torch()
    mat, om = _soil(), _soil_oracle(oracle)
    f, alpha = _trial_states(20, 3000)
    fo, ao, _ = _run_batch(torch, mat, f, alpha)
    ref_f = np.zeros_like(f)
    ref_a = np.zeros_like(alpha)
    for i in range(len(f)):
        ref_f[i], ref_a[i] = oracle.nacc_project(om, f[i], float(alpha[i]))
    scale = np.abs(ref_f).max(axis=1, keepdims=True)
    err = np.abs(fo - ref_f) / scale
    print(f"plain batch vs oracle: max rel {err.max():.3e}, alpha {np.abs(ao - ref_a).max():.3e}")
    assert err.max() < 2e-6
    assert np.abs(ao - ref_a).max() < 1e-5
    same = np.all(ref_f == f, axis=1)  # inside the yield surface: F unchanged
    assert same.sum() > 500 and np.array_equal(fo[same], f[same])


@pytest.mark.gpu
def test_gate_and_fallback():
    """z_max = 0 sends every yielding state to the plain return mapping: the
    surrogate kernel then equals the plain kernel bit for bit, and the counts
    add up."""
    torch = _torch()
    mat = _soil()
    f, alpha = _trial_states(21, 5000)
    fp, ap, _ = _run_batch(torch, mat, f, alpha)
    w, norm = _net(22)
    sg = eh.NaccSurrogate(0)
    norm[R.Z_MAX] = 0.0
    sg.set_params(w, norm)
    sg.counts(reset=True)
    fs, as_, _ = _run_batch(torch, mat, f, alpha, sg)
    c = sg.counts(reset=True)
    assert np.array_equal(fs, fp) and np.array_equal(as_, ap)
    yielding = int((~np.all(fp == f, axis=1)).sum())
    assert c["surrogate"] == 0 and c["fallback"] == c["yielding"] == yielding > 1000, c
    # with a wide range every Case C state goes through the network, the caps
    # (Cases A and B) keep the plain return bit for bit, elastic states keep F
    norm[R.Z_MAX] = 1e30
    sg.set_params(None, norm)
    fs, as_, _ = _run_batch(torch, mat, f, alpha, sg)
    c = sg.counts(reset=True)
    elastic = np.all(fp == f, axis=1)
    assert c["yielding"] == len(f) - elastic.sum(), c
    assert c["surrogate"] + c["fallback"] == c["yielding"] and c["surrogate"] > 100, c
    assert np.array_equal(fs[elastic], f[elastic])
    changed = ~np.all(fs == fp, axis=1)
    assert changed.sum() == c["surrogate"]  # a random network moves every state it returns
    sg.close()


def _harvest(sg, scene_fn, steps, cap):
    sim = scene_fn("gpu")
    sim.set_surrogate_mode(sg, "harvest", cap)
    sim.step_n(steps)
    n = len(sim.particles())
    subs = sim.stats().total_substeps
    sim.set_surrogate_mode(None, "plain")
    return sim, n, subs


@pytest.mark.gpu
def test_harvest_records_trial_states(oracle):
    """Every NACC particle is offered once per substep; each record's F and
    alpha project (plain, CPU oracle) to a finite state, and a record's
    material fields are the scene material's."""
    sg = eh.NaccSurrogate(1)
    sim, n, subs = _harvest(sg, S.nacc_block, 3, 1 << 20)
    offered, stored = sg.harvest_count()
    assert offered == stored == n * subs, (offered, n, subs)
    rec = sg.records()
    mat = _soil()
    assert np.all(rec[:, 5] == np.float32(mat.params.mu))
    assert np.all(rec[:, 7] == np.float32(mat.params.stretch_limit))
    assert np.all(rec[:, 8] == np.float32(mat.params.compression_limit))
    assert np.all(rec[:, 9] == np.float32(mat.params.hardening_exponent))
    assert np.all(np.isfinite(rec[:, :5]))
    # harvesting leaves the simulation as it is.  Two GPU runs differ only by
    # the order of their float atomics, which over the few hundred substeps
    # of three steps has moved a particle by up to ~1.2e-4 cell; a harvest
    # that wrote into the state would move them by far more
    plain = S.nacc_block("gpu")
    plain.step_n(3)
    dx = np.abs(plain.particles()["x"].astype(np.float64) - sim.particles()["x"]).max()
    assert dx < 1e-3, dx
    # the training set holds exactly the yielding records, labelled by the plain mapping
    n_set = sg.build_training_set()
    om = _soil_oracle(oracle)
    x, y = sg.training_set(0, min(n_set, 400))
    assert 0 < n_set <= stored
    for i in range(len(x)):
        assert np.isfinite(x[i]).all() and np.isfinite(y[i]).all()
    sim.close()
    plain.close()
    sg.close()


@pytest.mark.gpu
def test_training_reduces_loss_and_stress_error():
    """Adam on harvested states: the loss falls, and the trained network's
    Kirchhoff stress is far closer to the plain return mapping's than the
    trial state's (the method without a return mapping)."""
    torch = _torch()
    sg = eh.NaccSurrogate(2)
    sim, _, _ = _harvest(sg, S.nacc_block, 25, 1 << 21)
    n_set = sg.build_training_set()
    assert n_set > 2000, n_set
    loss = sg.train(epochs=60, batch=512, lr=5e-3, lr_final=5e-5, seed=3)
    print(f"training set {n_set}, loss first/last epoch {loss[0]:.4e} / {loss[-1]:.4e}")
    assert loss[-1] < 0.2 * loss[0], loss
    rec = sg.records()
    f = rec[:, :4].copy()
    alpha = rec[:, 4].copy()
    mat = _soil()
    fp, ap, _ = _run_batch(torch, mat, f, alpha)
    sg.counts(reset=True)
    fs, as_, _ = _run_batch(torch, mat, f, alpha, sg)
    c = sg.counts(reset=True)
    yl = ~np.all(fp == f, axis=1)
    mu, kappa = float(rec[0, 5]), float(rec[0, 6])  # the material's mu and kappa as the kernel read them
    tp = R.nacc_kirchhoff(fp[yl], mu, kappa)
    ts = R.nacc_kirchhoff(fs[yl], mu, kappa)
    tt = R.nacc_kirchhoff(f[yl], mu, kappa)
    e_s = np.linalg.norm(ts - tp) / np.linalg.norm(tp)
    e_t = np.linalg.norm(tt - tp) / np.linalg.norm(tp)
    print(f"yielding {yl.sum()}, counts {c}; stress rel err surrogate {e_s:.3e}, trial (no return) {e_t:.3e}")
    # the records are the training set's source: its Case C states go through the network
    assert c["surrogate"] >= 0.98 * n_set and c["surrogate"] + c["fallback"] == c["yielding"], (c, n_set)
    assert e_s < 0.1 * e_t
    sim.close()
    sg.close()


@pytest.mark.gpu
def test_substep_surrogate_mode():
    """The NACC scene stepped with the learned return mapping stays close to
    the plain one, and non-NACC scenes refuse the mode."""
    sg = eh.NaccSurrogate(4)
    sim, _, _ = _harvest(sg, S.nacc_block, 25, 1 << 21)
    sg.build_training_set()
    sg.train(epochs=60, batch=512, lr=5e-3, lr_final=5e-5, seed=5)
    plain = S.nacc_block("gpu")
    learned = S.nacc_block("gpu")
    learned.set_surrogate_mode(sg, "surrogate")
    sg.counts(reset=True)
    plain.step_n(10)
    learned.step_n(10)
    c = sg.counts(reset=True)
    xp, xl = plain.particles()["x"], learned.particles()["x"]
    dx = np.abs(xp.astype(np.float64) - xl).max()
    print(f"surrogate substeps: counts {c}, max |dx| after 10 steps {dx:.3e} cells")
    assert c["surrogate"] > 0
    assert dx < 0.5
    other = S.corotated_drop("gpu")
    other.set_surrogate_mode(sg, "surrogate")
    with pytest.raises(eh.EmergeError) as ei:
        other.step_n(1)
    assert ei.value.code == -2
    for s in (sim, plain, learned, other):
        s.close()
    sg.close()


@pytest.mark.gpu
def test_shared_harvest_buffer_cannot_grow_under_another_harvester():
    """Two scenes harvesting into one surrogate: growing the buffer while the other
    scene's device descriptor still points at it is refused (EMH_ERR_STATE); once
    that scene leaves harvest mode the buffer may grow, and both scenes harvest."""
    sg = eh.NaccSurrogate(2)
    a, b = S.nacc_block("gpu"), S.nacc_block("gpu")
    a.set_surrogate_mode(sg, "harvest", harvest_capacity=4096)
    b.set_surrogate_mode(sg, "harvest", harvest_capacity=4096)  # same size: no reallocation
    with pytest.raises(eh.EmergeError) as ei:
        b.set_surrogate_mode(sg, "harvest", harvest_capacity=1 << 16)
    assert ei.value.code == -5
    a.step_n(1)  # a still harvests into the buffer it was given
    offered, stored = sg.harvest_count()
    assert offered > 0 and stored > 0
    a.set_surrogate_mode(None, "plain")
    b.set_surrogate_mode(sg, "harvest", harvest_capacity=1 << 16)
    b.step_n(1)
    offered, stored = sg.harvest_count()
    assert offered == b.num_particles() * b.stats().last_substeps and stored == offered
    for s in (a, b):
        s.close()
    sg.close()


@pytest.mark.gpu
def test_device_adam_matches_reference_over_steps():
    """Four full-batch Adam steps on the device (k_sg_fwd_bwd + k_sg_wgrad + k_sg_adam,
    cosine learning rate) against the CPU reference: each step's gradient by the chain
    rule with the device's BF16 operand rounding emulated, then Adam with bias
    correction 1 / (1 - beta^t) at t = 1..4.  A planted mistake in the reference
    (bias correction left at t = 1) moves the weights far more than the device differs."""
    import surrogate_ref as R
    rng = np.random.default_rng(21)
    n = 256
    x = rng.normal(0.0, 1.0, (n, 6)).astype(np.float32)
    x[:, 3:] = x[:, 3:] * 0.1 + np.array([1.85, 0.5, 0.8], np.float32)
    y = (x[:, :3] + 0.05 * np.tanh(x[:, :3]) + rng.normal(0.0, 0.01, (n, 3))).astype(np.float32)
    sg = eh.NaccSurrogate(9)
    sg.set_training_set(x, y)
    w0, norm = sg.params()
    steps, lr, lr_final, b1, b2, eps = 4, 1e-2, 1e-3, 0.9, 0.999, 1e-8
    loss = sg.train(epochs=steps, batch=n, lr=lr, lr_final=lr_final, beta1=b1, beta2=b2, eps=eps, seed=5)
    assert len(loss) == steps
    w_dev, _ = sg.params()
    sg.close()

    def reference(bias_steps):
        w = w0.astype(np.float64)
        m = np.zeros_like(w)
        v = np.zeros_like(w)
        for t in range(1, steps + 1):
            g, _ = R.gradients(w.astype(np.float32), norm, x, y, exact=False)
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * (t - 1) / (steps - 1)))
            w, m, v = R.adam_step(w, m, v, g, bias_steps(t), lr_t, b1, b2, eps)
        return w

    w_ref = reference(lambda t: t)
    moved = np.linalg.norm(w_ref - w0)
    err = np.linalg.norm(w_dev - w_ref) / moved
    w_bad = reference(lambda t: 1)  # bias correction frozen at the first step
    err_bad = np.linalg.norm(w_bad - w_ref) / moved
    print(f"device vs reference after {steps} Adam steps: {err:.2e} of the update; planted mistake {err_bad:.2e}")
    assert err < 2e-2 and err_bad > 10 * err
