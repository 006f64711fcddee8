"""Device-resident access (GpuSimulation's buffer API) and the between-step
edits: particle / grid export into caller-owned device memory, set_gravity,
apply_impulse / apply_radial_impulse.  Device memory for the exports comes
from torch (plumbing only); every compute call goes through the C-ABI."""
import numpy as np
import pytest

import emerge_hip as eh
import scenes as S


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.gpu
def test_export_particles_matches_download():
    """The exported 128 B records equal the host download byte for byte, in
    caller order, also after reorders and with the cold fields (tags, flags)."""
    torch = _torch()
    sim = S.multi_material("gpu")
    p0 = sim.particles()
    p0["user_tag"] = np.arange(len(p0), dtype=np.uint32) * 7 + 3
    p0["temperature"] = np.linspace(-1.0, 1.0, len(p0), dtype=np.float32)
    sim.set_particles(p0)
    sim.set_reorder_interval(8)
    sim.step_n(12)  # the upload's reorder, then reorders every 8 substeps
    n = len(p0)
    buf = torch.zeros(n * 128 + 256, dtype=torch.uint8, device="cuda")
    sim.export_particles_device(buf.data_ptr(), n)
    got = buf.cpu().numpy()[: n * 128].view(eh.PARTICLE_DTYPE)
    ref = sim.particles()
    assert got.tobytes() == ref.tobytes()
    assert np.array_equal(got["user_tag"], p0["user_tag"])
    assert np.any(got["x"] != p0["x"])
    # the tail of the buffer is untouched
    assert not buf[n * 128:].any().item()
    with pytest.raises(eh.EmergeError) as ei:
        sim.export_particles_device(buf.data_ptr(), n - 1)
    assert ei.value.code == -1
    sim.close()


@pytest.mark.gpu
def test_export_grid_matches_get_grid():
    torch = _torch()
    sim = S.snow_drop("gpu")
    sim.step_n(3)
    r = sim.config.grid_res
    buf = torch.full((r * r, 4), 7.0, dtype=torch.float32, device="cuda")
    sim.export_grid_device(buf.data_ptr(), r * r)
    got = buf.cpu().numpy()
    mass, mom = sim.grid()
    assert np.array_equal(got[:, 2], mass)
    assert np.array_equal(got[:, :2], mom.reshape(-1, 2))
    assert np.all(got[:, 3] == 0.0) and mass.sum() > 0
    sim.close()


def _mirror(gpu_sim, backend_sim):
    """Gives the oracle the GPU's exact particle state."""
    backend_sim.set_particles(gpu_sim.particles())


@pytest.mark.gpu
@pytest.mark.parametrize("radial", [False, True])
def test_impulse_matches_oracle(radial, oracle):
    """solver/particles.rs:115-165 on identical state: the velocities after the
    impulse are bit-identical (one f32 expression per particle, sqrt correctly
    rounded on both sides); the steps after it take the oracle's substep
    schedule, so the loop's CFL bound saw the new velocities."""
    g = S.corotated_drop("gpu")
    o = S.corotated_drop("oracle32")
    g.step_n(2)
    _mirror(g, o)
    v_before = g.particles()["v"].copy()
    if radial:
        args = ((24.0, 8.0), 6.0, 40.0)
        g.apply_radial_impulse(*args)
        o.apply_radial_impulse(*args)
    else:
        args = ((22.0, 9.0), 5.0, (30.0, 55.0))
        g.apply_impulse(*args)
        o.apply_impulse(*args)
    pg, po = g.particles(), o.particles()
    assert np.array_equal(pg["v"], po["v"])
    moved = np.any(pg["v"] != v_before, axis=1)
    assert 10 < moved.sum() < len(pg)
    # the speed clamp of the impulse (grid_cell_size / min_dt = 1000) is not hit here
    t0g, t0o = g.stats().total_substeps, o.stats().total_substeps
    g.step_n(2)
    o.step_n(2)
    assert g.stats().total_substeps - t0g == o.stats().total_substeps - t0o
    e = S.rel_errors(S.state_matrix(g.particles()), S.state_matrix(o.particles()))
    assert e["x"] < 1e-5 and e["v"] < 1e-3 and e["F"] < 1e-4, e
    g.close()


@pytest.mark.gpu
def test_impulse_speed_clamp(oracle):
    """A huge impulse is clamped to grid_cell_size / min_dt (particles.rs:126-130)."""
    g = S.corotated_drop("gpu")
    o = S.corotated_drop("oracle32")
    g.step_n(1)
    _mirror(g, o)
    for sim in (g, o):
        sim.apply_impulse((24.0, 9.0), 4.0, (5.0e4, -2.0e4))
    pg, po = g.particles(), o.particles()
    assert np.array_equal(pg["v"], po["v"])
    speed = np.sqrt((pg["v"].astype(np.float64) ** 2).sum(1))
    assert speed.max() <= 1000.0 * (1 + 1e-6) and speed.max() > 900.0
    g.close()


@pytest.mark.gpu
def test_set_gravity(oracle):
    """set_gravity applies from the next step (lifecycle.rs:403-405)."""
    g = S.corotated_drop("gpu")
    o = S.corotated_drop("oracle32")
    g.step_n(1)
    _mirror(g, o)
    for sim in (g, o):
        sim.set_gravity((3.0, 4.0))
        sim.step_n(2)
    e = S.rel_errors(S.state_matrix(g.particles()), S.state_matrix(o.particles()))
    assert e["x"] < 1e-5 and e["v"] < 1e-3, e
    vg = g.particles()["v"].astype(np.float64).mean(0)
    assert vg[0] > 0.0  # the sideways gravity acted
    g.close()
