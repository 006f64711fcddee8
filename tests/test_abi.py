"""C-ABI checks that need no GPU: the library loads, exports every function
include/emerge_hip.h declares, its host-side constructors agree byte for byte
with the oracle's restatement of the Rust constructors, and validation and
error reporting behave like the reference's asserts (as error codes)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import emerge_hip as eh

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "emerge_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(emh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = eh.load_library()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", eh.library_path()], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (emh_[a-z0-9_]+)", out))
    assert set(_declared()) <= exported


def test_abi_version_and_struct_sizes():
    lib = eh.load_library()
    assert lib.emh_abi_version() == 1
    assert ctypes.sizeof(eh.CConfig) == 116
    assert eh.PARTICLE_DTYPE.itemsize == 128          # particle.rs:144
    assert ctypes.sizeof(eh.CMaterialParams) == 96    # params.rs:101


def test_config_presets_match_reference_defaults():
    """SimConfig::default / standard / unsafe_defaults  config.rs:123-185."""
    c = eh.SimConfig.standard(64, 0.05, (0.0, -1.0))
    assert (c.grid_res, c.dt, c.gravity) == (64, np.float32(0.05), (0.0, -1.0))
    assert c.cfl_coefficient == np.float32(0.9) and c.min_dt == np.float32(1e-3)
    assert c.max_substeps_per_step == 64 and c.boundary_thickness == 2 and c.j_max == 50.0
    assert c.adaptive_timestep == 1 and c.project_invalid_state == 1 and c.apic_blend == 1.0
    d = eh.SimConfig.default()
    assert d.gravity == (0.0, np.float32(-0.05)) and d.dt == 1.0
    u = eh.SimConfig.unsafe_defaults()
    assert u.adaptive_timestep == 0 and u.project_invalid_state == 0
    c.validate()


@pytest.mark.parametrize("field,value", [
    ("grid_res", 3), ("dt", 0.0), ("min_dt", 2.0), ("j_max", 1.0), ("apic_blend", 1.5),
    ("boundary_thickness", 0), ("boundary_thickness", 63), ("max_substeps_per_step", 0),
])
def test_config_validate_rejects(field, value):
    """SimConfig::validate (config.rs:247-299) panics; the C-ABI returns EMH_ERR_INVALID_ARG."""
    c = eh.SimConfig.standard(64, 0.05, (0.0, -1.0)).replace(**{field: value})
    with pytest.raises(eh.EmergeError) as ei:
        c.validate()
    assert ei.value.code == -1


def _bytes(m):
    return bytes(m.params), bytes(m.ext)


def test_material_constructors_match_oracle(oracle):
    """The product's constructors and the oracle's restatement produce identical bytes."""
    pairs = [
        (eh.NeoHookeanMaterial.new(400.0, 200.0), oracle.material("neohookean", 400.0, 200.0)),
        (eh.CorotatedMaterial.new(400.0, 200.0), oracle.material("corotated", 400.0, 200.0)),
        (eh.NewtonianFluidMaterial.low_viscosity(1000.0, 1e4), oracle.material("fluid", 1000.0, 1e-3, 1e4, 7.0)),
        (eh.DruckerPragerMaterial.new(201923.08, 134615.38), oracle.material("drucker_prager", 201923.08, 134615.38)),
        (eh.StomakhinMaterial.new(1.0, 2.0, 10.0, 0.025, 0.0075, 0.6, 20.0),
         oracle.material("snow", 1.0, 2.0, 10.0, 0.025, 0.0075, 0.6, 20.0)),
        (eh.NaccMaterial.new(76923.08, 192307.7, 1.85, 0.5, 0.8), oracle.material("nacc", 76923.08, 192307.7, 1.85, 0.5, 0.8)),
        (eh.VonMisesMaterial.new(2000.0, 3000.0, 100.0), oracle.material("von_mises", 2000.0, 3000.0, 100.0, 0.0)),
        (eh.VonMisesMaterial.with_hardening(2000.0, 3000.0, 100.0, 500.0),
         oracle.material("von_mises", 2000.0, 3000.0, 100.0, 500.0)),
        (eh.ViscoelasticMaterial.new(1000.0, 800.0, 50.0), oracle.material("viscoelastic", 1000.0, 800.0, 50.0)),
        (eh.RankineMaterial.new(2000.0, 3000.0, 100.0, 1.0), oracle.material("rankine", 2000.0, 3000.0, 100.0, 1.0)),
        (eh.MuIRheologyMaterial.new(2000.0, 3000.0), oracle.material("mu_i_default", 2000.0, 3000.0)),
        (eh.MuIRheologyMaterial.with_params(2000.0, 3000.0, 0.4, 0.7, 1.12),
         oracle.material("mu_i", 2000.0, 3000.0, 0.4, 0.7, 1.12)),
        (eh.GranularFluidMaterial.new(27777.78, 41666.67, 1.0, 200.0, 5.0, 0.4),
         oracle.material("granular_fluid", 27777.78, 41666.67, 1.0, 200.0, 5.0, 0.4)),
        (eh.BinghamFluidMaterial.new(1000.0, 0.5, 5000.0, 7.0, 100.0),
         oracle.material("bingham", 1000.0, 0.5, 5000.0, 7.0, 100.0)),
    ]
    for a, b in pairs:
        assert _bytes(a) == _bytes(b), (a, b)
    assert eh.lame_from_young(3.5e5, 0.3) == oracle.lame_from_young(3.5e5, 0.3)
    assert eh.to_radians(30.0) == oracle.to_radians(30.0)
    sn = eh.StomakhinMaterial.from_young_modulus(1.4e5, 0.2)
    lam, mu = oracle.lame_from_young(1.4e5, 0.2)
    assert _bytes(sn) == _bytes(oracle.material("snow", lam, mu, 10.0, 0.025, 0.0075, 0.6, 20.0))


def test_material_params_field_use():
    """Field encoding per model params() (sand.rs:321-343, fluid.rs:209-226, nacc.rs:300-314)."""
    dp = eh.DruckerPragerMaterial.new(100.0, 200.0)
    assert dp.params.model == 5 and dp.params.volume_ratio_min == np.float32(0.6)
    assert dp.params.dp_h0 == eh.to_radians(35.0) and dp.params.dp_h2 == np.float32(0.2)
    fl = eh.NewtonianFluidMaterial.low_viscosity(1000.0, 1e4)
    assert fl.params.model == 1 and fl.params.eos_power == 7.0 and fl.params.pressure_floor == np.float32(-0.1)
    assert fl.params.volume_ratio_max == 2.0 and fl.params.dynamic_viscosity == np.float32(1e-3)
    na = eh.NaccMaterial.new(3000.0, 5000.0, 1.2, 0.1, 2.0)
    assert na.params.model == 10 and na.params.lambda_ == 2000.0 and na.params.stretch_limit == np.float32(1.2)
    assert na.ext.nacc_kappa == 5000.0 and na.ext.nacc_hardening_enabled == 1


def test_new_model_params_field_use():
    """von_mises.rs:143-152 (yield stress in hardening_exponent), viscoelastic.rs:149-159
    (j_min in volume_ratio_min), and the with_hardening assert (von_mises.rs:61)."""
    vm = eh.VonMisesMaterial.with_hardening(2000.0, 3000.0, 100.0, 500.0)
    assert vm.params.model == 6 and vm.params.hardening_exponent == 100.0 and vm.params.hardening_modulus == 500.0
    sd = eh.VonMisesMaterial.soft_ductile(5e4, 0.35)
    assert sd.params.hardening_exponent == np.float32(250.0)
    ve = eh.ViscoelasticMaterial.new(1000.0, 800.0, 50.0)
    assert ve.params.model == 9 and ve.params.dynamic_viscosity == 50.0
    assert ve.params.volume_ratio_min == np.float32(0.01) and ve.params.active_stress_coeff == 0.0
    with pytest.raises(eh.EmergeError) as ei:
        eh.VonMisesMaterial.with_hardening(2000.0, 3000.0, 100.0, -1.0)
    assert ei.value.code == -1


def test_cell_key_matches_oracle(oracle):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        r = int(rng.integers(4, 700))
        cx, cy = (int(v) for v in rng.integers(-5, r + 5, size=2))
        assert eh.cell_key(cx, cy, r) == oracle.cell_key(cx, cy, r)


@pytest.mark.parametrize("r", [16, 64, 100])
def test_cell_key_layout(r):
    """Properties that fix the key without restating it: a one-to-one map of
    the R x R cells into [0, side^2), side the least power of two >= R;
    cells with consecutive keys are edge neighbours (a continuous curve); an
    aligned run of 4^j keys fills an aligned 2^j x 2^j square; so every run of
    64 keys covers at most two neighbouring 8 x 8 squares (bounding box <= 128
    cells).  A tiled or row order fails the neighbour check, a row order
    inside quadrants the square check."""
    keys = np.array([[eh.cell_key(x, y, r) for y in range(r)] for x in range(r)])
    side = 2
    while side < r:
        side *= 2
    assert len(np.unique(keys)) == r * r and keys.min() >= 0 and keys.max() < side * side
    cell_of = {int(keys[x, y]): (x, y) for x in range(r) for y in range(r)}
    for k, (x, y) in cell_of.items():
        if k + 1 in cell_of:
            nx, ny = cell_of[k + 1]
            assert abs(nx - x) + abs(ny - y) == 1, (k, (x, y), (nx, ny))
    if r == side:
        assert keys.max() == r * r - 1 and cell_of[0] == (0, 0)
        for j in (1, 2, 3):
            n = 4 ** j
            for k0 in range(0, r * r, n):
                c = np.array([cell_of[k] for k in range(k0, k0 + n)])
                lo = c.min(axis=0)
                assert np.all(c.max(axis=0) - lo == 2 ** j - 1) and np.all(lo % 2 ** j == 0), (j, k0)
        ordered = np.array([cell_of[k] for k in range(r * r)])
        worst = 0
        for s0 in range(0, len(ordered) - 63):
            wnd = ordered[s0:s0 + 64]
            worst = max(worst, int(np.ptp(wnd[:, 0]) + 1) * int(np.ptp(wnd[:, 1]) + 1))
        assert worst <= 128, worst
    # clamped outside the grid
    assert eh.cell_key(-3, 5, r) == eh.cell_key(0, 5, r) and eh.cell_key(r + 9, r, r) == eh.cell_key(r - 1, r - 1, r)


def _no_gpu():
    return eh.device_count() == 0


@pytest.mark.skipif("not _no_gpu()")
def test_simulation_fails_loudly_without_gpu():
    """No CPU fallback: creating a simulation without a device is an error."""
    with pytest.raises(eh.EmergeError) as ei:
        eh.Simulation.empty(eh.SimConfig.standard(64, 0.05, (0.0, -1.0)))
    assert ei.value.code == -3
    with pytest.raises(eh.EmergeError):
        eh.sort_cell_keys(np.zeros((4, 2), np.float32), 64)


def test_error_message_is_reported():
    c = eh.SimConfig.standard(64, 0.05, (0.0, -1.0)).replace(dt=-1.0)
    with pytest.raises(eh.EmergeError, match="dt must be positive"):
        c.validate()
