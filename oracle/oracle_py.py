"""ctypes wrapper of the CPU oracle (liboracle_mpm.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker -- never by the product library
or its Python package.  The oracle is a C++ restatement of the reference
crate's CPU Simulation::step (see mpm_oracle.hpp); precision=32 reproduces
the reference's f32 arithmetic, precision=64 is the FP64 oracle.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys
from typing import Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_REPO, "emerge-engine_amd"))

from emerge_hip import (CConfig, CMaterialExt, CMaterialParams, CSpawnRegion, CStepStats,  # noqa: E402
                        PARTICLE_DTYPE, Material, SimConfig, SpawnRegion, StepStats)

# Per-particle f64 export layout of orc_get_particles_f64.
F64_FIELDS = ("x0", "x1", "v0", "v1", "c00", "c10", "c01", "c11", "f00", "f10", "f01", "f11", "mass",
              "initial_volume", "volume", "density", "plastic_volume_ratio", "hardening_scale",
              "friction_hardening", "log_volume_strain")

_lib = None


def lib_path() -> str:
    return os.path.join(_HERE, "liboracle_mpm.so")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path()):
        build()
    lib = ctypes.CDLL(lib_path())
    P = ctypes.POINTER
    f = ctypes.c_float
    sig = {
        "orc_create": (ctypes.c_void_p, [ctypes.c_int, P(CConfig)]),
        "orc_create_from_spawn": (ctypes.c_void_p, [ctypes.c_int, P(CConfig), P(CSpawnRegion)]),
        "orc_destroy": (None, [ctypes.c_void_p]),
        "orc_set_material": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, P(CMaterialParams), P(CMaterialExt)]),
        "orc_reinit_particle_state": (ctypes.c_int, [ctypes.c_void_p]),
        "orc_add_slip_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "orc_clear_boundaries": (ctypes.c_int, [ctypes.c_void_p]),
        "orc_apply_impulse": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_float] * 6 + [ctypes.c_uint32]),
        "orc_sleep_tag": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
        "orc_active_count": (ctypes.c_size_t, [ctypes.c_void_p]),
        "orc_set_gravity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float]),
        "orc_add_friction_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float]),
        "orc_add_heightmap_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                      ctypes.c_float, ctypes.c_uint32]),
        "orc_add_predictive_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float]),
        "orc_boundary_velocity": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         P(ctypes.c_float)]),
        "orc_boundary_clamp": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p,
                                      ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_float)]),
        "orc_add_body": (ctypes.c_int, [ctypes.c_void_p, P(CSpawnRegion), P(ctypes.c_uint32)]),
        "orc_set_particles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "orc_num_particles": (ctypes.c_size_t, [ctypes.c_void_p]),
        "orc_get_particles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "orc_get_particles_f64": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "orc_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "orc_substeps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, f]),
        "orc_stats": (ctypes.c_int, [ctypes.c_void_p, P(CStepStats)]),
        "orc_substep_dts": (ctypes.c_size_t, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "orc_get_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "orc_p2g_only": (ctypes.c_int, [ctypes.c_void_p, f, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "orc_choose_substep_dt": (ctypes.c_double, [ctypes.c_void_p, f]),
        "orc_project_particle": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t]),
        "orc_set_mutation": (None, [ctypes.c_int]),
        "orc_set_threads": (None, [ctypes.c_int]),
        "orc_time_steps": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_uint32]),
        "orc_time_substeps": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_uint32, f]),
        "orc_quadratic_weights": (None, [f, f, P(ctypes.c_int32), P(f), P(f)]),
        "orc_svd2": (None, [P(f), P(f), P(f), P(f)]),
        "orc_svd2_f64": (None, [P(ctypes.c_double)] * 4),
        "orc_polar": (None, [P(f), P(f)]),
        "orc_kirchhoff": (None, [P(CMaterialParams), P(CMaterialExt), ctypes.c_void_p, P(f)]),
        "orc_update_particle": (None, [P(CMaterialParams), P(CMaterialExt), ctypes.c_void_p, f, ctypes.c_int]),
        "orc_timestep_bound": (f, [P(CMaterialParams), P(CMaterialExt), f, f, f, f, f]),
        "orc_dp_alpha": (f, [P(CMaterialParams), f]),
        "orc_nacc_project": (None, [P(CMaterialParams), P(CMaterialExt), P(f), f, P(f), P(f)]),
        "orc_nacc_project_batch": (None, [P(CMaterialParams), P(CMaterialExt), ctypes.c_int, ctypes.c_size_t,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        "orc_slip_velocity": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, P(f)]),
        "orc_lame_from_young": (None, [f, f, P(f), P(f)]),
        "orc_to_radians": (f, [f]),
        "orc_material_neohookean": (None, [f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_corotated": (None, [f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_von_mises": (None, [f, f, f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_viscoelastic": (None, [f, f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_rankine": (None, [f, f, f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_mu_i": (None, [f, f, f, f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_mu_i_default": (None, [f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_granular_fluid": (None, [f, f, f, f, f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_bingham": (None, [f, f, f, f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_fluid": (None, [f, f, f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_drucker_prager": (None, [f, f, P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_snow": (None, [f] * 7 + [P(CMaterialParams), P(CMaterialExt)]),
        "orc_material_nacc": (None, [f] * 5 + [P(CMaterialParams), P(CMaterialExt)]),
        "orc_cell_key": (ctypes.c_uint32, [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32]),
        "orc_stable_sort_cell_keys": (None, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p,
                                             ctypes.c_void_p]),
        "orc_lcg": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _f4(vals):
    return (ctypes.c_float * 4)(*vals)


class OracleSim:
    """The oracle's Simulation (same scene-building calls as emerge_hip.Simulation)."""

    def __init__(self, handle, config: SimConfig, precision: int):
        if not handle:
            raise ValueError("oracle rejected the configuration")
        self._h = handle
        self.config = config
        self.precision = precision

    @classmethod
    def empty(cls, config: SimConfig, precision: int = 32) -> "OracleSim":
        return cls(load().orc_create(precision, ctypes.byref(config.c)), config.copy(), precision)

    @classmethod
    def new(cls, config: SimConfig, spawn: SpawnRegion, precision: int = 32) -> "OracleSim":
        sc = spawn.to_c()
        return cls(load().orc_create_from_spawn(precision, ctypes.byref(config.c), ctypes.byref(sc)),
                   config.copy(), precision)

    def close(self):
        if self._h:
            load().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle call failed: {rc}")

    def set_material(self, mid: int, m: Material):
        self._ok(load().orc_set_material(self._h, mid, ctypes.byref(m.params), ctypes.byref(m.ext)))

    def with_default_material(self, m: Material):
        self.set_material(0, m)
        return self

    def with_material(self, mid: int, m: Material):
        self.set_material(mid, m)
        return self

    def with_boundary(self, b):
        lib = load()
        kind = type(b).__name__
        if kind == "SlipBoundary":
            self._ok(lib.orc_add_slip_boundary(self._h, b.thickness))
        elif kind == "FrictionBoundary":
            self._ok(lib.orc_add_friction_boundary(self._h, b.thickness, b.friction_coefficient))
        elif kind == "HeightmapBoundary":
            h = np.ascontiguousarray(b.heights, np.float32)
            self._ok(lib.orc_add_heightmap_boundary(self._h, h.ctypes.data, len(h), b.friction, b.wall_thickness))
        elif kind == "PredictiveBoundary":
            self._ok(lib.orc_add_predictive_boundary(self._h, b.thickness, b.predictive_wall_min))
        else:
            raise ValueError(kind)
        return self

    def clear_boundaries(self):
        self._ok(load().orc_clear_boundaries(self._h))

    def set_gravity(self, g):
        self._ok(load().orc_set_gravity(self._h, float(g[0]), float(g[1])))

    def sleep_tag(self, tag):
        self._ok(load().orc_sleep_tag(self._h, int(tag), 1))

    def wake_tag(self, tag):
        self._ok(load().orc_sleep_tag(self._h, int(tag), 0))

    def active_count(self) -> int:
        return int(load().orc_active_count(self._h))

    def apply_impulse(self, center, radius, force):
        self._ok(load().orc_apply_impulse(self._h, float(center[0]), float(center[1]), float(radius),
                                          float(force[0]), float(force[1]), 0.0, 0))

    def apply_radial_impulse(self, center, radius, strength):
        self._ok(load().orc_apply_impulse(self._h, float(center[0]), float(center[1]), float(radius), 0.0, 0.0,
                                          float(strength), 1))

    def reinit_all_particle_state(self):
        self._ok(load().orc_reinit_particle_state(self._h))

    def add_body(self, spawn: SpawnRegion) -> int:
        tag = ctypes.c_uint32()
        sc = spawn.to_c()
        self._ok(load().orc_add_body(self._h, ctypes.byref(sc), ctypes.byref(tag)))
        return tag.value

    def set_particles(self, particles: np.ndarray):
        arr = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        self._ok(load().orc_set_particles(self._h, arr.ctypes.data, arr.shape[0]))

    def num_particles(self) -> int:
        return int(load().orc_num_particles(self._h))

    def particles(self) -> np.ndarray:
        n = self.num_particles()
        out = np.zeros(n, dtype=PARTICLE_DTYPE)
        self._ok(load().orc_get_particles(self._h, out.ctypes.data, n))
        return out

    def particles_f64(self) -> np.ndarray:
        n = self.num_particles()
        out = np.zeros((n, 20), dtype=np.float64)
        self._ok(load().orc_get_particles_f64(self._h, out.ctypes.data, n))
        return out

    def step(self):
        self._ok(load().orc_step(self._h, 1))

    def step_n(self, n: int):
        self._ok(load().orc_step(self._h, n))

    def substeps(self, n: int, dt: float):
        self._ok(load().orc_substeps(self._h, n, dt))

    def choose_substep_dt(self, max_dt: float) -> float:
        """choose_substep_dt (step.rs:434-479) over the current state."""
        return float(load().orc_choose_substep_dt(self._h, max_dt))

    def project_particle(self, i: int) -> bool:
        """project_particle_state_to_admissible (step.rs:347-429) of particle i, in place."""
        r = int(load().orc_project_particle(self._h, i))
        if r < 0:
            raise IndexError(i)
        return r == 1

    def time_steps(self, n: int) -> float:
        return float(load().orc_time_steps(self._h, n))

    def time_substeps(self, n: int, dt: float) -> float:
        return float(load().orc_time_substeps(self._h, n, dt))

    def stats(self) -> StepStats:
        s = CStepStats()
        self._ok(load().orc_stats(self._h, ctypes.byref(s)))
        return StepStats(s.last_substeps, s.last_step_dt, s.last_sim_time_dropped, s.last_vel_clamp_count,
                         s.last_j_projection_count, s.total_substeps, s.frame_index)

    def substep_dts(self) -> np.ndarray:
        n = int(load().orc_substep_dts(self._h, None, 0))
        out = np.zeros(n, dtype=np.float64)
        load().orc_substep_dts(self._h, out.ctypes.data, n)
        return out

    def grid(self) -> Tuple[np.ndarray, np.ndarray]:
        r = self.config.grid_res
        m = np.zeros(r * r)
        mom = np.zeros((r * r, 2))
        self._ok(load().orc_get_grid(self._h, m.ctypes.data, mom.ctypes.data, r * r))
        return m, mom

    def p2g_grid(self, dt: float) -> Tuple[np.ndarray, np.ndarray]:
        r = self.config.grid_res
        m = np.zeros(r * r)
        mom = np.zeros((r * r, 2))
        self._ok(load().orc_p2g_only(self._h, dt, m.ctypes.data, mom.ctypes.data, r * r))
        return m, mom


# ------------------------------------------------------------ unit access
def quadratic_weights(x: float, y: float):
    b = (ctypes.c_int32 * 2)()
    wx = (ctypes.c_float * 3)()
    wy = (ctypes.c_float * 3)()
    load().orc_quadratic_weights(x, y, b, wx, wy)
    return (b[0], b[1]), np.array(wx[:], np.float32), np.array(wy[:], np.float32)


def svd2(f, precision: int = 32):
    if precision == 64:
        F = (ctypes.c_double * 4)(*f)
        u, s, vt = (ctypes.c_double * 4)(), (ctypes.c_double * 2)(), (ctypes.c_double * 4)()
        load().orc_svd2_f64(F, u, s, vt)
        return np.array(u[:]), np.array(s[:]), np.array(vt[:])
    F = _f4(f)
    u, s, vt = (ctypes.c_float * 4)(), (ctypes.c_float * 2)(), (ctypes.c_float * 4)()
    load().orc_svd2(F, u, s, vt)
    return np.array(u[:], np.float32), np.array(s[:], np.float32), np.array(vt[:], np.float32)


def polar(f):
    r = (ctypes.c_float * 4)()
    load().orc_polar(_f4(f), r)
    return np.array(r[:], np.float32)


def kirchhoff(m: Material, particle: np.ndarray) -> np.ndarray:
    arr = np.ascontiguousarray(particle, dtype=PARTICLE_DTYPE).reshape(1)
    tau = (ctypes.c_float * 4)()
    load().orc_kirchhoff(ctypes.byref(m.params), ctypes.byref(m.ext), arr.ctypes.data, tau)
    return np.array(tau[:], np.float32)


def update_particle(m: Material, particle: np.ndarray, dt: float, precision: int = 32) -> np.ndarray:
    arr = np.array(particle, dtype=PARTICLE_DTYPE).reshape(1)
    load().orc_update_particle(ctypes.byref(m.params), ctypes.byref(m.ext), arr.ctypes.data, dt, precision)
    return arr[0]


def timestep_bound(m: Material, density, h, cell=1.0, mcfl=0.5, vcfl=0.5) -> float:
    return float(load().orc_timestep_bound(ctypes.byref(m.params), ctypes.byref(m.ext), density, h, cell, mcfl, vcfl))


def dp_alpha(m: Material, q: float) -> float:
    return float(load().orc_dp_alpha(ctypes.byref(m.params), q))


def nacc_project(m: Material, f, alpha: float):
    fo = (ctypes.c_float * 4)()
    ao = ctypes.c_float()
    load().orc_nacc_project(ctypes.byref(m.params), ctypes.byref(m.ext), _f4(f), alpha, fo, ctypes.byref(ao))
    return np.array(fo[:], np.float32), ao.value


def nacc_project_batch(m: Material, f: np.ndarray, alpha: np.ndarray, precision: int = 64):
    """NaccMaterial::project over trial states f [n, 4] (column-major), alpha [n]; float64 results."""
    f = np.ascontiguousarray(f, dtype=np.float32).reshape(-1, 4)
    alpha = np.ascontiguousarray(alpha, dtype=np.float32).reshape(-1)
    fo = np.zeros((len(f), 4))
    ao = np.zeros(len(f))
    load().orc_nacc_project_batch(ctypes.byref(m.params), ctypes.byref(m.ext), precision, len(f), f.ctypes.data,
                                  alpha.ctypes.data, fo.ctypes.data, ao.ctypes.data)
    return fo, ao


BOUNDARY_KINDS = {"slip": 0, "friction": 1, "heightmap": 2, "predictive": 3}


def boundary_velocity(kind: str, cell_index: int, grid_res: int, v, thickness=2, mu=0.0, wall_min=0.0,
                      heights=None):
    vv = (ctypes.c_float * 2)(*v)
    h = None if heights is None else np.ascontiguousarray(heights, np.float32)
    load().orc_boundary_velocity(BOUNDARY_KINDS[kind], thickness, mu, wall_min,
                                 None if h is None else h.ctypes.data, 0 if h is None else len(h),
                                 cell_index, grid_res, vv)
    return (vv[0], vv[1])


def boundary_clamp(kind: str, pos, grid_res: int, thickness=2, wall_min=0.0, heights=None):
    pp = (ctypes.c_float * 2)(*pos)
    h = None if heights is None else np.ascontiguousarray(heights, np.float32)
    load().orc_boundary_clamp(BOUNDARY_KINDS[kind], thickness, wall_min,
                              None if h is None else h.ctypes.data, 0 if h is None else len(h), grid_res, pp)
    return (pp[0], pp[1])


def slip_velocity(thickness: int, cell_index: int, grid_res: int, v):
    vv = (ctypes.c_float * 2)(*v)
    load().orc_slip_velocity(thickness, cell_index, grid_res, vv)
    return (vv[0], vv[1])


def lame_from_young(e: float, nu: float):
    lam, mu = ctypes.c_float(), ctypes.c_float()
    load().orc_lame_from_young(e, nu, ctypes.byref(lam), ctypes.byref(mu))
    return lam.value, mu.value


def to_radians(d: float) -> float:
    return float(load().orc_to_radians(d))


def material(kind: str, *args) -> Material:
    p, e = CMaterialParams(), CMaterialExt()
    getattr(load(), f"orc_material_{kind}")(*args, ctypes.byref(p), ctypes.byref(e))
    return Material(p, e)


def set_threads(n: int) -> None:
    """Threads of G2P phase 1 (the reference's rayon loop, transfer/g2p.rs:189-196)."""
    load().orc_set_threads(int(n))


# Planted mistakes of the oracle (mpm_oracle.hpp Mutation), for its own tests.
MUT_NONE, MUT_CFL_NO_AFFINE, MUT_CFL_NO_MIN_CLAMP, MUT_PROJ_JSCALE_LINEAR = 0, 1, 2, 3
MUT_PROJ_NO_DENSITY_FLOOR, MUT_NACC_OTHER_ROOT, MUT_NACC_CASE_A_NO_ALPHA, MUT_NACC_HARDEN_SIGN = 4, 5, 6, 7


class mutation:
    """Context manager: the oracle runs with planted mistake `m` inside the block."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        load().orc_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        load().orc_set_mutation(MUT_NONE)
        return False


def cell_key(cx: int, cy: int, grid_res: int) -> int:
    return int(load().orc_cell_key(cx, cy, grid_res))


def stable_sort_cell_keys(xy: np.ndarray, grid_res: int):
    xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
    n = xy.shape[0]
    keys = np.zeros(n, np.uint32)
    perm = np.zeros(n, np.uint32)
    load().orc_stable_sort_cell_keys(xy.ctypes.data, n, grid_res, keys.ctypes.data, perm.ctypes.data)
    return keys, perm


def lcg(seed: int, n: int):
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    load().orc_lcg(seed, n, u.ctypes.data, f.ctypes.data)
    return u, f
