"""The device cell-key sort against std::stable_sort on identical keys.

The reference has no CPU sort (its WGSL sort is a coarse 16x16-block counting
sort with a non-deterministic order inside a block,
src/systems/gpu/shaders/particle_sort.wgsl:94-245), so the contract is the
build's own: keys are the tile-major cell key of floor(x), the permutation is
stable by input index.  Both must be bit-exact.
"""
import numpy as np
import pytest

import emerge_hip as eh


def _cases():
    rng = np.random.default_rng(42)
    yield "empty", np.zeros((0, 2), np.float32), 64
    yield "one", np.array([[10.5, 3.25]], np.float32), 64
    yield "uniform_64", rng.uniform(0, 64, (5000, 2)).astype(np.float32), 64
    yield "uniform_512", rng.uniform(2, 510, (200_000, 2)).astype(np.float32), 512
    # every particle in one cell: one bin of 4096 (stability inside a big bin)
    yield "one_cell", np.full((4096, 2), 17.3, np.float32), 128
    # 16 particles per cell, lattice order (how spawned bodies arrive)
    g = np.stack(np.meshgrid(np.arange(40, 104, 0.25), np.arange(30, 94, 0.25), indexing="ij"), -1)
    yield "lattice_16ppc", g.reshape(-1, 2).astype(np.float32), 128
    # out-of-range and negative coordinates clamp into the grid
    yield "out_of_range", rng.uniform(-20, 90, (3000, 2)).astype(np.float32), 64
    # grid side not a multiple of the 16-cell tile
    yield "ragged_res", rng.uniform(0, 100, (20000, 2)).astype(np.float32), 100
    yield "reversed", rng.uniform(0, 256, (50000, 2)).astype(np.float32)[::-1].copy(), 256


CASES = list(_cases())


@pytest.mark.gpu
@pytest.mark.parametrize("name,xy,res", CASES, ids=[c[0] for c in CASES])
def test_sort_bit_exact(name, xy, res, oracle):
    keys, perm = eh.sort_cell_keys(xy, res)
    okeys, operm = oracle.stable_sort_cell_keys(xy, res)
    assert np.array_equal(keys, okeys), f"{name}: keys differ"
    assert np.array_equal(perm, operm), f"{name}: permutation differs"
    # permutation validity (tests/gpu.rs:1540-1583 checks the same for the WGSL sort)
    assert np.array_equal(np.sort(perm), np.arange(len(xy), dtype=np.uint32))
    assert np.all(np.diff(keys[perm].astype(np.int64)) >= 0)


@pytest.mark.gpu
def test_sort_full_size_properties(oracle):
    """1,048,576 particles at 512^2 (the bench workload's size): bit-exact too."""
    rng = np.random.default_rng(7)
    xy = rng.uniform(2, 510, (1 << 20, 2)).astype(np.float32)
    keys, perm = eh.sort_cell_keys(xy, 512)
    okeys, operm = oracle.stable_sort_cell_keys(xy, 512)
    assert np.array_equal(keys, okeys)
    assert np.array_equal(perm, operm)
