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


@pytest.mark.gpu
def test_sort_clustered_full_size(oracle):
    """Every one of 1,048,576 keys in one cell (particles clamped into a corner), and a
    half-and-half split: a single bin of n entries, stable by index."""
    xy = np.full((1 << 20, 2), 2.5, np.float32)
    keys, perm = eh.sort_cell_keys(xy, 512)
    okeys, operm = oracle.stable_sort_cell_keys(xy, 512)
    assert np.array_equal(keys, okeys) and np.array_equal(perm, operm)
    assert np.array_equal(perm, np.arange(1 << 20, dtype=np.uint32))
    xy[::2] = 509.5
    keys, perm = eh.sort_cell_keys(xy, 512)
    okeys, operm = oracle.stable_sort_cell_keys(xy, 512)
    assert np.array_equal(keys, okeys) and np.array_equal(perm, operm)


def _stable_order(keys_in_slot_order):
    return np.argsort(keys_in_slot_order, kind="stable").astype(np.uint32)


@pytest.mark.gpu
def test_pipeline_reorder_permutation_is_the_stable_sort(oracle):
    """The permutation the substep actually stores the particles in, on the 1M-particle
    sand column: after the upload's reorder (first step) the device slot order is
    std::stable_sort of the uploaded positions' cell keys, bit for bit, and two runs
    give the same permutation."""
    import bench
    orders = []
    for _ in range(2):
        sim = bench.build_workload("gpu")
        xy = sim.particles()["x"].copy()
        sim.step_n(1)
        orders.append(sim.slot_order())
        sim.close()
    _, want = oracle.stable_sort_cell_keys(xy, 512)
    assert np.array_equal(orders[0], want)
    assert np.array_equal(orders[1], orders[0])


# constitutive model id of each material slot of scenes.multi_material:
# NeoHookean (2), Newtonian water (1), Stomakhin snow (4), corotated (3)
MULTI_MATERIAL_MODELS = np.array([2, 1, 4, 3], np.uint32)


def _cell_bits(res):
    n_bins = ((res + 15) // 16) ** 2 * 256
    return int(np.ceil(np.log2(n_bins)))


@pytest.mark.gpu
@pytest.mark.parametrize("segments", [True, False], ids=["model_segments", "cell_key_only"])
def test_pipeline_mid_run_reorder_is_stable_over_the_stored_order(segments, oracle):
    """A reorder at a later step start sorts the stored slots by the key of their
    current positions, ties in slot order: new order = old order composed with the
    stable argsort of the keys read in old slot order (multi-material scene with a
    density pass, reordering at every step).  With per-material segments the key is
    model << cell bits | cell key, and the stored order comes out grouped by model."""
    import scenes as S
    sim = S.multi_material("gpu")
    sim.set_tuning(1, False, segments)
    res = sim.config.grid_res
    sim.step_n(3)
    for _ in range(3):
        old = sim.slot_order()
        p = sim.particles()
        keys = oracle.stable_sort_cell_keys(p["x"], res)[0].astype(np.uint64)  # cell keys in caller order
        if segments:
            keys |= MULTI_MATERIAL_MODELS[p["material_id"]].astype(np.uint64) << _cell_bits(res)
        sim.step_n(1)
        new = sim.slot_order()
        assert np.array_equal(new, old[_stable_order(keys[old])])
        if segments:
            models = MULTI_MATERIAL_MODELS[p["material_id"][new]]
            assert np.all(np.diff(models.astype(np.int64)) >= 0)  # one contiguous segment per model
    sim.close()
