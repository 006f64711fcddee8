"""emh_sort_cell_keys_noalloc (the allocation-free variant of the cell-key sort)
against emh_sort_cell_keys on the same inputs: keys and permutation bit-exact,
on two calls in a row (the second smaller, so stale scratch from the first
would show), with device buffers the caller owns (torch tensors)."""
import numpy as np
import pytest

import emerge_hip as eh


def _noalloc(xy, res):
    import torch
    n = len(xy)
    d_xy = torch.from_numpy(np.ascontiguousarray(xy, np.float32)).to("cuda")
    d_keys = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    d_perm = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    stream.synchronize()
    eh.sort_cell_keys_noalloc(d_xy.data_ptr(), n, res, d_keys.data_ptr(), d_perm.data_ptr(), stream.cuda_stream)
    stream.synchronize()
    return d_keys.cpu().numpy().view(np.uint32)[:n], d_perm.cpu().numpy().view(np.uint32)[:n]


@pytest.mark.gpu
def test_noalloc_sort_matches_allocating_sort():
    rng = np.random.default_rng(11)
    cases = [
        ("full_size_512", rng.uniform(-3.0, 515.0, (eh.SORT_NOALLOC_MAX_KEYS, 2)).astype(np.float32), 512),
        ("smaller_ragged_100", rng.uniform(0.0, 100.0, (70_001, 2)).astype(np.float32), 100),
        ("one_cell", np.full((5000, 2), 7.5, np.float32), 1024),
    ]
    for name, xy, res in cases:
        print(f"case {name}: n={len(xy)} grid_res={res}", flush=True)
        want_keys, want_perm = eh.sort_cell_keys(xy, res)
        keys, perm = _noalloc(xy, res)
        assert np.array_equal(keys, want_keys), name
        assert np.array_equal(perm, want_perm), name


@pytest.mark.gpu
def test_noalloc_sort_rejects_oversize_and_accepts_empty():
    import torch
    buf = torch.zeros(16, dtype=torch.float32, device="cuda")
    with pytest.raises(eh.EmergeError):
        eh.sort_cell_keys_noalloc(buf.data_ptr(), eh.SORT_NOALLOC_MAX_KEYS + 1, 512, buf.data_ptr(),
                                  buf.data_ptr())
    with pytest.raises(eh.EmergeError):
        eh.sort_cell_keys_noalloc(0, 10, 512, 0, 0)
    eh.sort_cell_keys_noalloc(0, 0, 512, 0, 0)
