"""aq_hash_sort_device_noalloc, the hash sort that allocates nothing (its bins live in
a static device array), against aq_hash_sort_device on the same inputs: two calls in a
row, the second smaller (fewer records, a smaller grid), so the second call runs over
the first call's leftovers in the shared array. Bit-exact, as integer and data-move
work must be. Called through the C ABI with device pointers."""
import ctypes as C

import numpy as np
import pytest


def run_both(sizes, cs, rec, cell_count):
    import torch
    from aquarius_amd._abi import load
    from aquarius_amd.spatial_hash import SpatialHash
    lib = load()
    n, stride = rec.shape
    cells = max(cell_count, sizes[0] * sizes[1] * sizes[2])
    d_rec = torch.from_numpy(rec).cuda()

    def outputs():
        return dict(perm=torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda"),
                    hash=torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda"),
                    start=torch.full((cells,), -7, dtype=torch.int32, device="cuda"),
                    end=torch.full((cells,), -7, dtype=torch.int32, device="cuda"),
                    sorted=torch.full_like(d_rec, -7.0) if n else torch.empty(1, device="cuda"))
    ref, got = outputs(), outputs()
    torch.cuda.synchronize()
    sh = SpatialHash(sizes, cs, capacity=max(n, 1))
    sh.sort_device(d_rec.data_ptr(), n, stride, ref["perm"].data_ptr(), ref["hash"].data_ptr(),
                   ref["start"].data_ptr(), ref["end"].data_ptr(), ref["sorted"].data_ptr(),
                   cell_count=cell_count)
    sh.close()
    print(f"noalloc call: sizes {sizes}, {n} records of stride {stride}, cell_count {cell_count}", flush=True)
    rc = lib.aq_hash_sort_device_noalloc(sizes[0], sizes[1], sizes[2], C.c_float(cs), d_rec.data_ptr(), n, stride,
                                         got["perm"].data_ptr(), got["hash"].data_ptr(), got["start"].data_ptr(),
                                         got["end"].data_ptr(), cell_count, got["sorted"].data_ptr(), None)
    assert rc == 0, lib.aq_last_error()
    for k in ref:
        a, b = ref[k].cpu().numpy(), got[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k


@pytest.mark.gpu
def test_noalloc_matches_device_entry_point_twice():
    rng = np.random.default_rng(17)
    # first call: near the limits (128^3 cells, 3M records of the reference's 48-byte layout)
    n1 = 3_000_000
    rec1 = rng.standard_normal((n1, 12)).astype(np.float32)
    rec1[:, :3] = rng.uniform(-0.1, 10.1, (n1, 3))
    run_both((128, 128, 128), 10.0 / 128, rec1, 0)
    # second call, smaller: fewer cells and records, a longer cell array (tail of -1)
    n2 = 50_001
    rec2 = rng.uniform(0, 3, (n2, 3)).astype(np.float32)
    run_both((12, 9, 7), 0.25, rec2, 12 * 9 * 7 + 100)


@pytest.mark.gpu
def test_noalloc_rejects_sizes_beyond_its_array():
    from aquarius_amd._abi import load
    lib = load()
    rc = lib.aq_hash_sort_device_noalloc(129, 128, 128, C.c_float(1.0), None, 0, 3, None, None, None, None,
                                         0, None, None)
    assert rc != 0 and b"2097152" in lib.aq_last_error()
    rc = lib.aq_hash_sort_device_noalloc(8, 8, 8, C.c_float(1.0), None, (1 << 22) + 1, 3, None, None, None,
                                         None, 0, None, None)
    assert rc != 0 and b"4194304" in lib.aq_last_error()
