"""The spatial hashing service (include/aqflip_hash.h): its numpy oracle pinned by the
report's worked example and by the FLIP oracle's hash + sort (CPU), and the GPU
handle against the oracle, bit-exact, on the edge cases the reference's callers
produce (GPU)."""
import json
from pathlib import Path

import numpy as np
import pytest

from aquarius_amd import configs
from oracle import binding as ob
from oracle import spatial_hash_ref as R

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------- CPU: the oracle
def test_oracle_matches_report_example():
    kat = json.loads((GOLDEN / "report_spatial_index.json").read_text())
    pos = np.array([[x, y, 0.5] for x, y in kat["positions_xy"]], np.float32)
    out = R.spatial_hash(pos, (4, 4, 1), 1.0)
    unsorted = np.empty_like(out["sorted_hash"])
    unsorted[out["perm"]] = out["sorted_hash"]
    assert unsorted.tolist() == kat["hashes"]
    for c in range(16):
        want = kat["cell_ranges"].get(str(c), [-1, -1])
        assert [out["cell_start"][c], out["cell_end"][c]] == want


@pytest.mark.parametrize("scene", ["ball_drop_16", "ball_drop_64"])
def test_oracle_matches_flip_oracle_hash_sort(scene):
    cfg = configs.SCENES[scene]()
    pos, frac = ob.seed(cfg)
    o = ob.Oracle(cfg, pos=pos, frac4=frac, init=False)
    o.substage("hash_sort")
    start, end, hsh, perm = o.get_cells()
    X, Y, Z = cfg.size
    h = cfg.grid_physical_size / Y
    out = R.spatial_hash(pos, (X, Y, Z), np.float32(h))
    assert np.array_equal(out["perm"], perm)
    assert np.array_equal(out["sorted_hash"], hsh)
    assert np.array_equal(out["cell_start"], start[:X * Y * Z])
    assert np.array_equal(out["cell_end"], end[:X * Y * Z])


def test_oracle_clamps_and_handles_empty():
    out = R.spatial_hash(np.zeros((0, 3), np.float32), (2, 3, 4), 0.5)
    assert out["perm"].size == 0 and (out["cell_start"] == -1).all()
    pos = np.array([[-5, 0.1, 0.1], [100, 100, 100], [0.99, 1.4, 1.9]], np.float32)
    out = R.spatial_hash(pos, (2, 3, 4), 0.5)
    unsorted = np.empty_like(out["sorted_hash"])
    unsorted[out["perm"]] = out["sorted_hash"]
    assert unsorted.tolist() == [0, 1 * 12 + 2 * 4 + 3, 1 * 12 + 2 * 4 + 3]


# ---------------------------------------------------------------- GPU: the handle
def gpu_case(sizes, cs, rec):
    from aquarius_amd.spatial_hash import SpatialHash
    sh = SpatialHash(sizes, cs, capacity=max(1, len(rec) // 2))
    got = sh.sort(rec)
    sh.close()
    want = R.spatial_hash(rec, sizes, cs)
    for k in ("perm", "sorted_hash", "cell_start", "cell_end"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["sorted_records"].view(np.uint32), want["sorted_records"].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("stride", [3, 12])
def test_gpu_random_records(stride):
    rng = np.random.default_rng(stride)
    n = 200_003
    rec = rng.standard_normal((n, stride)).astype(np.float32)
    rec[:, :3] = rng.uniform(-0.5, 10.5, (n, 3))  # some outside the 20x16x12 grid of 0.5 cells
    rec[:, :3] = np.round(rec[:, :3] * 4) / 4 * (rng.random((n, 1)) < 0.3) + rec[:, :3] * (rng.random((n, 1)) >= 0.3)
    gpu_case((20, 16, 12), 0.5, rec)


@pytest.mark.gpu
def test_gpu_segments_of_17_to_32_records():
    # about 24 records per cell in random order (Poisson counts, mostly 17-32): each cell
    # receives runs from many waves, so its segment is sorted by the 32-entry network
    rng = np.random.default_rng(24)
    pos = rng.uniform(0, 16, (16 ** 3 * 24, 3)).astype(np.float32)
    counts = np.bincount((pos.astype(np.int64) * [256, 16, 1]).sum(1), minlength=16 ** 3)
    assert ((counts > 16) & (counts <= 32)).sum() > 3000
    gpu_case((16, 16, 16), 1.0, pos)


@pytest.mark.gpu
def test_gpu_edge_cases():
    gpu_case((3, 4, 5), 1.0, np.zeros((0, 3), np.float32))                      # empty
    gpu_case((3, 4, 5), 1.0, np.array([[2.5, 0.5, 4.9]], np.float32))            # one particle
    gpu_case((7, 1, 1), 0.25, np.full((5000, 4), 0.3, np.float32))               # one crowded cell
    pos = np.random.default_rng(1).uniform(0, 64, (300_000, 3)).astype(np.float32)
    gpu_case((64, 64, 64), 1.0, pos)                                              # dense grid
    from aquarius_amd.spatial_hash import SpatialHash
    sh = SpatialHash((4, 4, 4), 1.0)
    with pytest.raises(Exception, match="non-finite"):
        sh.sort(np.array([[np.nan, 0, 0]], np.float32))
    sh.close()


@pytest.mark.gpu
def test_gpu_device_entry_point():
    import torch
    from aquarius_amd.spatial_hash import SpatialHash
    rng = np.random.default_rng(3)
    n, stride, sizes = 100_000, 12, (32, 24, 16)
    rec = rng.uniform(0, 8, (n, stride)).astype(np.float32)
    want = R.spatial_hash(rec, sizes, 0.25)
    d_rec = torch.from_numpy(rec).cuda()
    d_perm = torch.empty(n, dtype=torch.int32, device="cuda")
    d_hash = torch.empty(n, dtype=torch.int32, device="cuda")
    cells = sizes[0] * sizes[1] * sizes[2]
    d_start = torch.empty(cells, dtype=torch.int32, device="cuda")
    d_end = torch.empty(cells, dtype=torch.int32, device="cuda")
    d_sorted = torch.empty_like(d_rec)
    torch.cuda.synchronize()
    sh = SpatialHash(sizes, 0.25)
    sh.sort_device(d_rec.data_ptr(), n, stride, d_perm.data_ptr(), d_hash.data_ptr(), d_start.data_ptr(),
                   d_end.data_ptr(), d_sorted.data_ptr())
    sh.close()
    assert np.array_equal(d_perm.cpu().numpy(), want["perm"])
    assert np.array_equal(d_hash.cpu().numpy(), want["sorted_hash"])
    assert np.array_equal(d_start.cpu().numpy(), want["cell_start"])
    assert np.array_equal(d_end.cpu().numpy(), want["cell_end"])
    assert np.array_equal(d_sorted.cpu().numpy(), want["sorted_records"])


@pytest.mark.gpu
def test_gpu_device_entry_point_follows_the_callers_stream():
    """The records are written by kernels still queued on the caller's stream (behind
    a long chain of matrix products) when the sort is called, with no synchronisation
    in between: the sort runs after them because it is queued on that stream. The
    cell arrays are longer than the grid (cell_count): the tail is set to -1."""
    import torch
    from aquarius_amd.spatial_hash import SpatialHash
    rng = np.random.default_rng(4)
    n, stride, sizes = 300_000, 4, (40, 30, 20)
    rec = rng.uniform(0, 10, (n, stride)).astype(np.float32)
    want = R.spatial_hash(rec, sizes, 0.25)
    cells = sizes[0] * sizes[1] * sizes[2]
    extra = 777
    side = torch.cuda.Stream()  # a non-default stream of the caller's (default priority)
    with torch.cuda.stream(side):
        src = torch.from_numpy(rec).cuda()
        d_rec = torch.full_like(src, -1.0)
        d_start = torch.full((cells + extra,), 5, dtype=torch.int32, device="cuda")
        d_end = torch.full((cells + extra,), 5, dtype=torch.int32, device="cuda")
        d_sorted = torch.empty_like(src)
        torch.cuda.synchronize()
        a = torch.randn(2048, 2048, device="cuda") / 64
        for _ in range(40):  # keep the stream busy for a few milliseconds
            a = torch.tanh(a @ a)
        d_rec.copy_(src + 0 * a[0, 0])  # written only after the products
        sh = SpatialHash(sizes, 0.25)
        sh.sort_device(d_rec.data_ptr(), n, stride, 0, 0, d_start.data_ptr(), d_end.data_ptr(),
                       d_sorted.data_ptr(), stream=side.cuda_stream, cell_count=cells + extra)
        got_start = d_start.cpu().numpy()
        got_end = d_end.cpu().numpy()
        got_sorted = d_sorted.cpu().numpy()
        sh.close()
    assert np.array_equal(got_start[:cells], want["cell_start"])
    assert np.array_equal(got_end[:cells], want["cell_end"])
    assert np.all(got_start[cells:] == -1) and np.all(got_end[cells:] == -1)
    assert np.array_equal(got_sorted, want["sorted_records"])


@pytest.mark.gpu
def test_cell_count_below_the_grid_is_rejected():
    from aquarius_amd.spatial_hash import SpatialHash
    sh = SpatialHash((4, 4, 4), 1.0)
    with pytest.raises(Exception, match="cell_count"):
        sh.sort(np.zeros((1, 3), np.float32), cell_count=10)
    sh.close()


@pytest.mark.gpu
def test_gpu_grid_with_more_than_1024_scan_blocks():
    """256 x 256 x 80 cells: 1,280 scan blocks of 4,096 cells, above the 1,024 for which
    each final block sums the totals before it, so the block totals are scanned by
    k_scan_parts first."""
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 1, (1_500_000, 3)).astype(np.float32) * np.array([256, 256, 80], np.float32) * 0.5
    gpu_case((256, 256, 80), 0.5, pos)
