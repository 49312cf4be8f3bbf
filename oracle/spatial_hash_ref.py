"""CPU oracle of the spatial hashing service -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of performSpatialHashing2 (Fluid/Fluid_3D_common.cuh:28-56) as
include/aqflip_hash.h specifies it:
  * calcHashImpl (Fluid/Fluid_3D_kernels.cuh:8-27): x = (int)(px / cell_size) with
    IEEE float32 division and truncation, likewise y, z; hash = x*(Y*Z) + y*Z + z.
    Each coordinate is clamped to the grid (the reference would index out of bounds).
  * thrust::sort_by_key(hash, indices) (Fluid/Fluid_3D_common.cuh:35) as a stable sort
    (numpy's stable argsort defines the order).
  * applySortImpl (Fluid/Fluid_3D_kernels.cuh:95-102): records gathered in sorted order.
  * findCellStartEndImpl (Fluid/Fluid_kernels.cuh:3-22): first and last (inclusive)
    sorted index per cell, -1 for empty cells (cudaMemset 0xFF, Fluid_3D_common.cuh:47-48).
Pinned by the report's worked example (tests/golden/report_spatial_index.json) and by
the FLIP oracle's hash_sort on seeded scenes (tests/test_oracle_pins.py).
"""
from __future__ import annotations

import numpy as np


def hash_records(records: np.ndarray, sizes, cell_size: float) -> np.ndarray:
    X, Y, Z = (int(v) for v in sizes)
    p = np.ascontiguousarray(records[:, :3], np.float32)
    q = p / np.float32(cell_size)  # IEEE float32 division, element by element
    c = np.trunc(q.astype(np.float64)).astype(np.int64)
    c[:, 0] = np.clip(c[:, 0], 0, X - 1)
    c[:, 1] = np.clip(c[:, 1], 0, Y - 1)
    c[:, 2] = np.clip(c[:, 2], 0, Z - 1)
    return (c[:, 0] * (Y * Z) + c[:, 1] * Z + c[:, 2]).astype(np.int32)


def spatial_hash(records: np.ndarray, sizes, cell_size: float) -> dict:
    X, Y, Z = (int(v) for v in sizes)
    n = records.shape[0]
    h = hash_records(records, sizes, cell_size) if n else np.zeros(0, np.int32)
    perm = np.argsort(h, kind="stable").astype(np.int32)
    sh = h[perm]
    cells = X * Y * Z
    start = np.full(cells, -1, np.int32)
    end = np.full(cells, -1, np.int32)
    if n:
        first = np.r_[True, sh[1:] != sh[:-1]]
        last = np.r_[sh[1:] != sh[:-1], True]
        idx = np.arange(n, dtype=np.int32)
        start[sh[first]] = idx[first]
        end[sh[last]] = idx[last]
    return dict(perm=perm, sorted_hash=sh, cell_start=start, cell_end=end,
                sorted_records=records[perm])
