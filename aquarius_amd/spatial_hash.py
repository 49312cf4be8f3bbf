"""Python mirror of the spatial hashing service (include/aqflip_hash.h), the
standalone form of performSpatialHashing2 (Fluid/Fluid_3D_common.cuh:28-56)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import check, load


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class SpatialHash:
    """Stable sort of particle records by grid cell on the GPU."""

    def __init__(self, sizes, cell_size: float, capacity: int = 0):
        self.lib = load()
        self.sizes = tuple(int(v) for v in sizes)
        h = C.c_void_p()
        check(self.lib.aq_hash_create(*self.sizes, float(cell_size), int(capacity), C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.aq_hash_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cells(self) -> int:
        return self.sizes[0] * self.sizes[1] * self.sizes[2]

    def sort(self, records: np.ndarray, cell_count: int = 0) -> dict:
        """records: (n, stride) float32, position in the first three columns.
        cell_count: length of the cell_start / cell_end outputs (0: the grid's cells;
        entries past the grid's cells are -1)."""
        rec = np.ascontiguousarray(records, np.float32)
        if rec.ndim != 2:
            raise ValueError("records must be (n, stride)")
        n, stride = rec.shape
        out = dict(perm=np.empty(n, np.int32), sorted_hash=np.empty(n, np.int32),
                   cell_start=np.empty(max(cell_count, self.cells), np.int32),
                   cell_end=np.empty(max(cell_count, self.cells), np.int32),
                   sorted_records=np.empty_like(rec))
        check(self.lib.aq_hash_sort_host(self.h, _ptr(rec, C.c_float), n, stride,
                                         _ptr(out["perm"], C.c_int32), _ptr(out["sorted_hash"], C.c_int32),
                                         _ptr(out["cell_start"], C.c_int32), _ptr(out["cell_end"], C.c_int32),
                                         int(cell_count), _ptr(out["sorted_records"], C.c_float)))
        return out

    def sort_device(self, d_records: int, n: int, stride: int, d_perm=0, d_sorted_hash=0, d_cell_start=0,
                    d_cell_end=0, d_sorted_records=0, stream: int = 0, cell_count: int = 0) -> None:
        """Device pointers (integers, e.g. from torch's data_ptr()); 0 skips an output.
        stream: the hipStream_t (an integer, e.g. torch.cuda.current_stream().cuda_stream;
        0 = the null stream) whose queued work the sort runs after. cell_count: length
        of the cell arrays (0: the grid's cells)."""
        check(self.lib.aq_hash_sort_device(self.h, d_records or None, n, stride, d_perm or None,
                                           d_sorted_hash or None, d_cell_start or None, d_cell_end or None,
                                           int(cell_count), d_sorted_records or None, stream or None))
