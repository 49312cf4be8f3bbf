"""ctypes binding of oracle/_build/liboracle_flip.so (TEST INFRASTRUCTURE ONLY).

The oracle is a C++17 restatement of the reference kernels (oracle/flip_oracle.hpp);
Oracle(cfg, fp64=False) is the FP32 parity reference, fp64=True the FP64 run.
Fields come back as float64 arrays in the same extents and layout the GPU
library exports (x-major, z fastest; 4-wide fields interleaved).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from aquarius_amd._abi import FIELDS, STAGES, Config

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle_flip.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    P, D, I, F = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_float)
    sig = {
        "orc_set_threads": (None, [C.c_int]),
        "orc_max_threads": (C.c_int, []),
        "orc_seed": (C.c_int, [C.POINTER(Config), F, F, C.c_int64, C.POINTER(C.c_int64)]),
        "orc_create": (P, [C.POINTER(Config), C.c_int]),
        "orc_destroy": (None, [P]),
        "orc_set_particles": (None, [P, F, F, F, C.c_int]),
        "orc_particle_count": (C.c_int, [P]),
        "orc_get_particles": (None, [P, D, D, D]),
        "orc_init": (None, [P]),
        "orc_step": (None, [P]),
        "orc_stage": (None, [P, C.c_int]),
        "orc_substage": (None, [P, C.c_int]),
        "orc_jacobi_sweeps": (None, [P, C.c_int]),
        "orc_field_elems": (C.c_size_t, [P, C.c_int]),
        "orc_get_field": (None, [P, C.c_int, D]),
        "orc_set_field": (None, [P, C.c_int, D]),
        "orc_get_cells": (None, [P, I, I, I, I]),
        "orc_residual": (C.c_double, [P]),
        "orc_physical_time": (C.c_double, [P]),
        "orc_time_steps": (C.c_double, [P, C.c_int]),
        "orc_time_step_stages": (C.c_double, [P, D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def set_threads(n: int) -> None:
    load().orc_set_threads(n)


def max_threads() -> int:
    return load().orc_max_threads()


def seed(cfg: Config) -> tuple[np.ndarray, np.ndarray]:
    lib = load()
    n = C.c_int64()
    lib.orc_seed(C.byref(cfg), None, None, 0, C.byref(n))
    pos = np.empty((n.value, 3), np.float32)
    frac = np.empty((n.value, 4), np.float32)
    rc = lib.orc_seed(C.byref(cfg), pos.ctypes.data_as(C.POINTER(C.c_float)),
                      frac.ctypes.data_as(C.POINTER(C.c_float)), n.value, C.byref(n))
    assert rc == 0
    return pos, frac


def _f(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


SUBSTAGES = {"hash_sort": 0, "reset": 1, "p2g_accum": 2, "p2g_divide": 3, "p2g_fractions": 4,
             "count": 5, "divergence": 6, "jacobi_sweep": 7, "extrapolate_once": 8,
             "diffusion_sweep": 9}


class Oracle:
    def __init__(self, cfg: Config, fp64: bool = False, pos=None, vel=None, frac4=None,
                 init: bool = True):
        self.lib = load()
        self.cfg = cfg
        self.fp64 = fp64
        self.size = tuple(cfg.size)
        self.h = self.lib.orc_create(C.byref(cfg), 1 if fp64 else 0)
        if pos is None:
            pos, frac4 = seed(cfg)
        self.set_particles(pos, vel, frac4)
        if init:
            self.lib.orc_init(self.h)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, pos, vel=None, frac4=None):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32)
        frac4 = None if frac4 is None else np.ascontiguousarray(frac4, np.float32)
        self.lib.orc_set_particles(self.h, _f(pos), _f(vel), _f(frac4), pos.shape[0])

    @property
    def particle_count(self) -> int:
        return self.lib.orc_particle_count(self.h)

    def get_particles(self):
        n = self.particle_count
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        frac = np.empty((n, 4))
        D = C.POINTER(C.c_double)
        self.lib.orc_get_particles(self.h, pos.ctypes.data_as(D), vel.ctypes.data_as(D),
                                   frac.ctypes.data_as(D))
        return pos, vel, frac

    def step(self, n: int = 1):
        for _ in range(n):
            self.lib.orc_step(self.h)

    def run_stage(self, name: str):
        self.lib.orc_stage(self.h, STAGES[name])

    def substage(self, name: str):
        self.lib.orc_substage(self.h, SUBSTAGES[name])

    def jacobi_sweeps(self, n: int):
        self.lib.orc_jacobi_sweeps(self.h, n)

    def field_shape(self, name: str):
        fid, dt, width, kind = FIELDS[name]
        X, Y, Z = self.size
        ext = (X + 1, Y + 1, Z + 1) if kind == "face" else (X, Y, Z)
        return ext + ((width,) if width > 1 else ())

    def get_field(self, name: str) -> np.ndarray:
        fid = FIELDS[name][0]
        out = np.empty(self.field_shape(name), np.float64)
        self.lib.orc_get_field(self.h, fid, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_field(self, name: str, value) -> None:
        fid = FIELDS[name][0]
        a = np.ascontiguousarray(value, np.float64).reshape(self.field_shape(name))
        self.lib.orc_set_field(self.h, fid, a.ctypes.data_as(C.POINTER(C.c_double)))

    def get_cells(self):
        X, Y, Z = self.size
        n = self.particle_count
        e1 = (X + 1) * (Y + 1) * (Z + 1)
        start = np.empty(e1, np.int32)
        end = np.empty(e1, np.int32)
        hsh = np.empty(n, np.int32)
        perm = np.empty(n, np.int32)
        I = C.POINTER(C.c_int32)
        self.lib.orc_get_cells(self.h, start.ctypes.data_as(I), end.ctypes.data_as(I),
                               hsh.ctypes.data_as(I), perm.ctypes.data_as(I))
        return start, end, hsh, perm

    def residual(self) -> float:
        return self.lib.orc_residual(self.h)

    @property
    def physical_time(self) -> float:
        return self.lib.orc_physical_time(self.h)

    def time_steps(self, n: int) -> float:
        return self.lib.orc_time_steps(self.h, n)

    STAGE_NAMES = ("g2p", "advect", "transfer_to_grid", "gravity_boundary", "divergence", "pressure",
                   "project", "extrapolate", "diffusion")

    def time_step_stages(self):
        """One step, timed: (total seconds, {stage: seconds})."""
        acc = np.zeros(9)
        total = self.lib.orc_time_step_stages(self.h, acc.ctypes.data_as(C.POINTER(C.c_double)))
        return total, dict(zip(self.STAGE_NAMES, acc.tolist()))
