"""Python host mirror of the reference's solver surface over the C ABI.

`Fluid` keeps the shape of Fluid_3D_FLIP::Fluid (Fluid/Fluid_3D_FLIP.cuh:53-156):
init(config), simulationStep(), draw(), physicalTime, getCenter(),
getContainerSize(). `FlipHandle` exposes the rest of the ABI (stages, fields,
cell bins, FP64 checks, timing) with numpy arrays in the reference extents.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import FIELDS, STAGES, STAGE_COUNT, Config, Stats, check, load


def _fp(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def seed_particles(cfg: Config) -> tuple[np.ndarray, np.ndarray]:
    """Host seeding (createSquareFluid / createSphereFluid). Returns (pos[n,3], frac4[n,4])."""
    lib = load()
    n = C.c_int64()
    check(lib.aq_flip_seed_particles(C.byref(cfg), None, None, 0, C.byref(n)))
    pos = np.empty((n.value, 3), np.float32)
    frac = np.empty((n.value, 4), np.float32)
    check(lib.aq_flip_seed_particles(C.byref(cfg), _fp(pos), _fp(frac), n.value, C.byref(n)))
    return pos, frac


class FlipHandle:
    """One simulation on one GPU (an aq_flip handle)."""

    def __init__(self, cfg: Config, pos=None, vel=None, frac4=None):
        self.lib = load()
        self.cfg = cfg
        h = C.c_void_p()
        if pos is None:
            check(self.lib.aq_flip_create(C.byref(cfg), C.byref(h)))
        else:
            pos = np.ascontiguousarray(pos, np.float32)
            vel = None if vel is None else np.ascontiguousarray(vel, np.float32)
            frac4 = None if frac4 is None else np.ascontiguousarray(frac4, np.float32)
            check(self.lib.aq_flip_create_with_particles(
                C.byref(cfg), _fp(pos), _fp(vel), _fp(frac4), pos.shape[0], C.byref(h)))
        self.h = h
        self.size = tuple(cfg.size)

    def save(self, path) -> None:
        """Checkpoint (aq_flip_save): everything a later step or export reads."""
        check(self.lib.aq_flip_save(self.h, str(path).encode()))

    @classmethod
    def load(cls, path, cfg: Config | None = None) -> "FlipHandle":
        """Resume from a checkpoint (aq_flip_load)."""
        self = cls.__new__(cls)
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.aq_flip_load(str(path).encode(), C.byref(h)))
        self.h = h
        self.cfg = cfg
        with open(path, "rb") as f:  # grid extents: the first configuration fields
            self.size = tuple(int(v) for v in np.frombuffer(f.read(24)[12:24], "<i4"))
        return self

    def close(self) -> None:
        if self.h:
            self.lib.aq_flip_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- stepping
    def step(self) -> None:
        check(self.lib.aq_flip_step(self.h))

    def steps(self, n: int) -> None:
        check(self.lib.aq_flip_steps(self.h, n))

    def run_stage(self, name: str) -> None:
        check(self.lib.aq_flip_run_stage(self.h, STAGES[name]))

    # ---- particles
    @property
    def particle_count(self) -> int:
        return int(self.lib.aq_flip_particle_count(self.h))

    def get_particles(self):
        n = self.particle_count
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        frac = np.empty((n, 4), np.float32)
        check(self.lib.aq_flip_get_particles(self.h, _fp(pos), _fp(vel), _fp(frac)))
        return pos, vel, frac

    def set_particles(self, pos, vel=None, frac4=None) -> None:
        pos = np.ascontiguousarray(pos, np.float32)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32)
        frac4 = None if frac4 is None else np.ascontiguousarray(frac4, np.float32)
        check(self.lib.aq_flip_set_particles(self.h, _fp(pos), _fp(vel), _fp(frac4), pos.shape[0]))

    # ---- grid fields (reference extents, x-major, z fastest)
    def field_shape(self, name: str) -> tuple:
        fid, dt, width, kind = FIELDS[name]
        X, Y, Z = self.size
        ext = (X + 1, Y + 1, Z + 1) if kind == "face" else (X, Y, Z)
        return ext + ((width,) if width > 1 else ())

    def get_field(self, name: str) -> np.ndarray:
        fid, dt, width, kind = FIELDS[name]
        out = np.empty(self.field_shape(name), dt)
        check(self.lib.aq_flip_get_field(self.h, fid, out.ctypes.data, out.nbytes))
        return out

    def set_field(self, name: str, value: np.ndarray) -> None:
        fid, dt, width, kind = FIELDS[name]
        a = np.ascontiguousarray(value, dt).reshape(self.field_shape(name))
        check(self.lib.aq_flip_set_field(self.h, fid, a.ctypes.data, a.nbytes))

    def get_cells(self):
        X, Y, Z = self.size
        n = self.particle_count
        e1 = (X + 1) * (Y + 1) * (Z + 1)
        start = np.empty(e1, np.int32)
        end = np.empty(e1, np.int32)
        hsh = np.empty(n, np.int32)
        perm = np.empty(n, np.int32)
        check(self.lib.aq_flip_get_cells(self.h, _ip(start), _ip(end), _ip(hsh), _ip(perm)))
        return start, end, hsh, perm

    # ---- FP64 checks and solves
    def residual(self) -> float:
        r = C.c_double()
        check(self.lib.aq_flip_residual_fp64(self.h, C.byref(r)))
        return r.value

    def stats(self) -> dict:
        s = Stats()
        check(self.lib.aq_flip_volume_stats(self.h, C.byref(s)))
        return dict(particles=s.particles, fluid_cells=s.fluid_cells,
                    phase_sum=list(s.phase_sum), mean_ppc=s.mean_ppc,
                    pos_min=list(s.pos_min), pos_max=list(s.pos_max))

    def solve_pressure(self, tol: float, max_iters: int, check_every: int = 10):
        it = C.c_int32()
        rel = C.c_double()
        st = self.lib.aq_flip_solve_pressure(self.h, tol, max_iters, check_every,
                                             C.byref(it), C.byref(rel))
        if st not in (_abi.AQ_OK, _abi.AQ_ESTALL):
            check(st)
        return it.value, rel.value, st == _abi.AQ_OK

    def solve_pressure_mgpcg(self, tol: float, max_iters: int = 200, check_every: int = 1,
                             from_current: bool = False):
        """Multigrid-preconditioned CG (FP64). Returns (iterations, FP64 residual of the
        stored FP32 pressure, converged)."""
        it = C.c_int32()
        rel = C.c_double()
        st = self.lib.aq_flip_solve_pressure_mgpcg(self.h, tol, max_iters, check_every,
                                                   1 if from_current else 0, C.byref(it), C.byref(rel))
        if st not in (_abi.AQ_OK, _abi.AQ_ESTALL):
            check(st)
        return it.value, rel.value, st == _abi.AQ_OK

    # ---- timing
    def set_timing(self, on: bool) -> None:
        check(self.lib.aq_flip_set_timing(self.h, 1 if on else 0))

    def reset_timing(self) -> None:
        check(self.lib.aq_flip_reset_timing(self.h))

    def timing(self) -> dict:
        ms = (C.c_double * STAGE_COUNT)()
        calls = (C.c_int64 * STAGE_COUNT)()
        check(self.lib.aq_flip_get_timing(self.h, ms, calls))
        return {name: (ms[i], calls[i]) for name, i in STAGES.items()}

    def paired_diffusion_computed(self) -> int:
        """Paired diffusion sweeps computed (not skipped as unchanged) since creation."""
        n = C.c_int64()
        check(self.lib.aq_flip_paired_diffusion_computed(self.h, C.byref(n)))
        return n.value

    @property
    def physical_time(self) -> float:
        return float(self.lib.aq_flip_physical_time(self.h))


class Fluid:
    """Fluid_3D_FLIP::Fluid surface (Fluid/Fluid_3D_FLIP.cuh:53-156) for Python callers."""

    def __init__(self):
        self._h: FlipHandle | None = None
        self.config: Config | None = None

    def init(self, config: Config) -> None:
        """Fluid::init (Fluid/Fluid_3D_FLIP.cu:435-505) without the renderers."""
        if self._h is not None:
            self._h.close()
        self.config = config
        self._h = FlipHandle(config)

    def simulationStep(self) -> None:
        """Fluid::simulationStep (Fluid/Fluid_3D_FLIP.cu:316-352)."""
        self._h.step()

    def draw(self, drawCommand=None) -> None:
        """Rendering is out of scope; kept for the interface."""

    @property
    def physicalTime(self) -> float:
        return 0.0 if self._h is None else self._h.physical_time

    def getCenter(self):
        c = (C.c_float * 3)()
        check(self._h.lib.aq_flip_get_center(self._h.h, c))
        return tuple(c)

    def getContainerSize(self) -> float:
        return float(self._h.lib.aq_flip_container_size(self._h.h))

    @property
    def handle(self) -> FlipHandle:
        return self._h
