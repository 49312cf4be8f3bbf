"""Scene configurations of the FLIP ball-drop workloads.

Every scene is the reference's pool + ball set-up (InitializationVolume,
Fluid/FluidConfig.cuh:20-39) with particles jittered uniformly within their
sub-cells (particle_jitter = 1.0), zero initial velocity and the reference's
10-unit domain (Fluid/Fluid_3D_FLIP.cuh:62). Scene coordinates are normalised to
the unit cube. Seeds are 0xA9A00000 + scene number.
"""
from __future__ import annotations

import ctypes as C

from ._abi import (AQ_PRESSURE_PREVIOUS, AQ_PRESSURE_ZERO, AQ_SHAPE_BOX, AQ_SHAPE_SPHERE, Config,
                   check, load)


def default_config() -> Config:
    """The reference defaults of FluidConfig (aq_flip_config_default)."""
    cfg = Config()
    check(load().aq_flip_config_default(C.byref(cfg)))
    return cfg


def _set_volumes(cfg: Config, vols: list[dict]) -> None:
    cfg.n_volumes = len(vols)
    for i, v in enumerate(vols):
        dst = cfg.volumes[i]
        dst.shape = v["shape"]
        for k in ("box_min", "box_max", "center"):
            vals = v.get(k, (0.0, 0.0, 0.0))
            for a in range(3):
                getattr(dst, k)[a] = vals[a]
        dst.radius = v.get("radius", 0.0)
        dst.phase = v.get("phase", 0)


def make_config(n: int, *, dt: float, pressure_iters: int, diffusion_iters: int = 100,
                pool_depth: float = 0.25, radius: float = 0.2, center=(0.5, 0.7, 0.5),
                sphere_phase: int = 0, phase_count: int = 1, diffusion_coeff: float = 1e-3,
                seed: int = 0xA9A00000, jitter: float = 1.0,
                pressure_init: int = AQ_PRESSURE_ZERO, ref_boundary_coverage: int = 1,
                flip_coeff: float = 0.95, extent=None) -> Config:
    cfg = Config()
    size = extent if extent is not None else (n, n, n)
    for a in range(3):
        cfg.size[a] = size[a]
    cfg.dt = dt
    cfg.flip_coeff = flip_coeff
    cfg.particle_jitter = jitter
    cfg.grid_physical_size = 10.0
    cfg.pressure_iters = pressure_iters
    cfg.diffusion_iters = diffusion_iters
    cfg.phase_count = phase_count
    cfg.gravity[0], cfg.gravity[1], cfg.gravity[2] = 0.0, -9.8, 0.0
    for a in range(4):
        cfg.phase_densities[a] = 1.0
    cfg.diffusion_coeff = diffusion_coeff
    cfg.pressure_init = pressure_init
    cfg.ref_boundary_coverage = ref_boundary_coverage
    cfg.seed = seed
    _set_volumes(cfg, [
        dict(shape=AQ_SHAPE_BOX, box_min=(0.0, 0.0, 0.0), box_max=(1.0, pool_depth, 1.0), phase=0),
        dict(shape=AQ_SHAPE_SPHERE, center=center, radius=radius, phase=sphere_phase),
    ])
    return cfg


def ball_drop_16() -> Config:
    """16^3, 8 ppc, pool 0.25 + ball r 0.2 at (0.5,0.7,0.5), dt 0.01, 60 Jacobi
    sweeps from zero, single phase; 50 steps."""
    return make_config(16, dt=0.01, pressure_iters=60, seed=0xA9A00000 + 1)


def ball_drop_64() -> Config:
    """64^3, 8 ppc, same scene, dt 0.005, 100 Jacobi sweeps from zero; 200 steps."""
    return make_config(64, dt=0.005, pressure_iters=100, seed=0xA9A00000 + 2)


def multiphase_128() -> Config:
    """128^3, 8 ppc (about 4.4M particles): a second-phase ball of radius 0.15 with
    concentration 1 over a phase-0 pool, diffusion coefficient 1e-3 with 100
    diffusion sweeps, 150 Jacobi sweeps from zero, dt 0.005; 100 steps."""
    return make_config(128, dt=0.005, pressure_iters=150, diffusion_iters=100, radius=0.15,
                       sphere_phase=1, phase_count=2, seed=0xA9A00000 + 3)


def ball_drop_256() -> Config:
    """256^3, 8 ppc (about 45M particles): ball r 0.2 into a pool of depth 0.3."""
    return make_config(256, dt=0.005, pressure_iters=150, pool_depth=0.3, radius=0.2,
                       seed=0xA9A00000 + 5)


SCENES = {
    "ball_drop_16": ball_drop_16,
    "ball_drop_64": ball_drop_64,
    "multiphase_128": multiphase_128,
    "ball_drop_256": ball_drop_256,
}

__all__ = ["default_config", "make_config", "SCENES", "ball_drop_16", "ball_drop_64",
           "multiphase_128", "ball_drop_256", "AQ_PRESSURE_ZERO", "AQ_PRESSURE_PREVIOUS"]
