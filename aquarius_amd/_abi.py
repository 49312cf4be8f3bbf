"""ctypes mirror of include/aqflip.h and the loader of libaqflip.so.

The library is built in-tree (python -m aquarius_amd.build). Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libaqflip.so"

AQ_OK, AQ_EINVAL, AQ_ENOMEM, AQ_EHIP, AQ_ESTALL = 0, 1, 2, 3, 4
STATUS_NAMES = {0: "AQ_OK", 1: "AQ_EINVAL", 2: "AQ_ENOMEM", 3: "AQ_EHIP", 4: "AQ_ESTALL"}

AQ_SHAPE_SPHERE, AQ_SHAPE_BOX = 0, 1
AQ_PRESSURE_ZERO, AQ_PRESSURE_PREVIOUS = 0, 1
AQ_MAX_VOLUMES = 16

# AQ_FIELD_* ids, element type, width and extent kind ("face" = (N+1)^3, "cell" = N^3)
FIELDS = {
    "content": (0, "i4", 1, "face"),
    "velocity": (1, "f4", 4, "face"),
    "new_velocity": (2, "f4", 4, "face"),
    "pressure": (3, "f4", 1, "cell"),
    "divergence": (4, "f4", 1, "cell"),
    "density": (5, "f4", 1, "cell"),
    "particle_count": (6, "i4", 1, "cell"),
    "has_velocity": (7, "i4", 4, "cell"),
    "volume_fractions": (8, "f4", 4, "cell"),
    "new_volume_fractions": (9, "f4", 4, "cell"),
}

STAGES = {
    "g2p": 0,
    "advect": 1,
    "transfer_to_grid": 2,
    "gravity": 3,
    "boundary": 4,
    "divergence": 5,
    "pressure": 6,
    "project": 7,
    "extrapolate": 8,
    "diffusion": 9,
    "pressure_diffusion": 10,  # paired pressure + diffusion sweeps (timing only)
    "p2g": 11,  # the P2G kernel of transfer_to_grid inside a step (timing only)
}
STAGE_COUNT = 12


class Volume(C.Structure):
    _fields_ = [
        ("shape", C.c_int32),
        ("box_min", C.c_float * 3),
        ("box_max", C.c_float * 3),
        ("center", C.c_float * 3),
        ("radius", C.c_float),
        ("phase", C.c_int32),
    ]


class Config(C.Structure):
    _fields_ = [
        ("size", C.c_int32 * 3),
        ("dt", C.c_float),
        ("flip_coeff", C.c_float),
        ("particle_jitter", C.c_float),
        ("grid_physical_size", C.c_float),
        ("pressure_iters", C.c_int32),
        ("diffusion_iters", C.c_int32),
        ("phase_count", C.c_int32),
        ("gravity", C.c_float * 3),
        ("phase_densities", C.c_float * 4),
        ("diffusion_coeff", C.c_float),
        ("pressure_init", C.c_int32),
        ("ref_boundary_coverage", C.c_int32),
        ("seed", C.c_uint64),
        ("n_volumes", C.c_int32),
        ("volumes", Volume * AQ_MAX_VOLUMES),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("particles", C.c_int64),
        ("fluid_cells", C.c_int64),
        ("phase_sum", C.c_double * 4),
        ("mean_ppc", C.c_double),
        ("pos_min", C.c_double * 3),
        ("pos_max", C.c_double * 3),
    ]


class UNetConfig(C.Structure):
    _fields_ = [
        ("channels", C.c_int32 * 3),
        ("seed", C.c_uint64),
        ("lr", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("target_scale", C.c_float),
        ("infer_extent", C.c_int32),
    ]


class Solver(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("tol", C.c_double),
        ("max_iters", C.c_int32),
        ("check_every", C.c_int32),
        ("stall_checks", C.c_int32),
        ("nn_refresh", C.c_int32),
    ]


class SolverStats(C.Structure):
    _fields_ = [
        ("solves", C.c_int64),
        ("sweeps", C.c_int64),
        ("fallbacks", C.c_int64),
        ("capped", C.c_int64),
        ("last_sweeps", C.c_int32),
        ("last_residual", C.c_double),
        ("last_guess_residual", C.c_double),
        ("last_fallback", C.c_int32),
        ("nn_corrections", C.c_int64),
        ("nn_rejected", C.c_int64),
    ]


AQ_SOLVE_FIXED, AQ_SOLVE_TOL_ZERO, AQ_SOLVE_TOL_NN, AQ_SOLVE_MGPCG = 0, 1, 2, 3

# name -> (restype, argtypes); every symbol declared in include/aqflip.h and
# include/aqflip_unet.h
_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
PROTOTYPES = {
    "aq_flip_config_default": (C.c_int, [C.POINTER(Config)]),
    "aq_flip_seed_particles": (C.c_int, [C.POINTER(Config), _FP, _FP, C.c_int64, C.POINTER(C.c_int64)]),
    "aq_flip_create": (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    "aq_flip_create_with_particles": (C.c_int, [C.POINTER(Config), _FP, _FP, _FP, C.c_int64, C.POINTER(_P)]),
    "aq_flip_destroy": (None, [_P]),
    "aq_flip_step": (C.c_int, [_P]),
    "aq_flip_steps": (C.c_int, [_P, C.c_int32]),
    "aq_flip_run_stage": (C.c_int, [_P, C.c_int32]),
    "aq_flip_particle_count": (C.c_int64, [_P]),
    "aq_flip_get_particles": (C.c_int, [_P, _FP, _FP, _FP]),
    "aq_flip_set_particles": (C.c_int, [_P, _FP, _FP, _FP, C.c_int64]),
    "aq_flip_field_bytes": (C.c_size_t, [_P, C.c_int32]),
    "aq_flip_get_field": (C.c_int, [_P, C.c_int32, C.c_void_p, C.c_size_t]),
    "aq_flip_set_field": (C.c_int, [_P, C.c_int32, C.c_void_p, C.c_size_t]),
    "aq_flip_get_cells": (C.c_int, [_P, _IP, _IP, _IP, _IP]),
    "aq_flip_residual_fp64": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "aq_flip_volume_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "aq_flip_solve_pressure": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32, _IP, C.POINTER(C.c_double)]),
    "aq_flip_set_timing": (C.c_int, [_P, C.c_int32]),
    "aq_flip_get_timing": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "aq_flip_reset_timing": (C.c_int, [_P]),
    "aq_flip_paired_diffusion_computed": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "aq_flip_physical_time": (C.c_double, [_P]),
    "aq_flip_get_center": (C.c_int, [_P, _FP]),
    "aq_flip_container_size": (C.c_float, [_P]),
    "aq_last_error": (C.c_char_p, []),
    "aq_flip_version": (C.c_char_p, []),
    # include/aqflip_unet.h
    "aq_unet_config_default": (C.c_int, [C.POINTER(UNetConfig)]),
    "aq_unet_create": (C.c_int, [C.POINTER(UNetConfig), C.POINTER(_P)]),
    "aq_unet_destroy": (None, [_P]),
    "aq_unet_param_count": (C.c_int64, [_P]),
    "aq_unet_get_params": (C.c_int, [_P, _FP]),
    "aq_unet_set_params": (C.c_int, [_P, _FP]),
    "aq_unet_forward_host": (C.c_int, [_P, _FP, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _FP]),
    "aq_unet_gradients_host": (C.c_int, [_P, _FP, _FP, _FP, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(C.c_double), _FP]),
    "aq_unet_adam_step_host": (C.c_int, [_P, _FP]),
    "aq_unet_set_lr": (C.c_int, [_P, C.c_float]),
    "aq_unet_get_adam_state": (C.c_int, [_P, _FP, _FP, C.POINTER(C.c_int64)]),
    "aq_unet_dataset_reserve": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "aq_unet_dataset_add_flip": (C.c_int, [_P, _P]),
    "aq_unet_dataset_size": (C.c_int32, [_P]),
    "aq_unet_train": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _FP]),
    "aq_unet_eval": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "aq_flip_nn_guess": (C.c_int, [_P, _P]),
    "aq_flip_nn_correct": (C.c_int, [_P, _P]),
    "aq_flip_solve_pressure_nn": (C.c_int, [_P, _P, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "aq_flip_set_solver": (C.c_int, [_P, C.POINTER(Solver), _P]),
    "aq_flip_solver_stats_get": (C.c_int, [_P, C.POINTER(SolverStats), C.c_int32]),
    "aq_unet_mfma_probe": (C.c_int, [_FP, _FP, _FP]),
    "aq_flip_solve_pressure_mgpcg": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                               C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "aq_flip_save": (C.c_int, [_P, C.c_char_p]),
    "aq_flip_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    # spatial hashing service (include/aqflip_hash.h)
    "aq_hash_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_int64, C.POINTER(_P)]),
    "aq_hash_destroy": (None, [_P]),
    "aq_hash_sort_host": (C.c_int, [_P, _FP, C.c_int64, C.c_int32, _IP, _IP, _IP, _IP, C.c_int64, _FP]),
    "aq_hash_sort_device": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P, _P, _P, C.c_int64, _P, _P]),
    "aq_hash_sort_device_noalloc": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_float, _P, C.c_int64,
                                              C.c_int32, _P, _P, _P, _P, C.c_int64, _P, _P]),
}

_lib = None


class AqError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load libaqflip.so (built in-tree) and bind every prototype."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m aquarius_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != AQ_OK:
        raise AqError(status, load().aq_last_error().decode())
