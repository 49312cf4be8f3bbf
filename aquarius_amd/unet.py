"""Python mirror of the learned warm start (include/aqflip_unet.h).

`UNet` wraps an aq_unet handle: parameters, host-tensor forward/gradients, the
device-resident training set and Adam training. `attach_solver` selects how a
FlipHandle solves for pressure (fixed sweeps, to a tolerance from zero, or to a
tolerance from the U-Net guess with zero-guess fallback).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import (AQ_SOLVE_FIXED, AQ_SOLVE_MGPCG, AQ_SOLVE_TOL_NN, AQ_SOLVE_TOL_ZERO, Solver, SolverStats,
                   UNetConfig, check, load)


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def default_config(**kw) -> UNetConfig:
    cfg = UNetConfig()
    check(load().aq_unet_config_default(C.byref(cfg)))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


class UNet:
    def __init__(self, cfg: UNetConfig | None = None, **kw):
        self.lib = load()
        self.cfg = cfg if cfg is not None else default_config(**kw)
        h = C.c_void_p()
        check(self.lib.aq_unet_create(C.byref(self.cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.aq_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def param_count(self) -> int:
        return int(self.lib.aq_unet_param_count(self.h))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.param_count, np.float32)
        check(self.lib.aq_unet_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.param_count
        check(self.lib.aq_unet_set_params(self.h, _fp(p)))

    def save(self, path) -> None:
        np.savez(path, params=self.get_params(), target_scale=np.float32(self.cfg.target_scale))

    def load_file(self, path) -> None:
        with np.load(path, allow_pickle=False) as z:
            if "target_scale" in z and float(z["target_scale"]) != float(self.cfg.target_scale):
                raise ValueError(f"{path}: weights were trained with target_scale "
                                 f"{float(z['target_scale'])}, this network uses {self.cfg.target_scale}")
            self.set_params(z["params"])

    def forward(self, x: np.ndarray) -> np.ndarray:
        """x: (B, D, H, W, 2) float -> (B, D, H, W) float."""
        x = np.ascontiguousarray(x, np.float32)
        B, D, H, W, _ = x.shape
        y = np.empty((B, D, H, W), np.float32)
        check(self.lib.aq_unet_forward_host(self.h, _fp(x), B, D, H, W, _fp(y)))
        return y

    def gradients(self, x, target, mask):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        B, D, H, W, _ = x.shape
        g = np.empty(self.param_count, np.float32)
        loss = C.c_double()
        check(self.lib.aq_unet_gradients_host(self.h, _fp(x), _fp(t), _fp(m), B, D, H, W,
                                              C.byref(loss), _fp(g)))
        return loss.value, g

    def adam_step(self, grads) -> None:
        """One Adam step of the FP32 master parameters with the given gradients."""
        g = np.ascontiguousarray(grads, np.float32)
        assert g.size == self.param_count
        check(self.lib.aq_unet_adam_step_host(self.h, _fp(g)))

    def set_lr(self, lr: float) -> None:
        check(self.lib.aq_unet_set_lr(self.h, float(lr)))
        self.cfg.lr = float(lr)

    def adam_state(self):
        """(m, v, t): the Adam moments and the step count."""
        m = np.empty(self.param_count, np.float32)
        v = np.empty(self.param_count, np.float32)
        t = C.c_int64()
        check(self.lib.aq_unet_get_adam_state(self.h, _fp(m), _fp(v), C.byref(t)))
        return m, v, t.value

    # ---- device training set
    def reserve(self, capacity: int, shape) -> None:
        check(self.lib.aq_unet_dataset_reserve(self.h, capacity, *shape))

    def add_sample(self, flip) -> None:
        check(self.lib.aq_unet_dataset_add_flip(self.h, flip.h))

    @property
    def dataset_size(self) -> int:
        return int(self.lib.aq_unet_dataset_size(self.h))

    def train(self, steps: int, batch: int, train_count: int, seed: int = 0) -> np.ndarray:
        hist = np.empty(max(steps, 1), np.float32)
        check(self.lib.aq_unet_train(self.h, steps, batch, train_count, seed, _fp(hist)))
        return hist[:steps]

    def eval(self, first: int, count: int) -> float:
        loss = C.c_double()
        check(self.lib.aq_unet_eval(self.h, first, count, C.byref(loss)))
        return loss.value


def nn_guess(flip, net: UNet) -> None:
    check(load().aq_flip_nn_guess(flip.h, net.h))


def nn_correct(flip, net: UNet) -> None:
    check(load().aq_flip_nn_correct(flip.h, net.h))


def solve_pressure_nn(flip, net: UNet, tol: float, max_iters: int, check_every: int = 10,
                      nn_refresh: int = 0):
    """The U-Net guess refined by Jacobi with a correction every nn_refresh sweeps
    (aq_flip_solve_pressure_nn). Returns (sweeps, residual, converged, corrections,
    corrections undone)."""
    from . import _abi
    it, res, nc, nr = C.c_int32(), C.c_double(), C.c_int32(), C.c_int32()
    st = load().aq_flip_solve_pressure_nn(flip.h, net.h, tol, max_iters, check_every, nn_refresh,
                                          C.byref(it), C.byref(res), C.byref(nc), C.byref(nr))
    if st not in (_abi.AQ_OK, _abi.AQ_ESTALL):
        check(st)
    return it.value, res.value, st == _abi.AQ_OK, nc.value, nr.value


def attach_solver(flip, mode: str, tol: float = 1e-4, max_iters: int = 100000,
                  check_every: int = 50, stall_checks: int = 20, net: UNet | None = None,
                  nn_refresh: int = 0) -> None:
    modes = {"fixed": AQ_SOLVE_FIXED, "tol_zero": AQ_SOLVE_TOL_ZERO, "tol_nn": AQ_SOLVE_TOL_NN,
             "mgpcg": AQ_SOLVE_MGPCG}
    s = Solver(modes[mode], tol, max_iters, check_every, stall_checks, nn_refresh)
    check(load().aq_flip_set_solver(flip.h, C.byref(s), net.h if net is not None else None))


def solver_stats(flip, reset: bool = False) -> dict:
    st = SolverStats()
    check(load().aq_flip_solver_stats_get(flip.h, C.byref(st), 1 if reset else 0))
    return {k: getattr(st, k) for k, _ in SolverStats._fields_}


def mfma_probe(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    D = np.empty((16, 16), np.float32)
    check(load().aq_unet_mfma_probe(_fp(A), _fp(B), _fp(D)))
    return D
