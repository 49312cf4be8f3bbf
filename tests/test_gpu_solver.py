"""Pressure-solve modes of aq_flip_step on the GPU: to a tolerance from zero, and
from the U-Net guess with zero-guess Jacobi as the fallback. Failure handling is
forced from the host (a network whose parameters are NaN), never by faulting."""
import numpy as np
import pytest

from aquarius_amd import configs
from aquarius_amd.fluid import FlipHandle
from aquarius_amd.unet import UNet, attach_solver, solver_stats

pytestmark = pytest.mark.gpu


def test_tolerance_solve_from_zero():
    h = FlipHandle(configs.ball_drop_16())
    attach_solver(h, "tol_zero", tol=1e-4, max_iters=50000, check_every=20)
    h.steps(3)
    st = solver_stats(h)
    print(st)
    assert st["solves"] == 3 and st["capped"] == 0 and st["fallbacks"] == 0
    assert st["last_residual"] <= 1e-4
    assert h.residual() <= 1e-4
    h.close()


def test_nan_network_falls_back_to_zero_guess():
    net = UNet(seed=2)
    net.set_params(np.full(net.param_count, np.nan, np.float32))
    h = FlipHandle(configs.ball_drop_16())
    attach_solver(h, "tol_nn", tol=1e-4, max_iters=50000, check_every=20, net=net)
    h.steps(2)
    st = solver_stats(h)
    print(st)
    assert st["solves"] == 2 and st["fallbacks"] == 2
    assert st["last_fallback"] == 1 and not np.isfinite(st["last_guess_residual"])
    assert st["last_residual"] <= 1e-4 and h.residual() <= 1e-4
    h.close()


def test_warm_start_reaches_tolerance_and_caps_are_reported():
    net = UNet(seed=4)
    h = FlipHandle(configs.ball_drop_16())
    attach_solver(h, "tol_nn", tol=1e-4, max_iters=50000, check_every=20, net=net)
    h.steps(2)
    st = solver_stats(h, reset=True)
    print(st)
    assert st["solves"] == 2 and st["last_residual"] <= 1e-4
    # an unreachable tolerance with a small cap is reported, and the step still completes
    attach_solver(h, "tol_zero", tol=1e-14, max_iters=40, check_every=20)
    h.step()
    st = solver_stats(h)
    assert st["capped"] == 1 and st["last_sweeps"] == 40
    h.close()


def test_grid_extent_must_suit_the_network():
    net = UNet(seed=1)
    h = FlipHandle(configs.make_config(0, dt=0.01, pressure_iters=10, extent=(12, 16, 18), seed=3))
    with pytest.raises(Exception, match="multiples of 4"):
        attach_solver(h, "tol_nn", net=net)
    h.close()
