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


def test_graph_replayed_tolerance_solves_match_eager(monkeypatch):
    """The tolerance solve replays captured graphs of its sweep chunks (one per buffer
    state and chunk length); every replay must give what eager launches give."""
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("AQFLIP_NO_GRAPH", flag)
        h = FlipHandle(configs.ball_drop_64())
        attach_solver(h, "tol_zero", tol=1e-4, max_iters=50000, check_every=30)
        h.steps(3)
        st = solver_stats(h)
        it, rel, ok = h.solve_pressure(1e-5, 4000, 25)  # direct solves replay graphs too
        out[flag] = (st["sweeps"], st["last_residual"], it, rel, h.get_field("pressure"))
        h.close()
    assert out["0"][:4] == out["1"][:4]
    assert np.array_equal(out["0"][4].view(np.uint32), out["1"][4].view(np.uint32))


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
    assert np.isfinite(st["last_guess_residual"])
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


def test_coarse_inference_path():
    """A grid larger than the network's resolution: the right-hand side is
    restricted for the network and the guess prolonged back (infer_extent 16 on a
    32^3 grid = one level); the guess is finite, zero on AIR, and the solve from
    it reaches the tolerance."""
    net = UNet(seed=6, infer_extent=16)
    h = FlipHandle(configs.make_config(32, dt=0.01, pressure_iters=20, seed=8))
    h.step()
    from aquarius_amd.unet import nn_guess
    nn_guess(h, net)
    g = h.residual()
    p = h.get_field("pressure")
    content = h.get_field("content")[:32, :32, :32]
    print(f"  coarse guess residual {g:.4f}")
    assert np.isfinite(g) and np.all(np.isfinite(p))
    assert np.all(p[content != 1] == 0)
    it, rel, ok = h.solve_pressure(1e-4, 50000, check_every=20)
    assert ok and rel <= 1e-4
    h.close()
