"""The trained warm start (aquarius_amd/weights/unet_warmstart_64.npz, produced by
`python -m aquarius_amd.warmstart train`) against plain Jacobi from zero: both are
refined to the same FP64 relative residual (1e-4), and the warm start must need
fewer sweeps -- the gain the learned part is for. In both cases the measured FP64
residual, never a prediction, decides when the solve has converged."""
import numpy as np
import pytest

from aquarius_amd import configs, warmstart
from aquarius_amd.fluid import FlipHandle

pytestmark = pytest.mark.gpu

# Measured gains are recorded in the bench output; the bars here are clear wins.
MIN_RATIO_64 = 1.25            # initial guess only
MIN_RATIO_64_CORRECTED = 2.0   # guess plus a network correction every 200 sweeps


def trained_net():
    from aquarius_amd.unet import UNet
    if not warmstart.WEIGHTS.exists():
        pytest.skip("no trained weights in aquarius_amd/weights")
    net = UNet()
    net.load_file(warmstart.WEIGHTS)
    return net


def test_warm_start_needs_fewer_sweeps_64():
    net = trained_net()
    rng = np.random.default_rng(2024)  # held-out scenes (training uses seed 0)
    plain = warm = corr = 0
    for i in range(3):
        cfg = warmstart.random_scene(rng, 64, seed=int(rng.integers(2**31)))
        h = FlipHandle(cfg)
        h.steps(int(rng.integers(5, 30)))
        r = warmstart.compare_on(h, net, tol=1e-4, refresh=200)
        h.close()
        print(f"  scene {i}: plain {r['plain_iters']} sweeps, warm {r['nn_iters']} sweeps "
              f"(guess residual {r['nn_guess_residual']:.3f}), corrected {r['corr_iters']} sweeps "
              f"({r['corrections']} corrections)", flush=True)
        assert r["plain_ok"] and r["nn_ok"] and r["corr_ok"]
        assert max(r["plain_residual"], r["nn_residual"], r["corr_residual"]) <= 1e-4
        plain += r["plain_iters"]
        warm += r["nn_iters"]
        corr += r["corr_iters"]
    print(f"  total: plain {plain}, warm {warm} ({plain / warm:.2f}x), corrected {corr} "
          f"({plain / corr:.2f}x)", flush=True)
    assert plain >= MIN_RATIO_64 * warm
    assert plain >= MIN_RATIO_64_CORRECTED * corr


def test_warm_start_solver_mode_reaches_tolerance_128():
    """The step-level controller on the 128^3 multiphase workload (warm start with
    a network correction every 200 sweeps): every step's pressure reaches 1e-4, and
    the warm start spends at least 2x fewer sweeps than zero."""
    from aquarius_amd.unet import attach_solver, solver_stats
    net = trained_net()
    cfg = configs.multiphase_128()
    sweeps = {}
    for mode in ("tol_zero", "tol_nn"):
        h = FlipHandle(cfg)
        nn = mode == "tol_nn"
        attach_solver(h, mode, tol=1e-4, max_iters=100000, check_every=50, net=net if nn else None,
                      nn_refresh=200 if nn else 0)
        h.steps(3)
        st = solver_stats(h)
        h.close()
        print(f"  {mode}: {st}", flush=True)
        assert st["solves"] == 3 and st["capped"] == 0
        assert st["last_residual"] <= 1e-4
        sweeps[mode] = st["sweeps"]
    assert 2 * sweeps["tol_nn"] <= sweeps["tol_zero"]


def test_corrections_near_the_round_off_floor_are_undone_128():
    """At the FP32 round-off floor of Jacobi (an unreachable 1e-7 on the 128^3 system of
    step 30, stall detection off, 30,000 sweeps, a correction every 200): once a chunk of
    sweeps ends within 1 % of the previous one, the next correction is judged; one that
    the following sweeps leave above the residual from before it is undone and the
    corrections stop, so far fewer than the 149 scheduled ones are applied. The solve
    ends at the floor: FP32 Jacobi settles on an iterate its sweeps no longer change, and
    which one depends on the path (observed 1.22e-4 from zero, 1.33e-4 and 1.48e-4 after
    corrections), so the bar is 30 % above plain Jacobi's floor."""
    from aquarius_amd.unet import attach_solver, solve_pressure_nn
    net = trained_net()
    cfg = configs.multiphase_128()
    h = FlipHandle(cfg)
    h.steps(30)
    attach_solver(h, "fixed", stall_checks=10**6)  # the pressure solve of the steps is unchanged
    zero = np.zeros(h.field_shape("pressure"), np.float32)
    h.set_field("pressure", zero)
    _, plain_floor, _ = h.solve_pressure(1e-7, 30000, check_every=100)
    it, res, ok, nc, nr = solve_pressure_nn(h, net, 1e-7, 30000, check_every=100, nn_refresh=200)
    print(f"  plain floor {plain_floor:.4e}; corrected {res:.4e} after {it} sweeps, {nc} corrections, "
          f"{nr} undone", flush=True)
    h.close()
    assert not ok and it == 30000 and nr == 1 and nc <= 40
    assert res <= 1.3 * plain_floor


def test_residual_equation_samples_16():
    """The residual samples the corrections train on: after k sweeps from zero,
    6 (J(p_k) - p_k) on FLUID cells is the residual b - A p_k (its norm over ||b|| is the
    library's FP64 relative residual at that point), and a sample holds r as its
    right-hand side and p* - p_k as its target; the handle is left as it was."""
    from aquarius_amd.unet import UNet
    cfg = configs.ball_drop_16()
    h = FlipHandle(cfg)
    h.steps(3)
    X, Y, Z = h.field_shape("pressure")
    fluid = h.get_field("content")[:X, :Y, :Z] == 1
    dens = h.get_field("density")
    b = -h.get_field("divergence").astype(np.float64) * dens
    warmstart.converge(h, 1e-6, 20000)
    p_star = h.get_field("pressure")
    div0 = h.get_field("divergence")
    k = 40
    h.set_field("pressure", np.zeros((X, Y, Z), np.float32))
    _, rel_k, _ = h.solve_pressure(0.0, k, k)
    pk = h.get_field("pressure").astype(np.float64)
    h.solve_pressure(0.0, 1, 1)
    r = np.where(fluid, 6.0 * (h.get_field("pressure") - pk), 0.0)
    rel_host = np.linalg.norm(r[fluid]) / np.linalg.norm(b[fluid])
    print(f"  relative residual after {k} sweeps: library {rel_k:.6e}, host {rel_host:.6e}")
    assert abs(rel_host - rel_k) <= 1e-5 * rel_k
    # the sample: right-hand side r, target p* - p_k; then the handle is restored
    net = UNet()
    net.reserve(2, (X, Y, Z))
    h.set_field("pressure", p_star)
    warmstart.add_residual_samples(net, h, p_star, (k,))
    assert net.dataset_size == 1
    np.testing.assert_array_equal(h.get_field("divergence"), div0)
    np.testing.assert_array_equal(h.get_field("pressure"), p_star)
    # re-create the sample's system by hand: its converged solution is p* - p_k
    safe = np.where(fluid, dens, 1.0)
    h.set_field("divergence", np.where(fluid, -r / safe, div0).astype(np.float32))
    warmstart.converge(h, 1e-6, 20000)
    e = h.get_field("pressure").astype(np.float64)
    err = np.linalg.norm((e - (p_star - pk))[fluid]) / np.linalg.norm((p_star - pk)[fluid])
    print(f"  solution of the residual equation against p* - p_k: {err:.3e}")
    assert err < 1e-4  # observed 1.1e-6; a wrong sign or scale of r gives O(1)
    h.close()
