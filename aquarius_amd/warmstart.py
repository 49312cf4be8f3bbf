"""Learned pressure warm start: training data, training and evaluation.

Training pairs come from the simulation: randomized ball-drop scenes (1-3
spheres of radius U[0.1, 0.25] above a pool of depth U[0.1, 0.4], seeded) are
stepped with the fixed-sweep solver, and at chosen steps the step's pressure
system is solved to convergence -- by default with plain FP32 Jacobi to a tight
FP64 residual, or with the MG-preconditioned CG solver (`--mgpcg`, FP64, to a 1e-10
CG residual); a pair is accepted only if the FP64 residual of its stored FP32
pressure is <= 1e-4;
(right-hand side, FLUID mask, converged pressure) is captured on the device.

    python -m aquarius_amd.warmstart train --pairs 4000 --out aquarius_amd/weights/unet_warmstart_64.npz
    python -m aquarius_amd.warmstart compare --weights aquarius_amd/weights/unet_warmstart_64.npz --n 128
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import configs
from ._abi import AQ_SHAPE_BOX, AQ_SHAPE_SPHERE
from .fluid import FlipHandle
from .unet import UNet

WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_warmstart_64.npz"


def random_scene(rng: np.random.Generator, n: int = 64, seed: int = 0):
    """Pool of depth U[0.1, 0.4] plus 1-3 spheres of radius U[0.1, 0.25] above it."""
    depth = float(rng.uniform(0.1, 0.4))
    k = int(rng.integers(1, 4))
    cfg = configs.make_config(n, dt=0.005, pressure_iters=100, diffusion_iters=0, pool_depth=depth,
                              seed=seed)
    vols = [dict(shape=AQ_SHAPE_BOX, box_min=(0, 0, 0), box_max=(1, depth, 1), phase=0)]
    for _ in range(k):
        r = float(rng.uniform(0.1, 0.25))
        lo_y = min(depth + r, 1 - r)
        c = (float(rng.uniform(r, 1 - r)), float(rng.uniform(lo_y, 1 - r)), float(rng.uniform(r, 1 - r)))
        vols.append(dict(shape=AQ_SHAPE_SPHERE, center=c, radius=r, phase=0))
    configs._set_volumes(cfg, vols)
    return cfg


def converge(h: FlipHandle, tol: float, cap: int, check_every: int = 100, start=None,
             stall_sweeps: int = 0):
    """Re-solve the handle's current pressure system with plain Jacobi, from zero or
    from `start` (a pressure field of the same extent). With stall_sweeps > 0 the solve
    also ends when that many sweeps improve the residual by less than 1 % (FP32
    Jacobi's round-off floor above tol)."""
    h.set_field("pressure", np.zeros(h.field_shape("pressure"), np.float32) if start is None else start)
    if stall_sweeps <= 0:
        return h.solve_pressure(tol, cap, check_every)
    total, best, flat = 0, np.inf, 0
    while True:
        it, rel, ok = h.solve_pressure(tol, min(stall_sweeps, cap - total), check_every)
        total += it
        if ok or total >= cap or not np.isfinite(rel):
            return total, rel, ok
        if rel < 0.99 * best:
            best, flat = rel, 0
        else:
            flat += 1
            if flat >= 2:
                return total, rel, False


def add_residual_samples(net: UNet, h: FlipHandle, p_star: np.ndarray, sweeps) -> None:
    """Samples of the residual equation the network's corrections solve. From zero, k
    Jacobi sweeps give p_k; with clamp-to-edge neighbours the residual is
    r = b - A p_k = 6 (J(p_k) - p_k) on FLUID cells (one more sweep gives J(p_k)), and
    the target is the error left, e = p* - p_k. The handle's right-hand side is set to r
    through its divergence field (rhs = -div * density) and its pressure to a multiple
    of e for the sample, then both are restored. The guess and the corrections scale the
    network's output by the energy-norm alpha, so only the error's shape matters; the
    multiple gives the sample's target (pressure / RMS of the right-hand side) the RMS
    the system's own sample has (the error left by a smooth residual is thousands of
    times its RMS, which would otherwise swamp the loss)."""
    X, Y, Z = h.field_shape("pressure")
    fluid = h.get_field("content")[:X, :Y, :Z] == 1
    div0 = h.get_field("divergence")
    dens = h.get_field("density")
    zero = np.zeros((X, Y, Z), np.float32)
    safe = np.where(fluid, dens, 1.0).astype(np.float64)

    def rms(a):
        return float(np.sqrt(np.mean(np.square(a[fluid], dtype=np.float64))))

    ratio = rms(p_star) / max(rms(-div0.astype(np.float64) * dens), 1e-30)  # the system's |p|/|b|
    for k in sweeps:
        h.set_field("pressure", zero)
        h.solve_pressure(0.0, k, k)
        pk = h.get_field("pressure").astype(np.float64)
        h.solve_pressure(0.0, 1, 1)
        r = np.where(fluid, 6.0 * (h.get_field("pressure") - pk), 0.0)
        e = np.where(fluid, p_star - pk, 0.0)
        c = ratio * rms(r) / max(rms(e), 1e-30)
        h.set_field("divergence", np.where(fluid, -r / safe, div0).astype(np.float32))
        h.set_field("pressure", (c * e).astype(np.float32))
        net.add_sample(h)
    h.set_field("divergence", div0)
    h.set_field("pressure", p_star)


def generate(net: UNet, pairs: int, n: int = 64, seed: int = 0, per_scene: int = 4,
             max_step: int = 60, tol: float = 1e-5, accept: float = 1e-4, cap: int = 60000,
             log=print, solver: str = "jacobi", stall_sweeps: int = 0, residual_sweeps=()):
    """Fill the network's device dataset with `pairs` converged systems. Within a
    scene each system is solved starting from the previous stop's converged pressure
    (the fixed point does not depend on the start, only the sweep count does). With
    stall_sweeps > 0 a Jacobi solve that stalls at its round-off floor above tol ends
    there (and is kept if its residual is <= accept). With residual_sweeps, each system
    is followed by one residual-equation sample per entry (add_residual_samples), so
    the dataset holds pairs * (1 + len(residual_sweeps)) samples, a system's together."""
    per = 1 + len(residual_sweeps)
    pairs *= per
    net.reserve(pairs, (n, n, n))
    rng = np.random.default_rng(seed)
    scene = 0
    t0 = time.time()
    sweeps = 0
    while net.dataset_size < pairs:
        cfg = random_scene(rng, n, seed=int(rng.integers(2**31)))
        h = FlipHandle(cfg)
        stops = sorted(set(int(s) for s in rng.integers(1, max_step + 1, per_scene)))
        done = 0
        prev = None
        for s in stops:
            h.steps(s - done)
            done = s
            if solver == "mgpcg":
                it, rel, _ = h.solve_pressure_mgpcg(1e-10, max_iters=300)
            else:
                it, rel, _ = converge(h, tol, cap, start=prev, stall_sweeps=stall_sweeps)
            sweeps += it
            if rel <= accept and net.dataset_size < pairs:
                net.add_sample(h)
                prev = h.get_field("pressure") if solver != "mgpcg" else None
                if residual_sweeps:
                    add_residual_samples(net, h, h.get_field("pressure"), residual_sweeps)
        h.close()
        scene += 1
        if scene % 50 == 0:
            log(f"  scenes {scene}, pairs {net.dataset_size}/{pairs}, mean {solver} iterations "
                f"{sweeps / max(1, net.dataset_size):.0f}, {time.time() - t0:.0f} s")
    return dict(scenes=scene, pairs=net.dataset_size, seconds=time.time() - t0, solver=solver,
                mean_iterations=sweeps / max(1, net.dataset_size))


def solve_with_corrections(h: FlipHandle, net: UNet, tol: float, cap: int, refresh: int,
                           check_every: int = 10):
    """From the U-Net guess: Jacobi with a network correction of the iterate every
    `refresh` sweeps -- the library's AQ_SOLVE_TOL_NN controller
    (aq_flip_solve_pressure_nn). Returns (sweeps, residual, converged, corrections)."""
    from .unet import solve_pressure_nn
    it, rel, ok, nc, _ = solve_pressure_nn(h, net, tol, cap, check_every, refresh)
    return it, rel, ok, nc


def compare_on(h: FlipHandle, net: UNet, tol: float = 1e-4, cap: int = 200000, check_every: int = 10,
               refresh: int | None = None):
    """Sweeps to `tol` for the handle's current system: from zero, from the U-Net
    guess, and (with `refresh`) from the guess with a network correction every
    `refresh` sweeps."""
    from .unet import nn_guess
    it0, r0, ok0 = converge(h, tol, cap, check_every)
    nn_guess(h, net)
    g = h.residual()
    it1, r1, ok1 = h.solve_pressure(tol, cap, check_every)
    out = dict(plain_iters=it0, plain_residual=r0, plain_ok=ok0, nn_guess_residual=g,
               nn_iters=it1, nn_residual=r1, nn_ok=ok1)
    if refresh:
        it2, r2, ok2, nc = solve_with_corrections(h, net, tol, cap, refresh, check_every)
        out.update(corr_iters=it2, corr_residual=r2, corr_ok=ok2, corrections=nc, refresh=refresh)
    return out


def evaluate(net: UNet, n: int = 64, scenes: int = 8, seed: int = 12345, steps: int = 30,
             tol: float = 1e-4, log=print, refresh: int | None = None):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(scenes):
        cfg = random_scene(rng, n, seed=int(rng.integers(2**31)))
        h = FlipHandle(cfg)
        h.steps(int(rng.integers(1, steps + 1)))
        rows.append(compare_on(h, net, tol, refresh=refresh))
        h.close()
        log("  ", rows[-1])
    p = np.array([r["plain_iters"] for r in rows], float)
    q = np.array([r["nn_iters"] for r in rows], float)
    out = dict(rows=rows, plain_mean=float(p.mean()), nn_mean=float(q.mean()),
               ratio=float(p.sum() / max(q.sum(), 1)))
    if refresh:
        c = np.array([r["corr_iters"] for r in rows], float)
        out.update(corr_mean=float(c.mean()), corr_ratio=float(p.sum() / max(c.sum(), 1)))
    return out


def evaluate_bench(net: UNet, stops=(1, 10, 30), tol: float = 1e-4, log=print, refresh: int | None = None,
                   cap: int = 60000):
    """Sweeps to `tol` on the 128^3 multiphase scene after `stops` steps."""
    h = FlipHandle(configs.multiphase_128())
    rows = []
    done = 0
    for s in stops:
        h.steps(s - done)
        done = s
        rows.append(dict(step=s, **compare_on(h, net, tol, cap=cap, check_every=50, refresh=refresh)))
        log("  ", rows[-1])
    h.close()
    p = sum(r["plain_iters"] for r in rows)
    q = sum(r["nn_iters"] for r in rows)
    out = dict(rows=rows, ratio=p / max(q, 1))
    if refresh:
        out["corr_ratio"] = p / max(sum(r["corr_iters"] for r in rows), 1)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=4000)
    t.add_argument("--held-out", type=int, default=64)
    t.add_argument("--n", type=int, default=64)
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--schedule", choices=("constant", "cosine"), default="constant",
                   help="Adam step size: constant, or cosine decay from --lr to --lr-min")
    t.add_argument("--lr-min", type=float, default=1e-5)
    t.add_argument("--keep-best", action="store_true",
                   help="keep the parameters with the lowest held-out loss seen (checked every 250 steps)")
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", type=Path, default=WEIGHTS)
    t.add_argument("--eval-scenes", type=int, default=8)
    t.add_argument("--tol", type=float, default=1e-5)
    t.add_argument("--eval-128", action="store_true")
    t.add_argument("--mgpcg", action="store_true", help="converge the training systems with MG-PCG")
    t.add_argument("--per-scene", type=int, default=4, help="systems captured per random scene")
    t.add_argument("--stall-sweeps", type=int, default=0,
                   help="end a Jacobi solve after this many sweeps without a 1%% improvement (0: never)")
    t.add_argument("--residual-sweeps", default="",
                   help="comma-separated Jacobi sweep counts k: each system also gives a sample of the "
                        "residual equation after k sweeps from zero (what corrections solve)")
    c = sub.add_parser("compare")
    c.add_argument("--weights", type=Path, default=WEIGHTS)
    c.add_argument("--n", type=int, default=64)
    c.add_argument("--scenes", type=int, default=4)
    c.add_argument("--bench", action="store_true")
    c.add_argument("--refresh", type=int, default=0)
    a = ap.parse_args(argv)

    if a.cmd == "train":
        net = UNet(lr=a.lr, seed=a.seed + 1)
        total = a.pairs + a.held_out
        rs = tuple(int(k) for k in a.residual_sweeps.split(",") if k)
        print(f"[warmstart] generating {total} systems at {a.n}^3"
              + (f", each with residual samples after {list(rs)} sweeps" if rs else ""), flush=True)
        gen = generate(net, total, n=a.n, seed=a.seed, tol=a.tol, solver="mgpcg" if a.mgpcg else "jacobi",
                       per_scene=a.per_scene, stall_sweeps=a.stall_sweeps, residual_sweeps=rs)
        print(f"[warmstart] data: {json.dumps(gen)}", flush=True)
        per = 1 + len(rs)
        train_n, held = a.pairs * per, a.held_out * per
        l0 = net.eval(train_n, held)
        t0 = time.time()
        hist = []
        chunk = 250
        best = (np.inf, None, 0)
        for s in range(0, a.steps, chunk):
            k = min(chunk, a.steps - s)
            if a.schedule == "cosine":  # from lr down to lr_min over the run, per chunk
                net.set_lr(a.lr_min + 0.5 * (a.lr - a.lr_min) * (1 + np.cos(np.pi * s / a.steps)))
            h = net.train(k, a.batch, train_n, seed=a.seed * 7919 + s)
            hist.append(float(h.mean()))
            ho = net.eval(train_n, held)
            if a.keep_best and ho < best[0]:
                best = (ho, net.get_params(), s + k)
            print(f"[warmstart] step {s + k}: lr {net.cfg.lr:.2e} train loss {hist[-1]:.4e} "
                  f"held-out {ho:.4e} ({time.time() - t0:.0f} s)", flush=True)
        if a.keep_best and best[1] is not None:
            net.set_params(best[1])
            print(f"[warmstart] kept the parameters of step {best[2]} (held-out {best[0]:.4e})", flush=True)
        l1 = net.eval(train_n, held)
        a.out.parent.mkdir(parents=True, exist_ok=True)
        net.save(a.out)
        print(f"[warmstart] held-out loss {l0:.4e} -> {l1:.4e}; saved {a.out}", flush=True)
        ev = evaluate(net, n=a.n, scenes=a.eval_scenes)
        print(f"[warmstart] sweeps to 1e-4: plain {ev['plain_mean']:.0f}, warm {ev['nn_mean']:.0f}, "
              f"ratio {ev['ratio']:.2f}", flush=True)
        if a.eval_128:
            eb = evaluate_bench(net)
            print(f"[warmstart] 128^3 multiphase: sweep ratio {eb['ratio']:.2f}", flush=True)
        return 0
    if a.cmd == "compare":
        net = UNet()
        net.load_file(a.weights)
        r = a.refresh or None
        ev = (evaluate_bench(net, refresh=r) if a.bench
              else evaluate(net, n=a.n, scenes=a.scenes, refresh=r))
        print(json.dumps({k: v for k, v in ev.items() if k != "rows"}), flush=True)
        return 0
    return 1


if __name__ == "__main__":
    sys.exit(main())
