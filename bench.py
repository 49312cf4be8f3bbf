#!/usr/bin/env python3
"""FLIP simulation-step benchmark on one MI355X.

Workload: the multiphase ball drop on a 128^3 MAC grid with 8 particles per cell
(about 4.43M particles, 2 phases), 150 Jacobi sweeps from a zero guess and 100
diffusion sweeps per step, dt 0.005 -- the configuration the headline metric
"FLIP steps/sec at 128^3, 8 ppc" is quoted on. A step is one full
Fluid_3D_FLIP::Fluid::simulationStep over device-resident state (synthetic,
seeded particles; nothing crosses PCIe inside the timed region).

Prints one JSON line with the contract fields plus:
  roofline      the dominant kernel's achieved algorithmic bandwidth against the
                8 TB/s HBM peak (per-launch duration from HIP events recorded on
                the library's own stream over the timed region)
  kernels       the same figure for every stage of the step
  jacobi_to_tol plain-Jacobi sweeps from zero to a 1e-4 relative FP64 residual on
                one fixed system -- the pressure system of step 110 of a fresh
                handle of the workload, whatever --steps is -- and (with the trained
                U-Net weights present) the sweeps from the U-Net guess alone and from
                the guess with a correction every 200 sweeps, with the corrections'
                cost in sweep-equivalents; MG-PCG on the same system
  tolerance_mode steps/s when every step solves to 1e-4: plain Jacobi from zero
                against the U-Net guess alone, the guess with corrections, and
                MG-PCG, over steps 2-11 (ball in free fall) and again over steps
                162-171 (after the ball has met the pool)
  fp64_error    step 111 of the workload on the GPU (FP32) against the FP64 CPU oracle
                stepping once from the GPU's state after step 110: relative L2 of
                pressure, face velocities, fractions and particle positions,
                velocities and fractions
  paired_diffusion  diffusion sweeps paired with pressure sweeps per step, how many of
                them computed their diffusion part in the profiled pass (a sweep after
                one that changed no bit skips the identical work), and the steps/s of
                the timed region with that skipping switched off
  cpu_baseline  the CPU oracle timed on a bounded sample of the same workload: after
                2 warm-up steps, the median of 5 steps on all the threads the box
                gives this process and the median of 5 on one thread, with per-stage
                milliseconds, the CPU model and nproc

--timed-only runs the warm-up, the timed region and the profiled pass and nothing
else (the form a rocprofv3 kernel trace of the timed configuration is taken with).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "FLIP steps/sec at 128^3, 8 ppc; Jacobi iters to 1e-4 residual, plain vs NN-warm"
# Sweeps between U-Net corrections of the iterate in the warm-start solves. A correction
# (inference through 64^3 and the energy-norm scaling) costs about as much as 280
# sweeps at 128^3, so fewer corrections trade sweeps for time. Measured in tolerance
# mode: every 100 / 150 / 200 / 300 / 400 / 600 sweeps gave 30.4 / 33.4 / 35.9 / 40.3 /
# 40.2 / 41.0 steps/s with 2720 / 3110 / 3300 / 3410 / 3750 / 4150 sweeps per step.
# The sweep-count comparison (jacobi_to_tol) corrects every 200 sweeps, the timed
# tolerance-mode run every 300.
NN_REFRESH = 200
NN_REFRESH_TIMED = 300


SWEEP_SPAN = 512  # cells per sweep chunk (the unit the listed sweeps skip or visit)


def swept_cells(content, size) -> int:
    """Cells a listed sweep visits: every cell of each 512-cell chunk (in the
    z-fastest interior order) that holds a FLUID cell."""
    import numpy as np
    X, Y, Z = size
    fluid = (content[:X, :Y, :Z] == 1).reshape(-1)
    pad = (-fluid.size) % SWEEP_SPAN
    chunks = np.concatenate([fluid, np.zeros(pad, bool)]).reshape(-1, SWEEP_SPAN).any(1)
    return int(min(chunks.sum() * SWEEP_SPAN, fluid.size))


def near_row_faces(content, size) -> int:
    """Faces the second and third extrapolation passes of a step visit: every face of
    each face row (x <= X, y <= Y) within 4 rows in x and in y of a cell row holding a
    FLUID cell (the rows P2G stamps)."""
    import numpy as np
    X, Y, Z = size
    rows = (content[:X, :Y, :Z] == 1).any(2)
    near = np.zeros((X + 1, Y + 1), bool)
    for x, y in zip(*rows.nonzero()):
        near[max(0, x - 4):x + 5, max(0, y - 4):y + 5] = True
    return int(near.sum()) * (Z + 1)


def algorithmic_bytes(cfg, n_particles: int, K: int, active_cells: float, sweeps: dict,
                      diffusion_computed: float | None = None,
                      near_faces: float | None = None) -> dict:
    """Bytes each stage must move once per launch (every operand read or written
    exactly once), per the data layout of the library. `sweeps` holds the launches
    per step of the sweep stages. Sweeps count the cells a launch visits: all C
    cells for the first two pressure sweeps of a step (both pressure buffers are
    rewritten in full after the FLUID set changes) and for the first diffusion sweep
    (it copies AIR cells across), the chunks holding FLUID otherwise. A paired sweep
    moves diffusion bytes only when it computed its diffusion part
    (`diffusion_computed` per step, the first sweep of a step always). Of a step's three
    extrapolation passes the first visits every face, the other two `near_faces` (all
    of them when None)."""
    X, Y, Z = cfg.size
    C = X * Y * Z
    Np = n_particles
    rec = 24 + 4 * K  # particle record bytes that matter (position, velocity, K fractions)
    paired = sweeps.get("pressure_diffusion", 0)
    alone = sweeps.get("pressure", 0)
    diff = sweeps.get("diffusion", 0)

    def per_launch(n, full):
        return (full * C + (n - full) * active_cells) / n if n else 0.0

    def paired_diffusion(n, computed):
        # cells whose diffusion a paired launch computes, on average: all C in the first
        # sweep of a step, the FLUID chunks in the others that were not skipped
        if not n:
            return 0.0
        computed = n if computed is None else max(1.0, min(float(n), computed))
        return (C + (computed - 1) * active_cells) / n

    return {
        # G2P + RK3 + hash + bin count: position and velocity read, position, velocity
        # and fractions written, hash and rank written; grid old/new face velocities
        # and fractions read once.
        "g2p": Np * (24 + 24 + 4 * K + 4 + 4) + C * (24 + 4 * K),
        # scan (counts in, offsets out), scatter (hash, rank in; permutation out),
        # segment sort (permutation checked), gather (permutation, record in and out)
        "transfer_to_grid": C * (4 + 4 + 4) + Np * (4 + 4 + 4) + Np * 4 + Np * (4 + 2 * rec),
        # P2G: sorted records and cell offsets in; velocity, new velocity, valid bits,
        # content, density and both fraction copies out
        "p2g": Np * rec + C * (4 + 12 + 12 + 1 + 1 + 4 + 8 * K),
        "gravity": C * (1 + 12 + 12 + 1),
        "divergence": C * (12 + 4 + 4 + 4 + 4),
        # per sweep: p 4 r + rhs 4 r + mask 1 r + p 4 w per visited cell
        "pressure": 13 * per_launch(alone, min(alone, max(0, 2 - paired))),
        "project": C * (1 + 4 + 4 + 12 + 12 + 1),
        # per pass: velocity and valid bits read and written per visited face
        "extrapolate": (12 + 1 + 12 + 1) * (C if near_faces is None else (C + 2 * min(near_faces, C)) / 3),
        "diffusion": (1 + 8 * K) * per_launch(diff, 1 if diff and not paired else 0),
        # one pressure sweep and one diffusion sweep in the same launch
        "pressure_diffusion": (13 * per_launch(paired, min(2, paired))
                               + (1 + 8 * K) * paired_diffusion(paired, diffusion_computed)),
    }


def whole_step(kernels: dict, ms_per_step: float) -> dict:
    """Algorithmic bytes of a whole step against its time: the step-level roofline."""
    per_step = sum(k["bytes_per_launch"] * k["launches_per_step"] for k in kernels.values())
    gbs = per_step / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_step": round(per_step), "achieved_GBs": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}


# Kernels of each stage, as rocprofv3 names them (template arguments stripped).
STAGE_KERNELS = {
    "g2p": ["aqf::k_g2p_advect_hash"],
    "transfer_to_grid": ["aqf::k_scan_reduce", "aqf::k_scan_final",
                         "aqf::k_scatter", "aqf::k_segsort", "aqf::k_gather"],
    "p2g": ["aqf::k_p2g"],
    "gravity": ["aqf::k_gravity_boundary"],
    "divergence": ["aqf::k_divergence"],
    "pressure": ["aqf::k_jacobi"],
    "project": ["aqf::k_project"],
    "extrapolate": ["aqf::k_extrapolate"],
    "diffusion": ["aqf::k_diffusion"],
    "pressure_diffusion": ["aqf::k_jacobi_diffusion"],
}


def measured_traffic(scene: str, stage: str):
    """HBM bytes per launch of `stage` from the committed FETCH_SIZE/WRITE_SIZE
    summary of this workload (copy-calibrated), or None when there is none."""
    path = ROOT / "profiles" / f"hbm_traffic_{scene}.md"
    if not path.exists():
        return None
    text = path.read_text()
    start = text.find("```json")
    if start < 0:
        return None
    data = json.loads(text[start + 7:text.find("```", start + 7)])
    # Instantiations of one kernel (template arguments) are averaged, weighted by
    # their dispatches; the different kernels of a stage add up.
    per_kernel = {}
    for name, row in data["kernels"].items():
        base = name.split("<")[0]
        if base in STAGE_KERNELS.get(stage, []):
            b, n = per_kernel.get(base, (0.0, 0))
            d = row["dispatches"]
            per_kernel[base] = (b + d * (row["fetch_bytes"] + row["write_bytes"]), n + d)
    if not per_kernel:
        return None
    return round(sum(b / n for b, n in per_kernel.values()))


def sweeps_paired_possible(cfg, K: int) -> bool:
    """Whether the step pairs pressure and diffusion sweeps (fixed-sweep solver, more
    than one phase, both chains non-empty)."""
    return K > 1 and cfg.pressure_iters > 0 and cfg.diffusion_iters > 0


def load_warmstart_net():
    """The trained warm-start U-Net shipped with the package, or None."""
    from aquarius_amd import warmstart
    if not warmstart.WEIGHTS.exists():
        return None
    from aquarius_amd.unet import UNet
    net = UNet()
    net.load_file(warmstart.WEIGHTS)
    return net


def tolerance_mode_steps(cfg, steps: int, cap: int, advance: int = 0, tol: float = 1e-4):
    """Steps/s with every step's pressure solved to a `tol` FP64 relative residual (or
    until FP32 Jacobi stalls at its round-off floor), one fresh handle per solver:
    plain Jacobi from zero; the U-Net guess alone; the guess with a network correction
    every NN_REFRESH_TIMED sweeps (a correction that does not lower the residual is
    undone and Jacobi continues alone); MG-preconditioned CG (FP64) to the same
    tolerance. `advance` fixed-sweep steps run first (graph replay, untimed)."""
    from aquarius_amd.fluid import FlipHandle
    from aquarius_amd.unet import attach_solver, solver_stats
    net = load_warmstart_net()
    modes = [("plain", "tol_zero", None, 0), ("mgpcg", "mgpcg", None, 0)]
    if net is not None:
        modes += [("nn_guess_only", "tol_nn", net, 0), ("nn", "tol_nn", net, NN_REFRESH_TIMED)]
    out = {"tol": tol, "steps": steps, "first_step": advance + 2, "check_every": 100,
           "nn_refresh": NN_REFRESH_TIMED}
    for label, mode, nnet, refresh in modes:
        h = FlipHandle(cfg)
        if advance:
            h.steps(advance)
        if mode == "mgpcg":
            attach_solver(h, mode, tol=tol, max_iters=500, check_every=1)
        else:
            attach_solver(h, mode, tol=tol, max_iters=cap, check_every=100, net=nnet, nn_refresh=refresh)
        h.steps(1)  # warm-up (first solve, lazy allocations)
        solver_stats(h, reset=True)
        t0 = time.perf_counter()
        h.steps(steps)
        secs = time.perf_counter() - t0
        st = solver_stats(h)
        h.close()
        row = {"steps_per_s": round(steps / secs, 3), "capped": st["capped"],
               "last_residual": st["last_residual"]}
        if mode == "mgpcg":
            row["mean_cg_iters"] = st["sweeps"] / max(st["solves"], 1)
        else:
            row.update({"mean_sweeps": st["sweeps"] / max(st["solves"], 1), "fallbacks": st["fallbacks"]})
        if nnet is not None:
            row.update({"nn_corrections_per_step": st["nn_corrections"] / max(st["solves"], 1),
                        "nn_rejected": st["nn_rejected"]})
        out[label] = row
        print(f"tolerance mode {label}: {row['steps_per_s']} steps/s", file=sys.stderr, flush=True)
    for label in ("nn_guess_only", "nn", "mgpcg"):
        if label in out:
            out[label]["speedup_vs_plain"] = round(out[label]["steps_per_s"] / out["plain"]["steps_per_s"], 3)
            if "mean_sweeps" in out[label]:
                out[label]["sweep_ratio_vs_plain"] = round(
                    out["plain"]["mean_sweeps"] / max(out[label]["mean_sweeps"], 1), 3)
    return out


def solve_comparison(cfg, at_step: int, cap: int) -> dict:
    """Sweeps to a 1e-4 FP64 relative residual on one fixed system: the pressure
    system after `at_step` steps of a fresh handle (fixed sweeps, graph replay)."""
    import numpy as np

    from aquarius_amd.fluid import FlipHandle
    h = FlipHandle(cfg)
    h.steps(at_step)
    zero = np.zeros(h.field_shape("pressure"), np.float32)
    h.set_field("pressure", zero)
    ts = time.perf_counter()
    it, rel, ok = h.solve_pressure(1e-4, cap, check_every=100)
    solve = {"system": f"pressure system of step {at_step} of a fresh handle", "tol": 1e-4,
             "plain_iters": it, "plain_residual": rel, "plain_converged": ok,
             "plain_seconds": round(time.perf_counter() - ts, 3), "cap": cap, "nn_warm_iters": None}
    net = load_warmstart_net()
    if net is not None:
        from aquarius_amd import warmstart
        from aquarius_amd.unet import nn_correct, nn_guess
        # guess only: one inference, then plain Jacobi
        ts = time.perf_counter()
        nn_guess(h, net)
        g = h.residual()
        it2, rel2, ok2 = h.solve_pressure(1e-4, cap, check_every=10)
        solve.update({"nn_warm_iters": it2, "nn_guess_residual": g, "nn_residual": rel2,
                      "nn_converged": ok2, "nn_seconds": round(time.perf_counter() - ts, 3),
                      "iter_ratio": round(it / max(it2, 1), 3)})
        # guess plus a network correction of the iterate every NN_REFRESH sweeps
        ts = time.perf_counter()
        it3, rel3, ok3, nc = warmstart.solve_with_corrections(h, net, 1e-4, cap, NN_REFRESH)
        solve.update({"nn_corrected_iters": it3, "nn_corrections": nc, "nn_corrected_residual": rel3,
                      "nn_corrected_converged": ok3,
                      "nn_corrected_seconds": round(time.perf_counter() - ts, 3),
                      "corrected_iter_ratio": round(it / max(it3, 1), 3), "nn_refresh": NN_REFRESH})
        # what a correction costs, in sweeps of this system: 10 corrections and 1000
        # sweeps timed on the host around synchronised calls
        h.set_field("pressure", zero)
        h.solve_pressure(0.0, 100, check_every=100)
        ts = time.perf_counter()
        for _ in range(10):
            nn_correct(h, net)
        t_corr = (time.perf_counter() - ts) / 10
        ts = time.perf_counter()
        h.solve_pressure(0.0, 1000, check_every=1000)
        t_sweep = (time.perf_counter() - ts) / 1000
        sweq = t_corr / t_sweep
        solve.update({"correction_ms": round(1e3 * t_corr, 4), "sweep_us": round(1e6 * t_sweep, 3),
                      "correction_sweep_equivalents": round(sweq, 1),
                      "guess_only_work_ratio": round(it / (it2 + sweq), 3),
                      "corrected_work_ratio": round(it / (it3 + (nc + 1) * sweq), 3)})
    # the accuracy reference: MG-preconditioned CG (FP64) from zero on the same system
    ts = time.perf_counter()
    itm, relm, okm = h.solve_pressure_mgpcg(1e-9, max_iters=200)
    solve.update({"mgpcg_iters": itm, "mgpcg_residual_of_fp32_field": relm, "mgpcg_converged": okm,
                  "mgpcg_seconds": round(time.perf_counter() - ts, 4)})
    h.close()
    return solve


# FLOP of one U-Net forward pass per voxel of its finest level: 2 * 27 * cin * cout per
# 3x3x3 convolution at each level (the finest level's voxels, 1/8 and 1/64 of them
# below), plus the 1x1x1 head.
UNET_LEVEL_CONVS = {1: [(8, 16), (16, 16), (48, 16), (16, 16)], 2: [(16, 32), (32, 32), (96, 32), (32, 32)],
                    3: [(32, 64), (64, 64)]}


def unet_flop_per_voxel() -> float:
    f = 0.0
    for lvl, convs in UNET_LEVEL_CONVS.items():
        f += sum(2 * 27 * ci * co for ci, co in convs) / 8 ** (lvl - 1)
    return f + 2 * 16


def unet_inference(cfg, at_step: int, reps: int = 50) -> dict:
    """U-Net inference on the pressure system of step `at_step` of a fresh handle: the
    guess (restriction to the network's 64^3 resolution, forward pass, prolongation,
    energy-norm scaling), timed over `reps` synchronised calls, with the FLOP rate of
    the forward pass's convolutions against the BF16 dense MFMA peak."""
    from aquarius_amd.fluid import FlipHandle
    from aquarius_amd.unet import nn_guess
    net = load_warmstart_net()
    if net is None:
        return {"unet": None}
    h = FlipHandle(cfg)
    h.steps(at_step)
    for _ in range(3):
        nn_guess(h, net)
    t0 = time.perf_counter()
    for _ in range(reps):
        nn_guess(h, net)
    ms = 1e3 * (time.perf_counter() - t0) / reps
    h.close()
    X, Y, Z = cfg.size
    n = max(X, Y, Z)
    vox = X * Y * Z
    while n > 64:
        n //= 2
        vox //= 8
    gflop = unet_flop_per_voxel() * vox / 1e9
    return {"unet_guess_ms": round(ms, 4), "network_voxels": vox, "forward_gflop": round(gflop, 2),
            "tflops_whole_guess": round(gflop / ms, 2), "peak_bf16_tflops": 2500.0,
            "frac_of_peak_whole_guess": round(gflop / ms / 2500.0, 4)}


STATE_FIELDS = ("content", "velocity", "new_velocity", "pressure", "divergence", "density", "particle_count",
                "has_velocity", "volume_fractions", "new_volume_fractions")


def fp64_error(cfg, at_step: int) -> dict:
    """The FP32 GPU step against an FP64 run of the same step: a fresh handle runs
    `at_step` steps, its whole state (particles in their order and every grid field)
    is handed to the FP64 CPU oracle, then both run one step. Relative L2 differences;
    particles are matched through the sort permutations (both start from the same
    order)."""
    import numpy as np

    from aquarius_amd.fluid import FlipHandle
    from oracle import binding as ob
    ob.set_threads(cpu_threads())
    g = FlipHandle(cfg)
    g.steps(at_step)
    pos, vel, frac = g.get_particles()
    o = ob.Oracle(cfg, fp64=True, pos=pos, vel=vel, frac4=frac, init=False)
    for f in STATE_FIELDS:
        o.set_field(f, g.get_field(f))
    g.step()
    o.step()

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out = {"reference": (f"CPU oracle in FP64 (the same kernels' logic) stepping once from the GPU state "
                         f"after step {at_step}")}
    for f in ("pressure", "new_velocity", "velocity", "new_volume_fractions"):
        out[f] = rel(g.get_field(f), o.get_field(f))
    gp, gv, gf = g.get_particles()
    op, ov, of = o.get_particles()
    _, _, gh, pg = g.get_cells()
    _, _, oh, po = o.get_cells()
    ig, io = np.empty_like(pg), np.empty_like(po)
    ig[pg] = np.arange(pg.size)
    io[po] = np.arange(po.size)
    out["particle_position"] = rel(gp[ig], op[io])
    out["particle_velocity"] = rel(gv[ig], ov[io])
    out["particle_fractions"] = rel(gf[ig], of[io])
    # particles that the FP32 and FP64 positions put in different cells
    out["particles_in_other_cells"] = int((gh[ig] != oh[io]).sum())
    g.close()
    o.close()
    return {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in out.items()}


def cpu_threads() -> int:
    """Threads the box gives this process (OMP_NUM_THREADS where set, else up to 16)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else min(16, os.cpu_count() or 1)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene: str, steps: int, warmup: int = 2) -> dict:
    """The CPU oracle (OpenMP) on the same workload: `warmup` steps, then the median of
    `steps` steps on all the process's threads and the median of `steps` on one
    thread, each step timed stage by stage with std::chrono::steady_clock."""
    import statistics

    from aquarius_amd import configs
    from oracle import binding as ob
    threads = cpu_threads()
    cfg = configs.SCENES[scene]()
    ob.set_threads(threads)
    o = ob.Oracle(cfg)
    for _ in range(warmup):
        o.step()

    def run(nthreads):
        ob.set_threads(nthreads)
        totals, stages = [], []
        for _ in range(steps):
            t, st = o.time_step_stages()
            totals.append(t)
            stages.append(st)
        med = statistics.median(totals)
        per_stage = {k: round(1e3 * statistics.median(s[k] for s in stages), 2) for k in stages[0]}
        return med, per_stage
    med_all, stage_all = run(threads)
    med_one, stage_one = run(1)
    o.close()
    return {"value": round(1.0 / med_all, 4), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"{scene}: {warmup} warm-up steps, then the median of {steps} steps with {threads} "
                       f"OpenMP threads and of {steps} more on 1 thread; CPU oracle (C++ restatement of "
                       f"the reference kernels, FP32)"),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "stage_ms": stage_all,
            "one_thread": {"value": round(1.0 / med_one, 4), "unit": "steps/s", "cores": 1, "stage_ms": stage_one}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scene", default="multiphase_128")
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed CPU oracle steps per thread count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--solve-cap", type=int, default=60000)
    ap.add_argument("--tol-steps", type=int, default=10,
                    help="steps per solver mode in the tolerance-mode comparison (0: skip)")
    ap.add_argument("--solve-step", type=int, default=110,
                    help="the step whose pressure system jacobi_to_tol solves (fresh handle)")
    ap.add_argument("--contact-step", type=int, default=160,
                    help="fixed-sweep steps before the post-contact tolerance-mode run")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--timed-only", action="store_true",
                    help="warm-up, timed region and profiled pass only (for kernel traces)")
    ap.add_argument("--scale-only", action="store_true",
                    help="only the 256^3 ball drop in tolerance mode (1e-3; plain, guess, corrections, "
                         "MG-PCG, 3 steps each) as its own JSON line")
    ap.add_argument("--unet-only", action="store_true",
                    help="time the U-Net inference on the workload's step-110 system only")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this FLIP step runs on one GPU only; a multi-GPU run (slab "
              "decomposition with halo exchange) is not part of this build", file=sys.stderr)
        return 2

    import numpy as np

    from aquarius_amd import configs
    from aquarius_amd.fluid import FlipHandle

    cfg = configs.SCENES[args.scene]()
    if args.unet_only:
        print(json.dumps(unet_inference(cfg, args.solve_step)), flush=True)
        return 0
    if args.scale_only:
        big = configs.ball_drop_256()
        print(json.dumps({"workload": "ball_drop_256: 256x256x256 MAC grid, 8 particles per cell, "
                                      "tolerance-mode steps 2-4 of fresh handles",
                          "tolerance_mode": tolerance_mode_steps(big, 3, 200000, tol=1e-3)}), flush=True)
        return 0
    h = FlipHandle(cfg)
    n = h.particle_count
    _, _, frac = h.get_particles()
    K = max(1, int((np.abs(frac).sum(0) > 0).nonzero()[0].max()) + 1) if n else 1
    K = max(K, cfg.phase_count)

    # Timed region: K steps replayed from the captured step graph, one stream
    # synchronisation before (end of warm-up) and after.
    h.steps(args.warmup)
    t0 = time.perf_counter()
    h.steps(args.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = 1000.0 * wall / args.steps

    # The same timed region with paired diffusion sweeps always computed (no skipping of
    # sweeps that would recompute unchanged values), on a second handle of the same scene.
    skip_off = None
    if sweeps_paired_possible(cfg, K) and not args.timed_only:
        os.environ["AQFLIP_NO_FIXED_POINT_SKIP"] = "1"
        try:
            h2 = FlipHandle(cfg)
        finally:
            del os.environ["AQFLIP_NO_FIXED_POINT_SKIP"]
        h2.steps(args.warmup)
        t2 = time.perf_counter()
        h2.steps(args.steps)
        skip_off = round(args.steps / (time.perf_counter() - t2), 3)
        h2.close()

    # Profiled pass: the same K steps launched eagerly with HIP events recorded on
    # the library's stream around every stage (read back once, at the end). The
    # library reports each stage's passes (sweeps, extrapolation passes, or runs),
    # one kernel launch per pass for the sweep stages.
    h.reset_timing()
    h.set_timing(True)
    content0 = h.get_field("content")
    active0 = swept_cells(content0, cfg.size)
    near0 = near_row_faces(content0, cfg.size)
    computed0 = h.paired_diffusion_computed()
    tp0 = time.perf_counter()
    h.steps(args.steps)
    tp1 = time.perf_counter()
    h.set_timing(False)
    timing = h.timing()
    eager_ms_per_step = 1000.0 * (tp1 - tp0) / args.steps
    # cells in chunks with FLUID, averaged over the start and the end of the pass
    content1 = h.get_field("content")
    active_cells = 0.5 * (active0 + swept_cells(content1, cfg.size))
    near_faces = 0.5 * (near0 + near_row_faces(content1, cfg.size))
    diffusion_computed = (h.paired_diffusion_computed() - computed0) / args.steps

    sweeps = {name: launches / args.steps for name, (ms, launches) in timing.items()}
    nbytes = algorithmic_bytes(cfg, n, K, active_cells, sweeps, diffusion_computed, near_faces)
    kernels = {}
    for name, (ms, launches) in timing.items():
        if name not in nbytes or launches == 0:
            continue
        avg_us = 1000.0 * ms / launches
        gbs = nbytes[name] / (avg_us * 1e-6) / 1e9
        kernels[name] = {"ms_per_step": round(ms / args.steps, 4),
                         "launches_per_step": round(launches / args.steps, 3),
                         "avg_launch_us": round(avg_us, 3), "bytes_per_launch": round(nbytes[name]),
                         "achieved_GBs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dominant]
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": dk["achieved_GBs"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dk["frac_of_hbm_peak"],
                "traffic": measured_traffic(args.scene, dominant)}

    h.close()
    solve = tol_mode = fp64 = None
    if not args.no_solve and not args.timed_only:
        solve = solve_comparison(cfg, args.solve_step, args.solve_cap)
        if args.tol_steps > 0:
            tol_mode = {"free_fall": tolerance_mode_steps(cfg, args.tol_steps, args.solve_cap),
                        "after_contact": tolerance_mode_steps(cfg, args.tol_steps, args.solve_cap,
                                                              advance=args.contact_step)}
    if not args.no_fp64 and not args.timed_only:
        fp64 = fp64_error(cfg, args.solve_step)
    unet = None if args.no_solve or args.timed_only else unet_inference(cfg, args.solve_step)

    out = {
        "metric": METRIC,
        "value": round(args.steps / wall, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{args.scene}: {cfg.size[0]}x{cfg.size[1]}x{cfg.size[2]} MAC grid, "
                                f"{n} particles (8 per cell), {K} phases, {cfg.pressure_iters} Jacobi "
                                f"sweeps from zero, {cfg.diffusion_iters} diffusion sweeps, dt {cfg.dt:g}"),
                   "particles": n, "grid": list(cfg.size), "pressure_iters": cfg.pressure_iters,
                   "diffusion_iters": cfg.diffusion_iters, "phases": K},
        "roofline": roofline,
        # every stage's algorithmic bytes per step (launches x bytes per launch, from the
        # profiled pass) over the timed step time
        "whole_step": whole_step(kernels, ms_per_step),
        "kernels": kernels,
        "sweep_cells": {"all": cfg.size[0] * cfg.size[1] * cfg.size[2], "visited_when_listed": round(active_cells)},
        "paired_diffusion": {"sweeps_per_step": sweeps.get("pressure_diffusion", 0),
                             "computed_per_step_profiled_pass": round(diffusion_computed, 2),
                             "steps_per_s_skip_off": skip_off},
        "jacobi_to_tol": solve,
        "tolerance_mode": tol_mode,
        "fp64_error": fp64,
        "unet_inference": unet,
        "profiled_pass_ms_per_step": round(eager_ms_per_step, 4),
    }
    if not args.no_cpu_baseline and not args.timed_only:
        out["cpu_baseline"] = cpu_baseline(args.scene, args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
