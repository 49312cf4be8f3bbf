"""GPU parity: the HIP library (through the C ABI) against the CPU oracle on the same
seeded inputs, stage by stage and over whole steps.

Bar: cell indices, sort order, content flags and every integer field bit-exact;
velocities, pressures and fractions within 1e-5 relative L2 (FP32) -- the kernels
keep the reference's operation order and no FMA contraction, so most fields are
expected to be bit-identical; the tests report how many values differ.
"""
import json
import subprocess
import sys

import numpy as np
import pytest

import aquarius_amd as aq
from aquarius_amd import configs
from aquarius_amd.fluid import FlipHandle
from oracle import binding as ob

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5
INT_FIELDS = ("content", "has_velocity", "particle_count")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def check_field(name, gpu, orc, what=""):
    g = gpu.get_field(name)
    o = orc.get_field(name)
    assert g.shape == o.shape, (name, g.shape, o.shape)
    if name in INT_FIELDS:
        bad = int((g != o).sum())
        assert bad == 0, f"{what}{name}: {bad} integer mismatches"
        return 0.0
    r = rel_l2(g, o)
    ndiff = int((g.astype(np.float64) != o).sum())
    log(f"  {what}{name}: rel_l2={r:.3e} differing={ndiff}/{g.size}")
    assert np.isfinite(g).all() == np.isfinite(o).all()
    assert r <= REL_TOL, (name, r)
    return r


def check_particles(gpu, orc, what=""):
    gp, gv, gf = gpu.get_particles()
    op, ov, of = orc.get_particles()
    for name, g, o in (("pos", gp, op), ("vel", gv, ov), ("frac", gf, of)):
        r = rel_l2(g, o)
        ndiff = int((g.astype(np.float64) != o).sum())
        log(f"  {what}particles.{name}: rel_l2={r:.3e} differing={ndiff}/{g.size}")
        assert r <= REL_TOL, (name, r)


def check_cells(gpu, orc, what=""):
    gs, ge, gh, gperm = gpu.get_cells()
    os_, oe, oh, operm = orc.get_cells()
    assert np.array_equal(gs, os_), f"{what}cellStart differs"
    assert np.array_equal(ge, oe), f"{what}cellEnd differs"
    assert np.array_equal(gh, oh), f"{what}hash differs"
    assert np.array_equal(gperm, operm), f"{what}sort permutation differs"


STAGE_FIELDS = {
    "g2p": [],
    "advect": [],
    "transfer_to_grid": ["content", "velocity", "new_velocity", "has_velocity", "density",
                         "volume_fractions", "new_volume_fractions"],
    "gravity": ["new_velocity"],
    "boundary": ["new_velocity", "content"],
    "divergence": ["divergence", "particle_count"],
    "pressure": ["pressure"],
    "project": ["new_velocity", "has_velocity"],
    "extrapolate": ["new_velocity", "has_velocity"],
    "diffusion": ["new_volume_fractions", "pressure"],
}
STEP_ORDER = ["g2p", "advect", "transfer_to_grid", "gravity", "boundary", "divergence", "pressure",
              "project", "extrapolate", "diffusion"]


def pair(cfg):
    pos, frac = ob.seed(cfg)
    gpu = FlipHandle(cfg, pos=pos, frac4=frac)
    orc = ob.Oracle(cfg, pos=pos, frac4=frac)
    return gpu, orc


def run_stagewise(cfg, steps, label):
    gpu, orc = pair(cfg)
    log(f"[{label}] init, {gpu.particle_count} particles")
    check_cells(gpu, orc, "init ")
    check_particles(gpu, orc, "init ")
    for f in STAGE_FIELDS["transfer_to_grid"] + ["content"]:
        check_field(f, gpu, orc, "init ")
    for s in range(steps):
        for st in STEP_ORDER:
            log(f"[{label}] step {s} stage {st}")
            gpu.run_stage(st)
            orc.run_stage(st)
            if st in ("g2p", "advect", "transfer_to_grid"):
                check_particles(gpu, orc, f"s{s} {st} ")
            if st == "transfer_to_grid":
                check_cells(gpu, orc, f"s{s} ")
            for f in STAGE_FIELDS[st]:
                check_field(f, gpu, orc, f"s{s} {st} ")
    gpu.close()


def test_stagewise_ball_drop_16():
    run_stagewise(configs.ball_drop_16(), steps=3, label="ball_drop_16")


def test_stagewise_ball_drop_64():
    run_stagewise(configs.ball_drop_64(), steps=1, label="ball_drop_64")


def test_stagewise_multiphase_24():
    cfg = configs.make_config(24, dt=0.01, pressure_iters=40, diffusion_iters=10, radius=0.15,
                              sphere_phase=1, phase_count=2, seed=21)
    run_stagewise(cfg, steps=2, label="multiphase_24")


def test_stagewise_edge_configs():
    # non-cubic grid, full boundary coverage, warm-started pressure, 4 phases
    cfg = configs.make_config(0, dt=0.01, pressure_iters=25, diffusion_iters=5, extent=(12, 16, 20),
                              sphere_phase=3, phase_count=4, ref_boundary_coverage=0,
                              pressure_init=aq.configs.AQ_PRESSURE_PREVIOUS, seed=5)
    run_stagewise(cfg, steps=2, label="edge_12x16x20")
    cfg = configs.make_config(3, dt=0.02, pressure_iters=5, diffusion_iters=2, pool_depth=0.4,
                              radius=0.3, center=(0.5, 0.6, 0.5), seed=9)
    run_stagewise(cfg, steps=2, label="tiny_3")


def test_full_steps_match_oracle_16():
    cfg = configs.ball_drop_16()
    gpu, orc = pair(cfg)
    log("[full steps 16] 5 steps")
    gpu.steps(5)
    orc.step(5)
    check_particles(gpu, orc, "5 steps ")
    check_cells(gpu, orc, "5 steps ")
    for f in ("pressure", "new_velocity", "velocity", "content", "divergence"):
        check_field(f, gpu, orc, "5 steps ")
    assert gpu.physical_time == pytest.approx(orc.physical_time, rel=0, abs=0)
    r_g, r_o = gpu.residual(), orc.residual()
    log(f"  residual gpu={r_g:.6e} oracle={r_o:.6e}")
    assert r_g == pytest.approx(r_o, rel=1e-9)
    st = gpu.stats()
    content = orc.get_field("content")[:16, :16, :16]
    assert st["fluid_cells"] == int((content == 1).sum())
    assert st["particles"] == orc.particle_count
    _, _, frac = orc.get_particles()
    np.testing.assert_allclose(st["phase_sum"][0], frac[:, 0].sum(), rtol=1e-12)
    gpu.close()


def free_running_parity(cfg, steps, full_every, label):
    """GPU and oracle stepped side by side from the same seeded particles with no state
    copied between them: the FP64 residual of both pressure fields after every step,
    and every particle and grid field (integer fields and the sort bit-exact) every
    `full_every` steps and at the end."""
    gpu, orc = pair(cfg)
    log(f"[{label}] {gpu.particle_count} particles, {steps} free-running steps")
    worst = 0.0
    for s in range(1, steps + 1):
        gpu.step()
        orc.step()
        r_g, r_o = gpu.residual(), orc.residual()
        assert np.isfinite(r_g) and r_g == pytest.approx(r_o, rel=1e-9), (s, r_g, r_o)
        worst = max(worst, abs(r_g - r_o) / max(abs(r_o), 1e-300))
        if s % full_every == 0 or s == steps:
            log(f"[{label}] step {s}: fp64 residual gpu={r_g:.6e} oracle={r_o:.6e}")
            check_cells(gpu, orc, f"s{s} ")
            check_particles(gpu, orc, f"s{s} ")
            for f in ("content", "velocity", "new_velocity", "pressure", "divergence", "has_velocity",
                      "particle_count", "density", "new_volume_fractions"):
                check_field(f, gpu, orc, f"s{s} ")
    log(f"[{label}] largest relative difference of the per-step FP64 residuals: {worst:.3e}")
    assert gpu.physical_time == orc.physical_time
    gpu.close()


def test_ball_drop_16_all_50_steps_match_oracle():
    """The 16^3 ball drop as stated: 60 Jacobi sweeps from zero, dt 0.01, all 50 steps."""
    free_running_parity(configs.ball_drop_16(), steps=50, full_every=10, label="ball_drop_16 x50")


def test_ball_drop_64_200_steps_match_oracle_with_fp64_residual_each_step():
    """The 64^3 ball drop as stated: 100 Jacobi sweeps from zero, dt 0.005, 200 steps,
    with the FP64 divergence residual compared after every step."""
    free_running_parity(configs.ball_drop_64(), steps=200, full_every=25, label="ball_drop_64 x200")


def test_fp64_error_statement_16():
    """FP32 GPU step against the FP64 oracle run of the same step (reported)."""
    cfg = configs.ball_drop_16()
    pos, frac = ob.seed(cfg)
    gpu = FlipHandle(cfg, pos=pos, frac4=frac)
    o64 = ob.Oracle(cfg, fp64=True, pos=pos, frac4=frac)
    gpu.step()
    o64.step()
    for f in ("pressure", "new_velocity"):
        r = rel_l2(gpu.get_field(f), o64.get_field(f))
        log(f"  fp32-gpu vs fp64-oracle {f}: rel_l2={r:.3e}")
        assert r < 1e-3
    gpu.close()


STATE_FIELDS = ("content", "velocity", "new_velocity", "pressure", "divergence", "density", "particle_count",
                "has_velocity", "volume_fractions", "new_volume_fractions")


@pytest.mark.parametrize("scene,at_step", [("ball_drop_16", 25), ("ball_drop_64", 100)])
def test_fp64_error_statement_mid_run(scene, at_step):
    """The FP32 GPU step against an FP64 run of the same step at the 16^3 and 64^3 ball
    drops: the GPU's whole state after `at_step` steps goes to the FP64 oracle and both
    run one step (reported; the bound only guards against a broken comparison)."""
    cfg = configs.SCENES[scene]()
    gpu = FlipHandle(cfg)
    gpu.steps(at_step)
    pos, vel, frac = gpu.get_particles()
    o64 = ob.Oracle(cfg, fp64=True, pos=pos, vel=vel, frac4=frac, init=False)
    for f in STATE_FIELDS:
        o64.set_field(f, gpu.get_field(f))
    gpu.step()
    o64.step()
    for f in ("pressure", "new_velocity", "velocity"):
        r = rel_l2(gpu.get_field(f), o64.get_field(f))
        log(f"  [{scene} step {at_step + 1}] fp32-gpu vs fp64-oracle {f}: rel_l2={r:.3e}")
        assert r < 1e-3
    gp, gv, _ = gpu.get_particles()
    op, ov, _ = o64.get_particles()
    _, _, gh, pg = gpu.get_cells()
    _, _, oh, po = o64.get_cells()
    ig, io = np.empty_like(pg), np.empty_like(po)
    ig[pg] = np.arange(pg.size)
    io[po] = np.arange(po.size)
    rp, rv = rel_l2(gp[ig], op[io]), rel_l2(gv[ig], ov[io])
    moved = int((gh[ig] != oh[io]).sum())
    log(f"  [{scene} step {at_step + 1}] particles: position rel_l2={rp:.3e}, velocity rel_l2={rv:.3e}, "
        f"{moved} of {pg.size} in another cell")
    assert rp < 1e-5 and rv < 1e-3
    gpu.close()


def test_solve_pressure_to_tolerance_16():
    cfg = configs.ball_drop_16()
    gpu = FlipHandle(cfg)
    gpu.step()
    it, rel, ok = gpu.solve_pressure(1e-4, 20000, check_every=50)
    log(f"  solve to 1e-4: iters={it} residual={rel:.3e} ok={ok}")
    assert ok and rel <= 1e-4
    # the cap is reported as a stall
    gpu2 = FlipHandle(cfg)
    gpu2.step()
    it2, rel2, ok2 = gpu2.solve_pressure(1e-12, 30, check_every=10)
    assert not ok2 and it2 == 30
    gpu.close()
    gpu2.close()


def test_cpp_driver_runs():
    from aquarius_amd.build import DRIVER
    out = subprocess.run([str(DRIVER), "--n", "16", "--steps", "5"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    log("  driver:", rec)
    assert rec["steps"] == 5 and rec["particles"] > 9000
    assert rec["physical_time"] == pytest.approx(0.05, rel=1e-5)
    assert rec["center"] == [5, 5, 5] and rec["container"] == 10


def test_python_fluid_surface():
    f = aq.Fluid()
    f.init(configs.ball_drop_16())
    for _ in range(3):
        f.simulationStep()
    assert f.physicalTime == pytest.approx(0.03, rel=1e-5)
    assert f.getCenter() == (5.0, 5.0, 5.0) and f.getContainerSize() == 10.0
    f.draw()


@pytest.mark.parametrize("pressure_iters,diffusion_iters", [(40, 10), (12, 30)])
def test_full_steps_match_oracle_multiphase_24(pressure_iters, diffusion_iters):
    """Whole steps of a two-phase scene through the captured step graph, whose
    diffusion sweeps share launches with the pressure sweeps (the longer chain
    finishing on its own: pressure in the first case, diffusion in the second)."""
    cfg = configs.make_config(24, dt=0.01, pressure_iters=pressure_iters, diffusion_iters=diffusion_iters,
                              radius=0.15, sphere_phase=1, phase_count=2, seed=21)
    gpu, orc = pair(cfg)
    log(f"[full steps multiphase 24 J{pressure_iters} D{diffusion_iters}] 4 steps")
    gpu.steps(4)
    orc.step(4)
    check_particles(gpu, orc, "4 steps ")
    check_cells(gpu, orc, "4 steps ")
    for f in ("pressure", "new_velocity", "new_volume_fractions", "volume_fractions", "density"):
        check_field(f, gpu, orc, "4 steps ")
    gpu.close()


@pytest.mark.parametrize("pressure_iters,diffusion_iters", [(40, 10), (12, 30)])
def test_full_steps_match_oracle_mixing_24(pressure_iters, diffusion_iters):
    """A two-phase ball resting in the pool: the phases mix from the first step, so
    every paired diffusion sweep changes the fractions and none is skipped."""
    cfg = configs.make_config(24, dt=0.01, pressure_iters=pressure_iters, diffusion_iters=diffusion_iters,
                              radius=0.15, center=(0.5, 0.2, 0.5), sphere_phase=1, phase_count=2,
                              diffusion_coeff=0.05, seed=23)
    gpu, orc = pair(cfg)
    log(f"[mixing 24 J{pressure_iters} D{diffusion_iters}] 4 steps")
    gpu.steps(4)
    orc.step(4)
    check_particles(gpu, orc, "4 steps ")
    check_cells(gpu, orc, "4 steps ")
    for f in ("pressure", "new_velocity", "new_volume_fractions", "volume_fractions", "density"):
        check_field(f, gpu, orc, "4 steps ")
    frac = gpu.get_field("new_volume_fractions")[:24, :24, :24]
    mixed = ((frac[..., 0] > 0) & (frac[..., 1] > 0)).sum()
    assert mixed > 0, "the scene must mix phases"
    gpu.close()


@pytest.mark.parametrize("scene", ["separated", "mixing"])
def test_fixed_point_skip_equals_full_compute_64(monkeypatch, scene):
    """Paired diffusion sweeps after a sweep that changed nothing skip their work; the
    steps give the same bits as computing every sweep."""
    center = (0.5, 0.7, 0.5) if scene == "separated" else (0.5, 0.2, 0.5)
    cfg = configs.make_config(64, dt=0.005, pressure_iters=60, diffusion_iters=40, radius=0.15,
                              center=center, sphere_phase=1, phase_count=2, seed=29)
    pos, frac = ob.seed(cfg)
    out, computed = {}, {}
    for flag in ("1", "0"):
        monkeypatch.setenv("AQFLIP_NO_FIXED_POINT_SKIP", flag)
        h = FlipHandle(cfg, pos=pos, frac4=frac)
        log(f"[64 {scene} no_skip={flag}] 3 steps")
        h.steps(3)
        out[flag] = {f: h.get_field(f) for f in ("pressure", "new_velocity", "new_volume_fractions")}
        out[flag]["pos"] = h.get_particles()[0]
        computed[flag] = h.paired_diffusion_computed()
        h.close()
    for k in out["0"]:
        assert np.array_equal(out["0"][k].view(np.uint32), out["1"][k].view(np.uint32)), k
    paired = 3 * min(cfg.pressure_iters, cfg.diffusion_iters)
    assert computed["1"] == paired  # skipping off: every paired sweep computed
    if scene == "separated":  # one-hot phases apart: only each step's first sweep computes
        assert computed["0"] == 3
    else:
        assert 3 <= computed["0"] <= paired


def test_listed_sweeps_on_eight_workgroups_64(monkeypatch):
    """Listed sweeps read each workgroup's first two chunks from a table and the rest
    from the chunk list. With 8 workgroups (AQFLIP_LISTED_GRID=8) each walks about 25
    chunks, nearly all from the list; the steps give the same bits as the default grids
    (at most two chunks per workgroup here)."""
    cfg = configs.make_config(64, dt=0.005, pressure_iters=60, diffusion_iters=40, radius=0.15,
                              center=(0.5, 0.2, 0.5), sphere_phase=1, phase_count=2, seed=29)
    pos, frac = ob.seed(cfg)
    out = {}
    for grid in ("0", "8"):
        monkeypatch.setenv("AQFLIP_LISTED_GRID", grid)
        h = FlipHandle(cfg, pos=pos, frac4=frac)
        log(f"[64 listed grid {grid}] 3 steps")
        h.steps(3)
        out[grid] = {f: h.get_field(f) for f in ("pressure", "new_velocity", "new_volume_fractions")}
        out[grid]["pos"] = h.get_particles()[0]
        active = int((h.get_field("content") == 1).sum())
        h.close()
    assert active > 8 * 512 * 2  # more FLUID cells than 8 workgroups' table entries cover
    for k in out["0"]:
        assert np.array_equal(out["0"][k].view(np.uint32), out["8"][k].view(np.uint32)), k


def test_concurrent_diffusion_equals_sequential_128(monkeypatch):
    """Fused pressure + diffusion launches, diffusion on a second stream, and the
    reference's sequential order give the same bits at 128^3."""
    cfg = configs.multiphase_128()
    pos, frac = ob.seed(cfg)
    out = {}
    for mode, (no_fuse, no_overlap) in {"fused": ("0", "0"), "overlap": ("1", "0"),
                                        "sequential": ("1", "1")}.items():
        monkeypatch.setenv("AQFLIP_NO_FUSE", no_fuse)
        monkeypatch.setenv("AQFLIP_NO_OVERLAP", no_overlap)
        h = FlipHandle(cfg, pos=pos, frac4=frac)
        log(f"[128 {mode}] 3 steps")
        h.steps(3)
        out[mode] = {f: h.get_field(f) for f in ("pressure", "new_velocity", "new_volume_fractions")}
        out[mode]["pos"] = h.get_particles()[0]
        h.close()
    for mode in ("fused", "overlap"):
        for k in out[mode]:
            assert np.array_equal(out[mode][k].view(np.uint32), out["sequential"][k].view(np.uint32)), (mode, k)


def test_stage_timing_counts_and_matches_graph(monkeypatch):
    """The profiled (eager, event-timed) step runs the same paired sweeps as the
    captured step, reports one pass per sweep, and gives the same bits."""
    cfg = configs.make_config(24, dt=0.01, pressure_iters=40, diffusion_iters=10, radius=0.15,
                              sphere_phase=1, phase_count=2, seed=21)
    pos, frac = ob.seed(cfg)
    out = {}
    for mode in ("graph", "timed"):
        h = FlipHandle(cfg, pos=pos, frac4=frac)
        log(f"[timing {mode}] 2 steps")
        if mode == "timed":
            h.set_timing(True)
        h.steps(2)
        if mode == "timed":
            t = h.timing()
            assert t["pressure_diffusion"][1] == 2 * 10
            assert t["pressure"][1] == 2 * 30
            assert t["diffusion"][1] == 0
            assert t["extrapolate"][1] == 2 * 3
            assert t["g2p"][1] == 2 and t["transfer_to_grid"][1] == 2 and t["p2g"][1] == 2
            assert all(ms >= 0 for ms, _ in t.values())
        out[mode] = {f: h.get_field(f) for f in ("pressure", "new_velocity", "new_volume_fractions")}
        out[mode]["pos"] = h.get_particles()[0]
        h.close()
    for k in out["graph"]:
        assert np.array_equal(out["graph"][k].view(np.uint32), out["timed"][k].view(np.uint32)), k
    monkeypatch.setenv("AQFLIP_NO_FUSE", "1")
    h = FlipHandle(cfg, pos=pos, frac4=frac)
    h.set_timing(True)
    h.steps(1)
    t = h.timing()
    h.close()
    assert t["pressure_diffusion"][1] == 0 and t["pressure"][1] == 40 and t["diffusion"][1] == 10


def test_full_steps_match_oracle_non_cubic_64_rows():
    """Whole steps on a 20 x 24 x 64 grid (X != Y, and z rows of 65 faces, so a wave of
    the last two extrapolation passes spans at most two face rows and tests them both
    with scalar loads) against the oracle: the near-row passes, their row stamps and the
    interleaved velocity copy the next step's G2P reads."""
    cfg = configs.make_config(0, dt=0.01, pressure_iters=30, diffusion_iters=5, extent=(20, 24, 64),
                              sphere_phase=1, phase_count=2, radius=0.25, center=(0.45, 0.65, 0.55), seed=11)
    gpu, orc = pair(cfg)
    log("[full steps 20x24x64] 6 steps")
    gpu.steps(6)
    orc.step(6)
    check_particles(gpu, orc, "6 steps ")
    check_cells(gpu, orc, "6 steps ")
    for f in ("pressure", "new_velocity", "velocity", "has_velocity", "content", "volume_fractions"):
        check_field(f, gpu, orc, "6 steps ")
    gpu.close()


def test_steps_after_a_field_upload_match_an_uninterrupted_run():
    """Uploading a velocity field (here unchanged) marks the interleaved (newVelocity,
    velocity) copy stale; the next step rebuilds it in full and must continue exactly as
    a run without the upload."""
    cfg = configs.ball_drop_64()
    a, b = FlipHandle(cfg), FlipHandle(cfg)
    a.steps(8)
    b.steps(7)
    b.set_field("new_velocity", b.get_field("new_velocity"))
    b.set_field("velocity", b.get_field("velocity"))
    b.steps(1)
    for f in ("pressure", "new_velocity", "velocity", "has_velocity", "content"):
        np.testing.assert_array_equal(a.get_field(f), b.get_field(f), err_msg=f)
    for x, y in zip(a.get_particles(), b.get_particles()):
        np.testing.assert_array_equal(x, y)
    a.close()
    b.close()
