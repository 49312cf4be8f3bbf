"""Multi-sweep launches (many Jacobi or diffusion sweeps per launch, separated by
grid-wide barriers) against the oracle and against one launch per sweep.

Enabled per handle by AQFLIP_MULTI_SWEEP=1 at creation. The per-cell arithmetic is
the single-sweep kernels', so every field must match bit for bit."""
import os

import numpy as np
import pytest

from aquarius_amd import configs
from aquarius_amd.fluid import FlipHandle
from oracle import binding as ob

from test_gpu_parity import check_cells, check_field, check_particles, log, run_stagewise

pytestmark = pytest.mark.gpu


@pytest.fixture
def multi(monkeypatch):
    monkeypatch.setenv("AQFLIP_MULTI_SWEEP", "1")


def test_multisweep_stagewise_multiphase_24(multi):
    cfg = configs.make_config(24, dt=0.01, pressure_iters=40, diffusion_iters=10, radius=0.15,
                              sphere_phase=1, phase_count=2, seed=21)
    run_stagewise(cfg, steps=2, label="multi multiphase_24")


def test_multisweep_full_steps_ball_drop_64(multi):
    cfg = configs.ball_drop_64()
    pos, frac = ob.seed(cfg)
    gpu = FlipHandle(cfg, pos=pos, frac4=frac)
    orc = ob.Oracle(cfg, pos=pos, frac4=frac)
    log("[multi 64] 3 steps")
    gpu.steps(3)
    orc.step(3)
    check_particles(gpu, orc, "3 steps ")
    check_cells(gpu, orc, "3 steps ")
    for f in ("pressure", "new_velocity", "content", "divergence"):
        check_field(f, gpu, orc, "3 steps ")
    gpu.close()


def test_multisweep_equals_single_launches_128(monkeypatch):
    """Full-size multiphase workload: 3 steps with and without multi-sweep launches."""
    cfg = configs.multiphase_128()
    pos, frac = ob.seed(cfg)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("AQFLIP_MULTI_SWEEP", flag)
        h = FlipHandle(cfg, pos=pos, frac4=frac)
        log(f"[128 multi={flag}] 3 steps")
        h.steps(3)
        out[flag] = {f: h.get_field(f) for f in ("pressure", "new_velocity", "new_volume_fractions")}
        out[flag]["pos"] = h.get_particles()[0]
        out[flag]["res"] = h.residual()
        h.close()
    for k in ("pressure", "new_velocity", "new_volume_fractions", "pos"):
        a, b = out["0"][k], out["1"][k]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    assert out["0"]["res"] == out["1"]["res"]


def test_multisweep_tolerance_solve(multi):
    cfg = configs.ball_drop_16()
    gpu = FlipHandle(cfg)
    gpu.step()
    it, rel, ok = gpu.solve_pressure(1e-4, 20000, check_every=50)
    log(f"  multi solve to 1e-4: iters={it} residual={rel:.3e}")
    assert ok and rel <= 1e-4
    gpu.close()
