"""Byte model of bench.py's roofline figures, checked against the closed forms per
stage (128^3 multiphase: 4,431,408 particles, 2 phases) and the whole-step sum."""
import bench
from aquarius_amd import configs


def test_algorithmic_bytes_closed_forms():
    cfg = configs.multiphase_128()
    Np, K = 4_431_408, 2
    C = 128 ** 3
    V = 563_712.0  # cells in the chunks holding FLUID
    sweeps = {"pressure_diffusion": 100, "pressure": 50}
    b = bench.algorithmic_bytes(cfg, Np, K, V, sweeps, diffusion_computed=1.0)
    rec = 24 + 4 * K
    assert b["p2g"] == Np * rec + C * (34 + 8 * K)
    assert b["g2p"] == Np * (24 + 24 + 4 * K + 4 + 4) + C * (24 + 4 * K)
    assert b["transfer_to_grid"] == 12 * C + 12 * Np + 4 * Np + Np * (4 + 2 * rec)
    # pressure alone: every sweep listed (the paired sweeps rewrite both buffers first)
    assert b["pressure"] == 13 * V
    # paired: 2 of 100 pressure sweeps full, diffusion computed once (the first, full)
    assert abs(b["pressure_diffusion"] - (13 * (2 * C + 98 * V) / 100 + (1 + 8 * K) * C / 100)) < 1e-6
    assert b["extrapolate"] == 26 * C
    # second and third passes over the near rows only
    b2 = bench.algorithmic_bytes(cfg, Np, K, V, sweeps, 1.0, near_faces=600_000)
    assert abs(b2["extrapolate"] - 26 * (C + 2 * 600_000) / 3) < 1e-6


def test_near_row_faces_box_of_nine_rows():
    import numpy as np
    X = Y = Z = 16
    content = np.zeros((X + 1, Y + 1, Z + 1), np.int32)
    content[8, 8, 3] = 1  # one FLUID cell: face rows 4..12 in x and in y
    assert bench.near_row_faces(content, (X, Y, Z)) == 81 * (Z + 1)
    content[0, 0, 0] = 1  # a corner: rows 0..4 (clipped), sharing row (4, 4) with the first box
    assert bench.near_row_faces(content, (X, Y, Z)) == (81 + 25 - 1) * (Z + 1)


def test_whole_step_sums_launches_times_bytes():
    kernels = {"a": {"bytes_per_launch": 1000, "launches_per_step": 3.0},
               "b": {"bytes_per_launch": 500, "launches_per_step": 1.0}}
    w = bench.whole_step(kernels, ms_per_step=0.001)  # 1 us per step
    assert w["bytes_per_step"] == 3500
    assert abs(w["achieved_GBs"] - 3500 / 1e-6 / 1e9) < 0.1
    assert abs(w["frac_of_hbm_peak"] - w["achieved_GBs"] / bench.HBM_PEAK_GBS) < 1e-4
