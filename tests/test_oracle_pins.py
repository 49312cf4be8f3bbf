"""Pins of the CPU oracle (oracle/flip_oracle.hpp) against what fixes each function
independently of the oracle's own code: the report's worked example, library
routines (stable sort, dense solve), brute force on tiny inputs, closed forms and
hand-built cases. CPU only.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from aquarius_amd import configs
from aquarius_amd._abi import AQ_SHAPE_BOX, AQ_SHAPE_SPHERE, Config
from oracle import binding as ob

GOLDEN = Path(__file__).resolve().parent / "golden"
AIR, FLUID, SOLID = 0, 1, 2


def tiny_config(X, Y, Z, **kw):
    cfg = configs.make_config(4, dt=kw.pop("dt", 0.01), pressure_iters=kw.pop("pressure_iters", 10),
                              extent=(X, Y, Z), **kw)
    return cfg


# ---------------------------------------------------------------- hashing and sort
def test_report_spatial_index_example():
    """report/ChapterImplementFLIP.tex:42-269: 4x4 grid, hash = x*4+y."""
    kat = json.loads((GOLDEN / "report_spatial_index.json").read_text())
    cfg = tiny_config(4, 4, 1)
    cfg.grid_physical_size = 4.0  # dx = 1: positions are in cell units
    pos = np.array([[x, y, 0.5] for x, y in kat["positions_xy"]], np.float32)
    o = ob.Oracle(cfg, pos=pos, init=False)
    o.substage("hash_sort")
    start, end, hsh, perm = o.get_cells()
    # hashes in the original order
    unsorted = np.empty_like(hsh)
    unsorted[perm] = hsh
    assert unsorted.tolist() == kat["hashes"]
    ranges = {int(k): v for k, v in kat["cell_ranges"].items()}
    for c in range(len(start)):
        if c in ranges:
            assert [start[c], end[c]] == ranges[c]
        else:
            assert start[c] == -1 and end[c] == -1
    # within a cell the previous order is kept (stable)
    assert perm.tolist() == np.argsort(np.array(kat["hashes"]), kind="stable").tolist()


@pytest.mark.parametrize("scene", ["ball_drop_16", "ball_drop_64"])
def test_sort_is_stable_sort_by_hash(scene):
    cfg = configs.SCENES[scene]()
    pos, frac = ob.seed(cfg)
    rng = np.random.default_rng(7)
    pos = pos[rng.permutation(len(pos))]
    o = ob.Oracle(cfg, pos=pos, frac4=None, init=False)
    o.substage("hash_sort")
    start, end, hsh, perm = o.get_cells()
    X, Y, Z = cfg.size
    h = np.float32(cfg.grid_physical_size) / np.float32(Y)
    cell = (pos / h).astype(np.int32)  # truncation, float32 division
    ref_hash = cell[:, 0] * (Y * Z) + cell[:, 1] * Z + cell[:, 2]
    ref_perm = np.argsort(ref_hash, kind="stable")
    assert np.array_equal(perm, ref_perm)
    assert np.array_equal(hsh, ref_hash[ref_perm])
    # inclusive cell ranges, -1 when empty, (X+1)(Y+1)(Z+1) entries
    counts = np.bincount(ref_hash, minlength=(X + 1) * (Y + 1) * (Z + 1))
    first = np.searchsorted(hsh, np.arange(len(counts)), side="left")
    occ = counts > 0
    assert np.array_equal(start[occ], first[occ]) and np.all(start[~occ] == -1)
    assert np.array_equal(end[occ], first[occ] + counts[occ] - 1) and np.all(end[~occ] == -1)


# ---------------------------------------------------------------- P2G brute force
def hat(r):
    r = np.abs(r)
    return np.where(r > 1, 0.0, 1 - r)


def brute_p2g(pos, vel, X, Y, Z, h):
    """Naive O(faces x particles) FP64 splat of the trilinear hat."""
    out = np.zeros((X + 1, Y + 1, Z + 1, 4))
    wsum = np.zeros((X + 1, Y + 1, Z + 1, 3))
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                centres = [(x, y + .5, z + .5), (x + .5, y, z + .5), (x + .5, y + .5, z)]
                for a, c in enumerate(centres):
                    r = (pos - np.array(c) * h) / h
                    w = hat(r[:, 0]) * hat(r[:, 1]) * hat(r[:, 2])
                    ws = w.sum()
                    wsum[x, y, z, a] = ws
                    out[x, y, z, a] = (w * vel[:, a]).sum() / ws if ws > 0 else 0.0
    return out, wsum


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 3, 5)])
def test_p2g_matches_brute_force(shape):
    X, Y, Z = shape
    cfg = tiny_config(X, Y, Z)
    h = cfg.grid_physical_size / Y
    rng = np.random.default_rng(3)
    n = 60
    pos = (rng.random((n, 3)) * np.array([X, Y, Z]) * h * 0.98 + 0.01 * h).astype(np.float32)
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    o = ob.Oracle(cfg, fp64=True, pos=pos, vel=vel, init=False)
    o.run_stage("transfer_to_grid")
    got = o.get_field("velocity")
    want, wsum = brute_p2g(pos.astype(np.float64), vel.astype(np.float64), X, Y, Z, h)
    np.testing.assert_allclose(got[:X, :Y, :Z, :3], want[:X, :Y, :Z, :3], rtol=1e-12, atol=1e-12)
    hv = o.get_field("has_velocity")
    assert np.array_equal(hv[..., :3].astype(bool), wsum[:X, :Y, :Z] > 0)
    # FP32 oracle agrees to single precision
    o32 = ob.Oracle(cfg, fp64=False, pos=pos, vel=vel, init=False)
    o32.run_stage("transfer_to_grid")
    np.testing.assert_allclose(o32.get_field("velocity")[:X, :Y, :Z, :3], want[:X, :Y, :Z, :3],
                               rtol=2e-5, atol=2e-6)


def test_single_particle_on_face_has_weight_one():
    """A particle sitting on a u-face position gets weight exactly 1 there."""
    cfg = tiny_config(4, 4, 4)
    h = cfg.grid_physical_size / 4
    pos = np.array([[2 * h, 1.5 * h, 2.5 * h]], np.float32)
    vel = np.array([[0.75, -0.25, 0.5]], np.float32)
    o = ob.Oracle(cfg, fp64=True, pos=pos, vel=vel, init=False)
    o.run_stage("transfer_to_grid")
    v = o.get_field("velocity")
    assert v[2, 1, 2, 0] == pytest.approx(0.75, abs=0)
    content = o.get_field("content")
    assert content[2, 1, 2] == FLUID and content[:4, :4, :4].sum() == FLUID


def test_zero_velocity_rest_state():
    cfg = tiny_config(6, 6, 6)
    o = ob.Oracle(cfg, fp64=False)
    o.run_stage("g2p")
    assert np.all(o.get_particles()[1] == 0)
    assert np.all(o.get_field("velocity") == 0)


# ---------------------------------------------------------------- pressure system
def build_dense_A(content, X, Y, Z):
    """Matrix of the legacy assembly (Fluid/Legacy/PressureEquation.cuh:216-328):
    diagonal 6 minus out-of-domain (or SOLID) neighbours, -1 per FLUID neighbour."""
    fluid = [(x, y, z) for x in range(X) for y in range(Y) for z in range(Z)
             if content[x, y, z] == FLUID]
    idx = {c: i for i, c in enumerate(fluid)}
    A = np.zeros((len(fluid), len(fluid)))
    for (x, y, z), i in idx.items():
        d = 6.0
        for dx, dy, dz in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
            n = (x + dx, y + dy, z + dz)
            if not (0 <= n[0] < X and 0 <= n[1] < Y and 0 <= n[2] < Z):
                d -= 1
            elif content[n] == FLUID:
                A[i, idx[n]] = -1.0
        A[i, i] = d
    return A, fluid


def pressure_problem(X=5, Y=4, Z=6, seed=0):
    rng = np.random.default_rng(seed)
    content = np.full((X + 1, Y + 1, Z + 1), SOLID, np.int32)
    inner = np.where(rng.random((X, Y, Z)) < 0.75, FLUID, AIR)
    inner[:, Y - 1, :] = AIR  # free surface on top: every fluid region sees air
    content[:X, :Y, :Z] = inner
    div = rng.standard_normal((X, Y, Z))
    dens = 0.5 + rng.random((X, Y, Z))
    cfg = tiny_config(X, Y, Z, pressure_iters=0)
    o = ob.Oracle(cfg, fp64=True, init=False)
    o.set_field("content", content)
    o.set_field("divergence", div)
    o.set_field("density", dens)
    o.set_field("pressure", np.zeros((X, Y, Z)))
    A, fluid = build_dense_A(content, X, Y, Z)
    b = np.array([-div[c] * dens[c] for c in fluid])
    return o, A, b, fluid, (X, Y, Z)


def test_jacobi_converges_to_dense_solution():
    o, A, b, fluid, (X, Y, Z) = pressure_problem()
    p_exact = np.linalg.solve(A, b)
    o.jacobi_sweeps(4000)
    p = o.get_field("pressure")
    got = np.array([p[c] for c in fluid])
    np.testing.assert_allclose(got, p_exact, rtol=1e-9, atol=1e-9)
    # AIR cells are Dirichlet zero
    content = o.get_field("content")
    assert np.all(p[content[:X, :Y, :Z] == AIR] == 0)
    assert o.residual() < 1e-10


def test_exact_solution_is_a_jacobi_fixed_point():
    o, A, b, fluid, (X, Y, Z) = pressure_problem(seed=4)
    p_exact = np.linalg.solve(A, b)
    p = np.zeros((X, Y, Z))
    for c, v in zip(fluid, p_exact):
        p[c] = v
    o.set_field("pressure", p)
    assert o.residual() < 1e-12
    o.jacobi_sweeps(1)
    np.testing.assert_allclose(o.get_field("pressure"), p, rtol=1e-12, atol=1e-12)


def test_residual_matches_dense_operator():
    o, A, b, fluid, (X, Y, Z) = pressure_problem(seed=9)
    rng = np.random.default_rng(1)
    p = rng.standard_normal((X, Y, Z))
    o.set_field("pressure", p)
    pv = np.array([p[c] for c in fluid])
    want = np.linalg.norm(b - A @ pv) / np.linalg.norm(b)
    assert o.residual() == pytest.approx(want, rel=1e-12)


# ---------------------------------------------------------------- gravity
def test_gravity_gather_equals_sequential_sweep():
    """applyGravityImpl (Fluid_3D_kernels.cu:4-39) as a sequential in-place sweep over
    source cells, with every gravity component non-zero."""
    X, Y, Z = 4, 5, 3
    cfg = tiny_config(X, Y, Z)
    cfg.gravity[0], cfg.gravity[1], cfg.gravity[2] = 1.5, -9.8, 2.5
    rng = np.random.default_rng(5)
    content = rng.choice([AIR, FLUID], size=(X + 1, Y + 1, Z + 1)).astype(np.int32)
    content[X, :, :] = rng.choice([AIR, SOLID], size=(Y + 1, Z + 1))
    nv = rng.standard_normal((X + 1, Y + 1, Z + 1, 4))
    nv[..., 3] = 0
    o = ob.Oracle(cfg, fp64=True, init=False)
    o.set_field("content", content)
    o.set_field("new_velocity", nv)
    o.run_stage("gravity")
    got = o.get_field("new_velocity")
    want = nv.copy()
    add = np.array([1.5, -9.8, 2.5, 0.0], np.float32).astype(np.float64) * np.float64(np.float32(cfg.dt))
    add = np.array([np.float64(np.float32(g)) * np.float64(np.float32(cfg.dt)) for g in (1.5, -9.8, 2.5)] + [0.0])
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                if content[x, y, z] != FLUID:
                    continue
                want[x, y, z] += add
                for d in range(3):
                    t = [x, y, z]
                    t[d] += 1
                    if content[tuple(t)] == AIR:
                        want[tuple(t)] += add
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


# ---------------------------------------------------------------- boundary coverage
def launched_pairs(A, B):
    total = A * B
    nt = min(1024, total)
    blocks = total // nt + (1 if total % nt else 0)
    return min(blocks * nt, (A + 1) * (B + 1))


@pytest.mark.parametrize("n,missing", [(16, 33), (50, 0)])
def test_fix_boundary_reference_coverage(n, missing):
    """fixBoundaryX launches N^2 threads for (N+1)^2 pairs (Fluid_3D_common.cu:20-37)."""
    cfg = tiny_config(n, n, n)
    o = ob.Oracle(cfg, fp64=False, init=False)
    nv = -np.ones((n + 1, n + 1, n + 1, 4))
    o.set_field("new_velocity", nv)
    o.run_stage("boundary")
    got = o.get_field("new_velocity")
    untouched_x0 = (got[0, :, :, 0] == -1).sum()
    assert untouched_x0 == missing == (n + 1) ** 2 - launched_pairs(n, n)
    content = o.get_field("content")
    assert (content[n, :, :] == SOLID).sum() == (n + 1) ** 2 - missing
    if missing:
        # X misses rows y = N-1 (z >= 1) and y = N
        assert np.all(got[0, n, :, 0] == -1) and np.all(got[0, n - 1, 1:, 0] == -1)
    cfg.ref_boundary_coverage = 0
    o2 = ob.Oracle(cfg, fp64=False, init=False)
    o2.set_field("new_velocity", nv)
    o2.run_stage("boundary")
    assert (o2.get_field("new_velocity")[0, :, :, 0] == -1).sum() == 0


# ---------------------------------------------------------------- divergence
def test_divergence_closed_form():
    X, Y, Z = 4, 4, 4
    cfg = tiny_config(X, Y, Z)
    o = ob.Oracle(cfg, fp64=True, pos=np.array([[0.1, 0.1, 0.1]], np.float32), init=False)
    o.substage("hash_sort")
    ix, iy, iz = np.meshgrid(np.arange(X + 1), np.arange(Y + 1), np.arange(Z + 1), indexing="ij")
    nv = np.stack([0.1 * ix, 0.2 * iy, 0.3 * iz, 0 * ix], -1).astype(np.float64)
    o.set_field("new_velocity", nv)
    o.run_stage("divergence")
    div = o.get_field("divergence")
    # one particle in cell 0 -> particleCount 0 (inclusive end), like every empty cell
    assert np.all(o.get_field("particle_count") == 0)
    np.testing.assert_allclose(div, 0.6 + 0.008, rtol=1e-12)


# ---------------------------------------------------------------- extrapolation
def test_extrapolation_direction_test():
    """extrapolateVelocityByOne (Fluid_3D_kernels.cu:210-422) on a 3^3 hand case."""
    cfg = tiny_config(3, 3, 3)
    for u, expect_left in [(0.5, True), (1.5, False)]:
        o = ob.Oracle(cfg, fp64=True, init=False)
        hv = np.zeros((3, 3, 3, 4))
        hv[1, 1, 1, 0] = 1
        nv = np.zeros((4, 4, 4, 4))
        nv[1, 1, 1, 0] = u
        o.set_field("has_velocity", hv)
        o.set_field("new_velocity", nv)
        o.substage("extrapolate_once")
        got = o.get_field("new_velocity")
        h2 = o.get_field("has_velocity")
        # the -x neighbour takes u only if u < 1 - 1e-6 (its +x neighbour must not point away)
        assert (got[0, 1, 1, 0] == u) == expect_left and bool(h2[0, 1, 1, 0]) == expect_left
        # the +x neighbour needs u > -1 + 1e-6: always true here
        assert got[2, 1, 1, 0] == u and h2[2, 1, 1, 0] == 1
        # y and z neighbours test the neighbour's own v / w (zero): pass
        for c in [(1, 0, 1), (1, 2, 1), (1, 1, 0), (1, 1, 2)]:
            assert got[c + (0,)] == u
        # diagonal cells are two steps away
        assert got[0, 0, 1, 0] == 0 and h2[0, 0, 1, 0] == 0
        # components y, z nowhere set
        assert np.all(h2[..., 1:3] == 0)


# ---------------------------------------------------------------- diffusion
def test_diffusion_two_phase_closed_form():
    cfg = tiny_config(3, 3, 3, phase_count=2)
    lam = np.float64(np.float32(cfg.diffusion_coeff) * np.float32(cfg.dt)) / (np.float64(cfg.grid_physical_size / 3) ** 2)
    o = ob.Oracle(cfg, fp64=True, init=False)
    content = np.full((4, 4, 4), SOLID)
    content[:3, :3, :3] = FLUID
    f = np.zeros((3, 3, 3, 4))
    f[..., 0] = 1
    f[1, 1, 1] = [0, 1, 0, 0]
    o.set_field("content", content)
    o.set_field("new_volume_fractions", f)
    o.substage("diffusion_sweep")
    got = o.get_field("new_volume_fractions")
    lam = o_lambda(cfg)
    np.testing.assert_allclose(got[1, 1, 1, :2], [6 * lam / (1 + 6 * lam), 1 / (1 + 6 * lam)], rtol=1e-12)
    np.testing.assert_allclose(got[0, 1, 1, :2], [(5 * lam + 1) / (6 * lam + 1), lam / (6 * lam + 1)], rtol=1e-12)
    np.testing.assert_allclose(got.sum(-1), 1.0, rtol=1e-14)


def o_lambda(cfg):
    h = np.float64(cfg.grid_physical_size) / cfg.size[1]
    return (np.float64(np.float32(cfg.diffusion_coeff)) * np.float64(np.float32(cfg.dt))) / (h * h)


def test_single_phase_diffusion_is_identity():
    cfg = configs.ball_drop_16()
    o = ob.Oracle(cfg)
    before = o.get_field("new_volume_fractions")
    o.run_stage("diffusion")
    assert np.array_equal(o.get_field("new_volume_fractions"), before)


# ---------------------------------------------------------------- G2P and RK3
def test_g2p_reproduces_linear_field_and_blend():
    X, Y, Z = 6, 6, 6
    cfg = tiny_config(X, Y, Z, flip_coeff=0.95)
    h = cfg.grid_physical_size / Y
    rng = np.random.default_rng(2)
    pos = (rng.uniform(1.2, 4.5, (50, 3)) * h).astype(np.float32)
    vp = rng.standard_normal((50, 3)).astype(np.float32)
    o = ob.Oracle(cfg, fp64=True, pos=pos, vel=vp, init=False)
    ix, iy, iz = np.meshgrid(np.arange(X + 1), np.arange(Y + 1), np.arange(Z + 1), indexing="ij")

    def lin(a, b, c, d, x, y, z):
        return a + b * x + c * y + d * z
    # face positions in cell units: u at (i, j+.5, k+.5), v at (i+.5, j, k+.5), w at (i+.5, j+.5, k)
    new = np.stack([lin(0.1, 0.2, -0.3, 0.4, ix, iy + .5, iz + .5),
                    lin(-0.2, 0.1, 0.5, -0.1, ix + .5, iy, iz + .5),
                    lin(0.3, -0.2, 0.1, 0.2, ix + .5, iy + .5, iz), 0 * ix], -1)
    old = 0.5 * new
    o.set_field("new_velocity", new)
    o.set_field("velocity", old)
    o.set_field("new_volume_fractions", np.tile([1.0, 0, 0, 0], (X, Y, Z, 1)))
    o.run_stage("g2p")
    q = pos.astype(np.float64) / h
    at = np.stack([lin(0.1, 0.2, -0.3, 0.4, *q.T), lin(-0.2, 0.1, 0.5, -0.1, *q.T),
                   lin(0.3, -0.2, 0.1, 0.2, *q.T)], -1)
    a = np.float64(np.float32(0.95))
    want = a * (vp + (at - 0.5 * at)) + (1 - a) * at
    np.testing.assert_allclose(o.get_particles()[1], want, rtol=1e-10, atol=1e-12)


def test_rk3_uniform_field_and_wall_bounce():
    X = Y = Z = 8
    cfg = tiny_config(X, Y, Z, dt=0.01)
    h = cfg.grid_physical_size / Y
    pos = np.array([[4 * h, 4 * h, 4 * h], [0.3 * h, 4 * h, 4 * h]], np.float32)
    vel = np.array([[1.0, 2.0, 3.0], [-5.0, 0, 0]], np.float32)
    o = ob.Oracle(cfg, fp64=True, pos=pos, vel=vel, init=False)
    u = np.array([1.0, -2.0, 0.5])
    nv = np.zeros((X + 1, Y + 1, Z + 1, 4))
    nv[..., :3] = u
    o.set_field("new_velocity", nv)
    o.run_stage("advect")
    p, v, _ = o.get_particles()
    dt = np.float64(np.float32(0.01))
    # 2/9 + 3/9 + 4/9 = 1: a uniform field moves the particle by dt*u
    np.testing.assert_allclose(p[0], pos[0].astype(np.float64) + dt * u, rtol=1e-13)
    assert np.array_equal(v[0], vel[0])
    # second particle is pushed below dx/4 in x: clamped and its x velocity reflected
    nv[..., :3] = [-40.0, 0, 0]
    o2 = ob.Oracle(cfg, fp64=True, pos=pos, vel=vel, init=False)
    o2.set_field("new_velocity", nv)
    o2.run_stage("advect")
    p2, v2, _ = o2.get_particles()
    assert p2[1, 0] == pytest.approx(h / 4, rel=1e-15)
    assert v2[1, 0] == 5.0


# ---------------------------------------------------------------- invariants and determinism
def test_steps_keep_invariants_multiphase():
    cfg = configs.make_config(20, dt=0.01, pressure_iters=40, diffusion_iters=20, radius=0.15,
                              sphere_phase=1, phase_count=2, seed=11)
    o = ob.Oracle(cfg)
    n0 = o.particle_count
    o.step(5)
    pos, vel, frac = o.get_particles()
    h = cfg.grid_physical_size / 20
    assert o.particle_count == n0
    assert pos.min() >= h / 4 - 1e-6 and pos.max() <= 10 - h / 4 + 1e-6
    np.testing.assert_allclose(frac.sum(1), 1.0, rtol=1e-5)
    assert np.isfinite(vel).all()
    f = o.get_field("new_volume_fractions")
    c = o.get_field("content")[:20, :20, :20]
    np.testing.assert_allclose(f[c == FLUID].sum(-1), 1.0, rtol=1e-5)
    assert o.physical_time == pytest.approx(5 * 0.01, rel=1e-6)


def test_oracle_is_thread_count_independent():
    cfg = configs.ball_drop_16()
    res = []
    for t in (1, 4):
        ob.set_threads(t)
        o = ob.Oracle(cfg)
        o.step(3)
        res.append((o.get_particles(), o.get_field("pressure"), o.get_field("new_velocity")))
    ob.set_threads(0 or 8)
    (pa, pp, nv), (pb_, pp2, nv2) = res
    for a, b in zip(pa, pb_):
        assert np.array_equal(a, b)
    assert np.array_equal(pp, pp2) and np.array_equal(nv, nv2)


def test_fp32_against_fp64_one_step():
    """Every FP32 result is stated against the FP64 run of the same step."""
    cfg = configs.ball_drop_16()
    pos, frac = ob.seed(cfg)
    o32 = ob.Oracle(cfg, pos=pos, frac4=frac)
    o64 = ob.Oracle(cfg, fp64=True, pos=pos, frac4=frac)
    o32.step(1)
    o64.step(1)
    for name in ("pressure", "new_velocity"):
        a, b = o32.get_field(name), o64.get_field(name)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 1e-4, (name, rel)


# ---------------------------------------------------------------- P2G fractions and density
def brute_fractions(pos, frac, X, Y, Z, h, rho):
    """transferVolumeFractionsToCell (Fluid/Fluid_3D_FLIP.cu:118-165) by brute force in
    FP64: each cell averages the fractions of its OWN particles with trilinear weights
    to the cell centre, renormalises them to sum 1 and takes density = sum rho_k f_k;
    cells without particles get zero fractions and density 0."""
    own = (pos.astype(np.float32) / np.float32(h)).astype(np.int64)  # the hash's truncation
    f = np.zeros((X, Y, Z, 4))
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                sel = np.all(own == (x, y, z), axis=1)
                if not sel.any():
                    continue
                r = (pos[sel].astype(np.float64) - (np.array([x, y, z]) + 0.5) * h) / h
                w = hat(r[:, 0]) * hat(r[:, 1]) * hat(r[:, 2])
                avg = (w[:, None] * frac[sel]).sum(0) / w.sum()
                f[x, y, z] = avg / avg.sum()
    return f, f @ np.asarray(rho, np.float64)


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 3, 5)])
def test_p2g_fractions_and_density_match_brute_force(shape):
    X, Y, Z = shape
    rho = (1.0, 2.5, 0.5, 4.0)
    cfg = tiny_config(X, Y, Z, phase_count=4)
    for a in range(4):
        cfg.phase_densities[a] = rho[a]
    h = cfg.grid_physical_size / Y
    rng = np.random.default_rng(8)
    n = 70
    # positions kept 5 % away from cell walls so the own cell is unambiguous
    cell = rng.integers(0, [X, Y, Z], size=(n, 3))
    pos = ((cell + 0.05 + 0.9 * rng.random((n, 3))) * h).astype(np.float32)
    frac = np.zeros((n, 4), np.float32)
    onehot = rng.random(n) < 0.5
    frac[onehot, rng.integers(0, 4, onehot.sum())] = 1.0
    # fractional rows, deliberately not summing to 1 (the cell result is renormalised)
    frac[~onehot] = rng.random(((~onehot).sum(), 4)).astype(np.float32) * 0.8
    o = ob.Oracle(cfg, fp64=True, pos=pos, frac4=frac, init=False)
    o.run_stage("transfer_to_grid")
    want_f, want_rho = brute_fractions(pos, frac.astype(np.float64), X, Y, Z, h, rho)
    for name in ("volume_fractions", "new_volume_fractions"):
        np.testing.assert_allclose(o.get_field(name), want_f, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(o.get_field("density"), want_rho, rtol=1e-12, atol=1e-14)
    assert (want_rho == 0).any() and (want_rho > 0).any()  # empty and occupied cells both present


def test_p2g_fractions_two_particle_hand_case():
    """One particle at the cell centre (weight 1, phase 0) and one offset by h/4 on every
    axis (weight 0.75^3 = 0.421875, phase 1): f = (1, 0.421875) / 1.421875 and
    density = (1*1 + 2.5*0.421875) / 1.421875 with phase densities (1, 2.5)."""
    cfg = tiny_config(2, 2, 2, phase_count=2)
    cfg.phase_densities[0], cfg.phase_densities[1] = 1.0, 2.5
    h = cfg.grid_physical_size / 2  # 5.0
    pos = np.array([[1.5 * h, 0.5 * h, 0.5 * h], [1.75 * h, 0.75 * h, 0.75 * h]], np.float32)
    frac = np.array([[1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    o = ob.Oracle(cfg, fp64=True, pos=pos, frac4=frac, init=False)
    o.run_stage("transfer_to_grid")
    f = o.get_field("volume_fractions")
    np.testing.assert_allclose(f[1, 0, 0], [1 / 1.421875, 0.421875 / 1.421875, 0, 0], rtol=1e-15)
    assert o.get_field("density")[1, 0, 0] == pytest.approx(2.0546875 / 1.421875, rel=1e-15)
    # every other cell is empty: zero fractions, density 0
    d = o.get_field("density")
    d[1, 0, 0] = 0
    assert np.all(d == 0) and np.all(f.reshape(-1, 4)[[0, 1, 2, 3, 5, 6, 7]] == 0)


# ---------------------------------------------------------------- projection
def test_projection_density_branches_hand_case():
    """updateVelocityWithPressureImpl (Fluid/Fluid_3D_kernels.cu:118-207) on a hand-built
    3^3 case. Cells (x, y, z) with (content, density, pressure):
      (0,1,1) FLUID 2 3;  (1,1,1) FLUID 4 1;  (2,1,1) AIR 0 0;
      (1,0,1) SOLID 7 5;  (0,1,0) AIR 9 0.5;  (2,0,1) AIR 0 0.
    Expected face updates (u -= (p - p_lower_neighbour) / rho_used):
      u(1,1,1): FLUID|FLUID, rho (2+4)/2 = 3 -> u + 2/3
      u(2,1,1): AIR|FLUID(left), rho_left = 4 -> u + 1/4
      v(1,1,1): FLUID over SOLID, rho_this = 4 -> v + 1      (the SOLID's pressure is read)
      w(0,1,1): FLUID over AIR, rho_this = 2 -> w - 5/4       (the AIR's density is ignored)
      v(2,1,1): AIR over AIR -> unchanged, hasVelocity 0
    hasVelocity is overwritten on every cell with 'face touches FLUID'."""
    X = Y = Z = 3
    cfg = tiny_config(X, Y, Z)
    o = ob.Oracle(cfg, fp64=True, init=False)
    content = np.full((4, 4, 4), AIR, np.int32)
    content[3, :, :] = SOLID
    content[:, 3, :] = SOLID
    content[:, :, 3] = SOLID
    dens = np.zeros((3, 3, 3))
    p = np.zeros((3, 3, 3))
    for c, (k, d, pr) in {(0, 1, 1): (FLUID, 2, 3), (1, 1, 1): (FLUID, 4, 1), (2, 1, 1): (AIR, 0, 0),
                          (1, 0, 1): (SOLID, 7, 5), (0, 1, 0): (AIR, 9, 0.5),
                          (2, 0, 1): (AIR, 0, 0)}.items():
        content[c], dens[c], p[c] = k, d, pr
    rng = np.random.default_rng(12)
    nv = rng.standard_normal((4, 4, 4, 4))
    o.set_field("content", content)
    o.set_field("density", dens)
    o.set_field("pressure", p)
    o.set_field("new_velocity", nv)
    o.set_field("has_velocity", np.ones((3, 3, 3, 4)))
    o.run_stage("project")
    got = o.get_field("new_velocity")
    assert got[1, 1, 1, 0] == pytest.approx(nv[1, 1, 1, 0] + 2 / 3, rel=1e-15)
    assert got[2, 1, 1, 0] == pytest.approx(nv[2, 1, 1, 0] + 0.25, rel=1e-15)
    assert got[1, 1, 1, 1] == pytest.approx(nv[1, 1, 1, 1] + 1.0, rel=1e-15)
    assert got[0, 1, 1, 2] == pytest.approx(nv[0, 1, 1, 2] - 1.25, rel=1e-15)
    assert got[2, 1, 1, 1] == nv[2, 1, 1, 1]
    # faces not touching FLUID keep their value; the 4th channel is untouched
    fl = content[:3, :3, :3] == FLUID
    touch = np.zeros((3, 3, 3, 3), bool)
    touch[1:, :, :, 0] = fl[1:] | fl[:-1]
    touch[:, 1:, :, 1] = fl[:, 1:] | fl[:, :-1]
    touch[:, :, 1:, 2] = fl[:, :, 1:] | fl[:, :, :-1]
    hv = o.get_field("has_velocity")
    assert np.array_equal(hv[..., :3].astype(bool), touch) and np.all(hv[..., 3] == 0)
    assert touch.sum() == 10  # 6 faces of (1,1,1), 4 of (0,1,1) (its x face is at x = 0)
    assert np.array_equal(got[:3, :3, :3, :3][~touch], nv[:3, :3, :3, :3][~touch])
    assert np.array_equal(got[..., 3], nv[..., 3])
    # the face-extent padding layer (index 3) is outside the launch and unchanged
    assert np.array_equal(got[3], nv[3]) and np.array_equal(got[:, 3], nv[:, 3])


# ---------------------------------------------------------------- G2P fraction read
def test_g2p_reads_fractions_of_the_particles_own_cell():
    """transferToParticlesImpl (Fluid/Fluid_3D_FLIP.cu:208-224): a particle takes
    newVolumeFractions of cell floor(p/h) -- no interpolation -- renormalised to sum 1
    when the sum is positive. The field varies from cell to cell, so an interpolated or
    rounded read would give other values; the old volumeFractions field is not read."""
    X = Y = Z = 4
    cfg = tiny_config(X, Y, Z, phase_count=2)
    h = cfg.grid_physical_size / Y  # 2.5, exact
    cells = np.array([[1, 2, 1], [1, 2, 1], [2, 0, 3], [3, 3, 0], [0, 1, 2], [2, 2, 2]])
    offs = np.array([[0.9, 0.9, 0.9], [0.1, 0.5, 0.05], [0.5, 0.95, 0.5], [0.99, 0.01, 0.6],
                     [0.0, 0.0, 0.0], [0.3, 0.7, 0.2]])  # (0,1,2)+0 lies exactly on cell walls
    pos = ((cells + offs) * h).astype(np.float32)
    assert np.array_equal(np.floor(pos.astype(np.float64) / h).astype(int), cells)
    nvf = np.zeros((X, Y, Z, 4))
    idx = np.arange(X * Y * Z).reshape(X, Y, Z)
    nvf[..., 0] = idx + 1.0
    nvf[..., 1] = 2.0
    nvf[3, 3, 0] = 0.0  # a zero-sum cell: fractions copied unnormalised (all zero)
    o = ob.Oracle(cfg, fp64=True, pos=pos, init=False)
    o.set_field("new_volume_fractions", nvf)
    o.set_field("volume_fractions", np.full((X, Y, Z, 4), 0.25))
    o.run_stage("g2p")
    _, _, f = o.get_particles()
    # the oracle may keep particles in another order than given: match by position
    gp, _, _ = o.get_particles()
    order = [int(np.argmin(np.abs(gp - q).sum(1))) for q in pos.astype(np.float64)]
    for k, c in enumerate(map(tuple, cells)):
        s = nvf[c].sum()
        want = nvf[c] / s if s > 0 else nvf[c]
        np.testing.assert_allclose(f[order[k]], want, rtol=1e-15, atol=0)
    assert np.all(f[order[3]] == 0)
    np.testing.assert_allclose(f[order[0]], [26 / 28, 2 / 28, 0, 0], rtol=1e-15)  # cell (1,2,1): index 25
