"""Checkpoint and resume (aq_flip_save / aq_flip_load): a run interrupted by a
save/load continues bit for bit like an uninterrupted one, exports after a load
equal those before the save, and malformed files are rejected with AQ_EINVAL."""
import numpy as np
import pytest

import aquarius_amd as aq
from aquarius_amd import configs
from aquarius_amd.fluid import FlipHandle

from test_gpu_parity import log

pytestmark = pytest.mark.gpu

FIELDS = ("content", "velocity", "new_velocity", "pressure", "divergence", "density", "particle_count",
          "has_velocity", "volume_fractions", "new_volume_fractions")


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def snapshot_state(h):
    out = {f: h.get_field(f) for f in FIELDS}
    for k, v in zip(("pos", "vel", "frac"), h.get_particles()):
        out[k] = v
    out["cells"] = h.get_cells()
    out["t"] = h.physical_time
    return out


@pytest.mark.parametrize("name", ["multiphase_24", "ball_drop_64_previous"])
def test_resume_is_bit_identical(tmp_path, name):
    if name == "multiphase_24":
        cfg = configs.make_config(24, dt=0.01, pressure_iters=40, diffusion_iters=10, radius=0.15,
                                  sphere_phase=1, phase_count=2, seed=21)
    else:
        cfg = configs.ball_drop_64()
        cfg.pressure_init = aq.configs.AQ_PRESSURE_PREVIOUS
    a = FlipHandle(cfg)
    a.steps(4)
    path = tmp_path / "state.aqsnap"
    a.save(path)
    before = snapshot_state(a)
    a.steps(3)
    want = snapshot_state(a)
    a.close()
    b = FlipHandle.load(path)
    log(f"[{name}] loaded {b.particle_count} particles at t={b.physical_time:.3f}")
    got = snapshot_state(b)
    for k in before:
        if k == "cells":
            assert all(same(x, y) for x, y in zip(got[k], before[k]))
        elif k == "t":
            assert got[k] == before[k]
        else:
            assert same(got[k], before[k]), k
    b.steps(3)
    got = snapshot_state(b)
    for k in want:
        if k == "cells":
            assert all(same(x, y) for x, y in zip(got[k], want[k]))
        elif k == "t":
            assert got[k] == want[k]
        else:
            assert same(got[k], want[k]), k
    b.close()


def test_oracle_step_from_a_restored_state_matches(tmp_path):
    """A snapshot restored into a new handle is handed to the CPU oracle (particles in
    their stored order, every grid field); one step of each from there agrees bit for
    bit on the sort and within the parity bar on every float (observed: identical)."""
    from oracle import binding as ob
    from test_gpu_parity import check_cells, check_field, check_particles
    cfg = configs.make_config(24, dt=0.01, pressure_iters=40, diffusion_iters=10, radius=0.15,
                              sphere_phase=1, phase_count=2, seed=21)
    cfg.pressure_init = aq.configs.AQ_PRESSURE_PREVIOUS  # the restored pressure is the first guess
    a = FlipHandle(cfg)
    a.steps(6)
    path = tmp_path / "state.aqsnap"
    a.save(path)
    a.close()
    b = FlipHandle.load(path)
    pos, vel, frac = b.get_particles()
    o = ob.Oracle(cfg, pos=pos, vel=vel, frac4=frac, init=False)
    for f in FIELDS:
        o.set_field(f, b.get_field(f))
    log(f"[snapshot -> oracle] {b.particle_count} particles at t={b.physical_time:.3f}")
    b.step()
    o.step()
    check_cells(b, o, "restored+1 ")
    check_particles(b, o, "restored+1 ")
    for f in FIELDS:
        check_field(f, b, o, "restored+1 ")
    b.close()


def test_bad_snapshots_are_rejected(tmp_path):
    bad = tmp_path / "bad.aqsnap"
    bad.write_bytes(b"NOTASNAP" + bytes(64))
    with pytest.raises(Exception, match="not a libaqflip snapshot"):
        FlipHandle.load(bad)
    h = FlipHandle(configs.ball_drop_16())
    good = tmp_path / "good.aqsnap"
    h.save(good)
    h.close()
    data = good.read_bytes()
    (tmp_path / "short.aqsnap").write_bytes(data[: len(data) // 2])
    with pytest.raises(Exception, match="truncated|corrupt"):
        FlipHandle.load(tmp_path / "short.aqsnap")
    v2 = bytearray(data)
    v2[8] = 99  # version
    (tmp_path / "v99.aqsnap").write_bytes(bytes(v2))
    with pytest.raises(Exception, match="version"):
        FlipHandle.load(tmp_path / "v99.aqsnap")
