"""Particle seeding on the GPU (aq_flip_create) against the host seeding of the same
scene (aq_flip_seed_particles, which the oracle's seeding pins bit for bit): the
initial particle set, its sorted order and the initial grid must be identical."""
import numpy as np
import pytest

import aquarius_amd as aq
from aquarius_amd import configs
from aquarius_amd.fluid import FlipHandle, seed_particles

from test_gpu_parity import log

pytestmark = pytest.mark.gpu


def scenes():
    yield "ball_drop_16", configs.ball_drop_16()
    yield "ball_drop_64", configs.ball_drop_64()
    yield "multiphase_128", configs.multiphase_128()
    yield "edge_12x16x20_4ph", configs.make_config(0, dt=0.01, pressure_iters=5, extent=(12, 16, 20),
                                                   sphere_phase=3, phase_count=4, seed=5)
    cfg = configs.make_config(40, dt=0.01, pressure_iters=5, pool_depth=0.35, radius=0.3,
                              center=(0.2, 0.75, 0.6), seed=0xDEADBEEF12345678)
    cfg.particle_jitter = 1.0
    yield "jitter1_offcentre_40", cfg


@pytest.mark.parametrize("name,cfg", list(scenes()), ids=lambda v: v if isinstance(v, str) else "")
def test_gpu_seeding_equals_host_seeding(name, cfg):
    pos, frac = seed_particles(cfg)
    log(f"[{name}] {len(pos)} particles")
    g = FlipHandle(cfg)                          # seeded on the device
    hst = FlipHandle(cfg, pos=pos, frac4=frac)   # seeded on the host, uploaded
    assert g.particle_count == hst.particle_count == len(pos)
    for a, b in zip(g.get_particles(), hst.get_particles()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for f in ("content", "velocity", "new_volume_fractions", "density"):
        assert np.array_equal(g.get_field(f).view(np.uint32), hst.get_field(f).view(np.uint32)), f
    g.close()
    hst.close()
