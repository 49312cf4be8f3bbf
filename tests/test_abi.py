"""The C ABI library and its host surfaces, checked without a GPU: every symbol
include/aqflip.h declares is exported, the ctypes mirror has the C layout,
bad configurations return error statuses, host seeding matches the oracle, and the
C++ drop-in surface compiles against the header."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import aquarius_amd as aq
from aquarius_amd import _abi, configs
from oracle import binding as ob

ROOT = Path(__file__).resolve().parent.parent
HEADERS = [ROOT / "include" / "aqflip.h", ROOT / "include" / "aqflip_unet.h",
           ROOT / "include" / "aqflip_hash.h"]


def declared_functions():
    names = set()
    for header in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(aq_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.PROTOTYPES), set(names) ^ set(_abi.PROTOTYPES)
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (aq_\w+)", out))
    assert set(names) <= exported


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "aqflip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(aq_flip_config), offsetof(aq_flip_config, seed),
         offsetof(aq_flip_config, volumes), sizeof(aq_flip_volume), sizeof(aq_flip_stats),
         offsetof(aq_flip_config, ref_boundary_coverage));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    vals = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                        check=True).stdout.split()))
    assert vals == [C.sizeof(_abi.Config), _abi.Config.seed.offset, _abi.Config.volumes.offset,
                    C.sizeof(_abi.Volume), C.sizeof(_abi.Stats),
                    _abi.Config.ref_boundary_coverage.offset]


def test_default_config_is_reference_default():
    """FluidConfig defaults (Fluid/FluidConfig.cuh:41-97)."""
    c = aq.default_config()
    assert list(c.size) == [50, 50, 50]
    assert (c.pressure_iters, c.diffusion_iters) == (100, 100)
    assert c.dt == pytest.approx(0.033) and c.flip_coeff == pytest.approx(0.95)
    assert c.particle_jitter == pytest.approx(0.05) and c.grid_physical_size == 10
    assert list(c.gravity) == pytest.approx([0, -9.8, 0]) and c.phase_count == 2
    assert c.diffusion_coeff == pytest.approx(1e-3)
    assert c.pressure_init == _abi.AQ_PRESSURE_PREVIOUS and c.ref_boundary_coverage == 1
    assert c.n_volumes == 2
    assert c.volumes[0].shape == _abi.AQ_SHAPE_BOX
    assert list(c.volumes[0].box_max) == pytest.approx([1, 0.1875, 1])
    assert c.volumes[1].shape == _abi.AQ_SHAPE_SPHERE
    assert list(c.volumes[1].center) == pytest.approx([0.5, 0.75, 0.5])
    assert c.volumes[1].radius == pytest.approx(0.125)


@pytest.mark.parametrize("mutate,msg", [
    (lambda c: setattr(c, "phase_count", 5), "Only supports up to 4 phases"),
    (lambda c: c.size.__setitem__(0, 0), "grid size"),
    (lambda c: setattr(c, "pressure_iters", -1), "iteration"),
    (lambda c: setattr(c, "n_volumes", 17), "n_volumes"),
    (lambda c: setattr(c, "pressure_init", 7), "pressure_init"),
])
def test_bad_config_returns_einval(mutate, msg):
    cfg = configs.ball_drop_16()
    mutate(cfg)
    lib = _abi.load()
    h = C.c_void_p()
    assert lib.aq_flip_create(C.byref(cfg), C.byref(h)) == _abi.AQ_EINVAL
    assert msg in lib.aq_last_error().decode()
    assert not h.value


def test_empty_scene_returns_einval():
    cfg = configs.ball_drop_16()
    cfg.n_volumes = 0
    lib = _abi.load()
    h = C.c_void_p()
    assert lib.aq_flip_create(C.byref(cfg), C.byref(h)) == _abi.AQ_EINVAL


def test_out_of_range_volumes_are_skipped_like_the_reference():
    """checkCoordValid (Fluid/Fluid_3D_FLIP.cu:542-556) silently skips the volume."""
    cfg = configs.ball_drop_16()
    cfg.volumes[1].center[1] = 1.5
    pos, _ = aq.seed_particles(cfg)
    assert len(pos) == 16 * 4 * 16 * 8  # only the pool remains


@pytest.mark.parametrize("scene", ["ball_drop_16", "ball_drop_64", "multiphase_128"])
def test_host_seeding_matches_oracle(scene):
    cfg = configs.SCENES[scene]()
    p1, f1 = aq.seed_particles(cfg)
    p2, f2 = ob.seed(cfg)
    assert np.array_equal(p1, p2) and np.array_equal(f1, f2)
    X, Y, Z = cfg.size
    h = np.float32(10.0) / np.float32(Y)
    assert p1.min() >= h / 4 and p1.max() <= np.float32(X) * h - h / 4
    # pool: 2x2x2 lattice in every cell below the pool depth
    pool = int(X * Y * 0.25) * Z * 8 if scene != "ball_drop_256" else None
    assert (f1[:, 0] == 1).sum() >= pool - (0 if scene != "multiphase_128" else 0)


def test_cpp_surface_compiles(tmp_path):
    """The drop-in Fluid_3D_FLIP::Fluid header builds against include/aqflip.h."""
    src = tmp_path / "use.cpp"
    src.write_text("""
#include "Fluid_3D_FLIP.hpp"
#include <memory>
int main() {
  FluidConfig config;
  config.FLIP.sizeX = 8;
  std::shared_ptr<Fluid_3D> fluid = std::make_shared<Fluid_3D_FLIP::Fluid>();
  aq_flip_config a = Fluid_3D_FLIP::toAbi(config);
  return a.size[0] == 8 && a.n_volumes == 2 ? 0 : 1;
}
""")
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}",
                    f"-I{ROOT / 'aquarius_amd' / 'csrc' / 'host'}", str(src), "-o", str(exe),
                    f"-L{_abi.LIB_PATH.parent}", "-laqflip", f"-Wl,-rpath,{_abi.LIB_PATH.parent}"],
                   check=True)
    assert subprocess.run([str(exe)]).returncode == 0
