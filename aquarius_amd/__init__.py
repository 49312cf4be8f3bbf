"""aquarius_amd: MI355X-native FLIP simulation step behind a C ABI (libaqflip).

The compute path is hand-written HIP for gfx950 (aquarius_amd/csrc); this package
holds the build script, the ctypes mirror of include/aqflip.h and a Python host
class with the reference's Fluid_3D_FLIP::Fluid surface.
"""
from ._abi import FIELDS, STAGES, AqError, Config, load  # noqa: F401
from .configs import (SCENES, ball_drop_16, ball_drop_64, ball_drop_256,  # noqa: F401
                      default_config, make_config, multiphase_128)
from .fluid import Fluid, FlipHandle, seed_particles  # noqa: F401

__version__ = "0.1.0"
