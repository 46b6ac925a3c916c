"""libfluid_amd -- MI355X-native weakly compressible SPH timestep.

The solver is libfluid_amd/lib/libfluid_hip.so (HIP kernels for gfx950 behind
the C ABI of include/lf_abi.h); this package is its Python binding:

  Scene        one device-resident scene (lf_create / lf_step / lf_download)
  Batch        many independent scenes stepped together (lf_batch_*)
  serve        continuous-batching server of scene jobs (lf_server_*)
  FluidSystem  mirror of the reference's FluidSystem host API
  configs      seeded synthetic scenes (dam break, column, spray, double dam)
"""
from . import _lib, configs, serve
from ._lib import LIB_PATH, LfError, LfParams, device_count
from .fluidsystem import FluidSystem
from .scene import Batch, Scene, default_params, fluid_block, jittered_lattice, make_params, uniform_points, write_geo

__all__ = [
    "LIB_PATH", "LfError", "LfParams", "device_count", "FluidSystem", "Scene", "Batch", "configs", "serve",
    "default_params", "fluid_block", "jittered_lattice", "make_params", "uniform_points", "write_geo", "_lib",
]
