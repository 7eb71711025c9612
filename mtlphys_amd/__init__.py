"""mtlphys_amd -- MI355X (gfx950) Position-Based Fluids step.

The PBF step of rakdcolon/mtlphys (predict, spatial hash, density/lambda,
delta, finalize) as hand-written HIP kernels behind a C ABI
(include/mtlp.h, built into mtlphys_amd/libmtlp.so), with a Python mirror of
the reference's Engine interface.
"""
from . import _capi as capi
from ._capi import MtlpError, Params, default_params, grid_dims, scene_params
from .engine import Batch, DeviceBuffer, Engine, generate_scene

__all__ = ["capi", "Engine", "Batch", "DeviceBuffer", "MtlpError", "Params", "default_params",
           "scene_params", "grid_dims", "generate_scene"]
