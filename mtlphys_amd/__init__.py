"""mtlphys_amd -- MI355X (gfx950) Position-Based Fluids step.

The PBF step of rakdcolon/mtlphys (predict, spatial hash, density/lambda,
delta, finalize, foam, the sphere-drop scene's sphere, the mouse pulse) as
hand-written HIP kernels behind a C ABI (include/mtlp.h, built into
mtlphys_amd/libmtlp.so), with a Python mirror of the reference's Engine
interface, batched ensembles, the step's adjoint (autograd.py) and a state
file format (state.py).
"""
from . import _capi as capi
from ._capi import MtlpError, Params, default_params, grid_dims, scene_params
from .engine import Batch, DeviceBuffer, Engine, generate_scene
from .state import read_state, write_state

__all__ = ["capi", "Engine", "Batch", "DeviceBuffer", "MtlpError", "Params", "default_params",
           "scene_params", "grid_dims", "generate_scene", "read_state", "write_state"]
