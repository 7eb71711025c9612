"""ctypes binding of libmtlp.so (include/mtlp.h).

The library is loaded from this package directory only; a missing library is
an error (there is no fallback implementation).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmtlp.so")

MTLP_OK = 0
STATUS_NAMES = {0: "MTLP_OK", 1: "MTLP_E_INVALID", 2: "MTLP_E_RANGE", 3: "MTLP_E_NOMEM",
                4: "MTLP_E_HIP", 5: "MTLP_E_DEVICE"}

SCENE_CUBE_DROP = 0
SCENE_DAM_BREAK = 1
SCENE_SPHERE_DROP = 2
SCENE_JITTER_CUBE = 3
SCENE_JITTER_DAM = 4
SCENE_EMPTY = 255

BUF_POSITIONS = 0
BUF_VELOCITIES = 1
BUF_PREV_POSITIONS = 2
BUF_CELL_HASHES = 3
BUF_CELL_START = 4
BUF_SORTED_INDEX = 5
BUF_NEIGHBOR_COUNTS = 6
BUF_LAMBDAS = 7
BUF_SORTED_POSITIONS = 8
BUF_FOAM_INTENSITY = 9

TAPE_POSITIONS = 0
TAPE_VELOCITIES = 1
TAPE_SORTED_INDEX = 2
TAPE_CELL_START = 3
TAPE_ITERATE = 4
TAPE_GRAD_SUMS = 5
TAPE_DELTA = 6

MAX_SCAN_LENGTH = 0x7FFFFFFF
MAX_PARTICLES = 0x40000000


class Params(ctypes.Structure):
    """mtlp_params (include/mtlp.h)."""
    _fields_ = [
        ("gravity", ctypes.c_float * 3),
        ("bounds_min", ctypes.c_float * 3),
        ("bounds_max", ctypes.c_float * 3),
        ("h", ctypes.c_float),
        ("cell_size", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("scorr_k", ctypes.c_float),
        ("scorr_dq", ctypes.c_float),
        ("max_step", ctypes.c_float),
        ("damping", ctypes.c_float),
        ("wall_eps", ctypes.c_float),
        ("floor_friction", ctypes.c_float),
        ("wall_friction", ctypes.c_float),
        ("vmax", ctypes.c_float),
        ("scorr_n", ctypes.c_uint32),
        ("iterations", ctypes.c_uint32),
    ]

    def copy(self):
        c = Params()
        ctypes.pointer(c)[0] = self
        return c

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


# Exported symbols and their signatures: (restype, argtypes).
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_F = ctypes.c_float
SIGNATURES = {
    "mtlp_status_string": (ctypes.c_char_p, [_I]),
    "mtlp_last_error": (ctypes.c_char_p, []),
    "mtlp_default_params": (_I, [ctypes.POINTER(Params)]),
    "mtlp_scene_params": (_I, [_U32, _U32, ctypes.POINTER(Params)]),
    "mtlp_generate_scene": (_I, [_U32, _U32, ctypes.c_uint64, ctypes.POINTER(Params), _P, _P]),
    "mtlp_grid_dims": (_I, [ctypes.POINTER(Params), ctypes.POINTER(_U32)]),
    "mtlp_create": (_I, [_I, ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    "mtlp_destroy": (_I, [_P]),
    "mtlp_reset": (_I, [_P, _U32, _U32, ctypes.c_uint64]),
    "mtlp_set_params": (_I, [_P, ctypes.POINTER(Params)]),
    "mtlp_get_params": (_I, [_P, ctypes.POINTER(Params)]),
    "mtlp_upload_state": (_I, [_P, _P, _P, _U32]),
    "mtlp_step": (_I, [_P, _P, _F]),
    "mtlp_build_spatial_hash": (_I, [_P, _P]),
    "mtlp_scan_scratch_bytes": (ctypes.c_size_t, [_U32]),
    "mtlp_exclusive_scan": (_I, [_P, _P, _P, _P, _P, _U32]),
    "mtlp_exclusive_scan_noalloc": (_I, [_P, _P, _P, _U32]),
    "mtlp_set_tile_pull": (_I, [_P, _U32]),
    "mtlp_set_packed_positions": (_I, [_P, ctypes.c_int]),
    "mtlp_launch_config": (_I, [_P, ctypes.POINTER(_U32), ctypes.POINTER(_U32), ctypes.POINTER(ctypes.c_int)]),
    "mtlp_get_buffer": (_I, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_size_t)]),
    "mtlp_read_buffer": (_I, [_P, _I, _P, ctypes.c_size_t]),
    "mtlp_write_buffer": (_I, [_P, _I, _P, ctypes.c_size_t]),
    "mtlp_particle_count": (_U32, [_P]),
    "mtlp_total_cells": (_U32, [_P]),
    "mtlp_cell_size": (_F, [_P]),
    "mtlp_set_profiling": (_I, [_P, _I]),
    "mtlp_stage_times": (_I, [_P, ctypes.POINTER(_F), ctypes.POINTER(ctypes.c_char_p), _U32,
                              ctypes.POINTER(_U32)]),
    "mtlp_error_flags": (_I, [_P, ctypes.POINTER(_U32)]),
    "mtlp_batch_create": (_I, [_I, _U32, ctypes.POINTER(_U32), ctypes.POINTER(Params),
                               ctypes.POINTER(_P)]),
    "mtlp_batch_seed": (_I, [_P, _U32, ctypes.c_uint64]),
    "mtlp_batch_info": (_I, [_P, ctypes.POINTER(_U32), ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "mtlp_batch_set_params": (_I, [_P, _U32, ctypes.POINTER(Params)]),
    "mtlp_step_n": (_I, [_P, _P, _F, _U32]),
    "mtlp_set_graph": (_I, [_P, _I]),
    "mtlp_autotune": (_I, [_P, _P, _F, ctypes.c_char_p, ctypes.POINTER(_U32)]),
    "mtlp_set_stencil_block": (_I, [_P, _U32]),
    "mtlp_set_neighbor_lists": (_I, [_P, _I]),
    "mtlp_save_state": (_I, [_P, ctypes.c_char_p]),
    "mtlp_load_state": (_I, [_P, ctypes.c_char_p]),
    "mtlp_integrate": (_I, [_P, _P, _F]),
    "mtlp_apply_mouse_pulse": (_I, [_P, _P, ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F), _F]),
    "mtlp_set_sphere": (_I, [_P, _I, ctypes.POINTER(_F), ctypes.POINTER(_F), _F]),
    "mtlp_get_sphere": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F)]),
    "mtlp_set_state_device": (_I, [_P, _P, _P, _P]),
    "mtlp_get_state_device": (_I, [_P, _P, _P, _P]),
    "mtlp_rollout": (_I, [_P, _P, _F, _U32, _U32]),
    "mtlp_adjoint_rollout": (_I, [_P, _P, _F, _U32, _U32, _P, _P, _P, _P, _P]),
    "mtlp_density": (_I, [_P, _P, _P, _P]),
    "mtlp_density_vjp": (_I, [_P, _P, _P, _P, _P]),
    "mtlp_tape_record": (_I, [_P, _P, _F, _U32]),
    "mtlp_tape_buffer": (_I, [_P, _I, _U32, _U32, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_size_t)]),
    "mtlp_tape_read": (_I, [_P, _I, _U32, _U32, _P, ctypes.c_size_t]),
}

_lib = None


class MtlpError(RuntimeError):
    def __init__(self, status, where, message=""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{where}: {name}" + (f" ({message})" if message else ""))


def _share_torch_hip_runtime():
    """A process must hold one HIP runtime.  torch's wheel carries its own
    libamdhip64.so (SONAME libamdhip64.so.7, like /opt/rocm's); loading torch
    first makes the dynamic loader resolve libmtlp.so's libamdhip64.so.7 to
    that same copy.  Loaded the other way round, the process would hold two
    runtimes and the second finds no device."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libmtlp.so, failing loudly if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m mtlphys_amd.build` (or __graft_entry__.build())")
        _share_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != MTLP_OK:
        msg = lib().mtlp_last_error()
        raise MtlpError(status, where, msg.decode() if msg else "")
    return status


def default_params():
    p = Params()
    check(lib().mtlp_default_params(ctypes.byref(p)), "mtlp_default_params")
    return p


def scene_params(scene, n):
    p = Params()
    check(lib().mtlp_scene_params(scene, n, ctypes.byref(p)), "mtlp_scene_params")
    return p


def grid_dims(params):
    dims = (_U32 * 3)()
    check(lib().mtlp_grid_dims(ctypes.byref(params), dims), "mtlp_grid_dims")
    return tuple(dims)
