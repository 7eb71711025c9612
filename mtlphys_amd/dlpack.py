"""Zero-copy DLPack views of device (or host) memory, for handing engine
buffers to torch without a copy: `Engine.tensor(buffer_id)`.

A view is a DLManagedTensor (DLPack's legacy "dltensor" capsule) built with
ctypes around an address the caller guarantees alive: the engine's buffers
live until mtlp_reset / mtlp_destroy, and the view keeps a reference to its
owner until the consumer calls the deleter.
"""
import ctypes
import threading

import numpy as np

kDLCPU = 1
kDLROCM = 10

_DTYPES = {np.dtype(np.float32): (2, 32), np.dtype(np.uint32): (1, 32), np.dtype(np.int32): (0, 32),
           np.dtype(np.float64): (2, 64)}


class DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class DLManagedTensor(ctypes.Structure):
    pass


_Deleter = ctypes.CFUNCTYPE(None, ctypes.POINTER(DLManagedTensor))
DLManagedTensor._fields_ = [("dl_tensor", DLTensor), ("manager_ctx", ctypes.c_void_p), ("deleter", _Deleter)]

_live = {}            # key -> (managed tensor, shape array, owner)
_lock = threading.Lock()
_next_key = [1]


@_Deleter
def _delete(ptr):
    # may run during interpreter teardown, after this module's globals are gone
    try:
        key = ptr.contents.manager_ctx
        with _lock:
            _live.pop(key, None)
    except Exception:
        pass


# Consumers keep the deleter's address in views that can outlive this
# module's globals at interpreter exit; an extra reference keeps the thunk.
ctypes.pythonapi.Py_IncRef(ctypes.py_object(_delete))


_PyCapsule_New = ctypes.pythonapi.PyCapsule_New
_PyCapsule_New.restype = ctypes.py_object
_PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]


def capsule(address, shape, dtype, device_type, device_id=0, owner=None):
    """A "dltensor" capsule viewing `address` as a C-contiguous array."""
    code, bits = _DTYPES[np.dtype(dtype)]
    shp = (ctypes.c_int64 * len(shape))(*[int(s) for s in shape])
    mt = DLManagedTensor()
    mt.dl_tensor.data = ctypes.c_void_p(int(address))
    mt.dl_tensor.device = DLDevice(device_type, device_id)
    mt.dl_tensor.ndim = len(shape)
    mt.dl_tensor.dtype = DLDataType(code, bits, 1)
    mt.dl_tensor.shape = shp
    mt.dl_tensor.strides = None  # compact row-major
    mt.dl_tensor.byte_offset = 0
    with _lock:
        key = _next_key[0]
        _next_key[0] += 1
        _live[key] = (mt, shp, owner)
    mt.manager_ctx = key
    mt.deleter = _delete
    return _PyCapsule_New(ctypes.addressof(mt), b"dltensor", None)


def live_views():
    """Number of views not yet released by their consumer."""
    with _lock:
        return len(_live)


__all__ = ["capsule", "live_views", "kDLCPU", "kDLROCM"]
