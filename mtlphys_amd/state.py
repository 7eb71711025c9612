"""The engine's state file (include/mtlp.h, "state I/O") read and written
with numpy alone, for fixtures and for moving states between machines
without a GPU.  Engine.save_state / Engine.load_state use the C ABI.

Layout (little-endian): b"MTLPSTAT", u32 version, n_scenes, n, flags (bit 0:
batch), scene id, sphere on, f32 sphere centre[3], velocity[3], radius; per
scene an mtlp_params (92 bytes) and a u32 count; float4 positions[n];
float4 velocities[n].
"""
import ctypes
import struct

import numpy as np

from . import _capi as C

MAGIC = b"MTLPSTAT"
VERSION = 1
_HEADER = struct.Struct("<8s6I7f")


def write_state(path, params, counts, x4, v4, scene=C.SCENE_EMPTY, batch=False, sphere=None):
    """params: list of Params (one per scene); counts: particles per scene;
    x4, v4: (n, 4) float32; sphere: None or (centre[3], velocity[3], radius)."""
    counts = [int(c) for c in counts]
    if len(params) != len(counts):
        raise ValueError("one parameter set per scene")
    x4 = np.ascontiguousarray(x4, np.float32).reshape(-1, 4)
    v4 = np.ascontiguousarray(v4, np.float32).reshape(-1, 4)
    n = sum(counts)
    if len(x4) != n or len(v4) != n:
        raise ValueError("state arrays must hold sum(counts) particles")
    c, vel, r = sphere if sphere is not None else ((0.0, 4.5, 0.0), (0.0, 0.0, 0.0), 0.6)
    head = _HEADER.pack(MAGIC, VERSION, len(counts), n, 1 if batch else 0, int(scene),
                        1 if sphere is not None else 0, *c, *vel, r)
    with open(path, "wb") as f:
        f.write(head)
        for p, cnt in zip(params, counts):
            f.write(bytes(p))
            f.write(struct.pack("<I", cnt))
        f.write(x4.tobytes())
        f.write(v4.tobytes())


def read_state(path):
    """-> dict(params, counts, x, v, scene, batch, sphere (None or tuple))."""
    with open(path, "rb") as f:
        data = f.read()
    magic, version, ns, n, flags, scene, sphere_on, *sph = _HEADER.unpack_from(data, 0)
    if magic != MAGIC:
        raise ValueError(f"{path} is not an mtlp state file")
    if version != VERSION:
        raise ValueError(f"unsupported state file version {version}")
    off = _HEADER.size
    params, counts = [], []
    psize = ctypes.sizeof(C.Params)
    for _ in range(ns):
        params.append(C.Params.from_buffer_copy(data[off:off + psize]))
        off += psize
        counts.append(struct.unpack_from("<I", data, off)[0])
        off += 4
    if sum(counts) != n:
        raise ValueError("state file counts do not add up")
    x = np.frombuffer(data, np.float32, 4 * n, off).reshape(n, 4).copy()
    off += 16 * n
    v = np.frombuffer(data, np.float32, 4 * n, off).reshape(n, 4).copy()
    sphere = (tuple(sph[0:3]), tuple(sph[3:6]), sph[6]) if sphere_on else None
    return dict(params=params, counts=counts, x=x, v=v, scene=scene, batch=bool(flags & 1), sphere=sphere)


__all__ = ["read_state", "write_state", "MAGIC", "VERSION"]
