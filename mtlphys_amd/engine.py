"""Python mirror of the reference engine interface (engine/include/mtlphys/Engine.hpp).

`Engine` wraps the C ABI of libmtlp.so with the reference's member names:
reset, step, build_spatial_hash, run_exclusive_scan, the buffer accessors,
particle_count, total_cells and cell_size.  Every GPU computation runs in the
HIP kernels of libmtlp.so; this module only moves handles and host arrays.

Streams: methods take `stream=None` (the null stream), an integer hipStream_t
handle, or any object with a `.cuda_stream` attribute (a torch stream).
"""
import ctypes

import numpy as np

from . import _capi as C

# buffer id -> (numpy dtype, components per particle / cell)
_BUFFER_LAYOUT = {
    C.BUF_POSITIONS: (np.float32, 4),
    C.BUF_VELOCITIES: (np.float32, 4),
    C.BUF_PREV_POSITIONS: (np.float32, 4),
    C.BUF_CELL_HASHES: (np.uint32, 1),
    C.BUF_CELL_START: (np.uint32, 1),
    C.BUF_SORTED_INDEX: (np.uint32, 1),
    C.BUF_NEIGHBOR_COUNTS: (np.uint32, 1),
    C.BUF_LAMBDAS: (np.float32, 1),
    C.BUF_SORTED_POSITIONS: (np.float32, 4),
    C.BUF_FOAM_INTENSITY: (np.float32, 1),
}


def _stream_handle(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(int(stream.cuda_stream))
    return ctypes.c_void_p(int(stream))


def _ptr(arr):
    return arr.ctypes.data_as(ctypes.c_void_p)


def _dev(p):
    """A device address: None, an int, or an object with data_ptr() (a torch tensor)."""
    if p is None:
        return None
    if hasattr(p, "data_ptr"):
        return ctypes.c_void_p(int(p.data_ptr()))
    return ctypes.c_void_p(int(p))


def generate_scene(scene, n, seed=0, params=None):
    """Host-side scene seeding (Engine::seedParticles, Engine.cpp:454-549).

    Returns (x4, v4) as float32 arrays of shape (n, 4).  Needs no GPU."""
    p = params if params is not None else C.scene_params(scene, n)
    x = np.zeros((n, 4), np.float32)
    v = np.zeros((n, 4), np.float32)
    C.check(C.lib().mtlp_generate_scene(scene, n, seed, ctypes.byref(p), _ptr(x), _ptr(v)),
            "mtlp_generate_scene")
    return x, v


class DeviceBuffer:
    """Borrowed view of one engine buffer (the reference's MTL::Buffer* accessors)."""

    def __init__(self, engine, which):
        self._engine = engine
        self.which = which
        ptr = ctypes.c_void_p()
        nbytes = ctypes.c_size_t()
        C.check(C.lib().mtlp_get_buffer(engine._h, which, ctypes.byref(ptr), ctypes.byref(nbytes)),
                "mtlp_get_buffer")
        self.ptr = ptr.value or 0
        self.nbytes = nbytes.value
        self.dtype, self.components = _BUFFER_LAYOUT[which]

    def read(self):
        """Synchronous device->host copy (tests/TestUtils.hpp:82-97)."""
        itemsize = np.dtype(self.dtype).itemsize * self.components
        count = self.nbytes // itemsize
        shape = (count, self.components) if self.components > 1 else (count,)
        out = np.empty(shape, self.dtype)
        C.check(C.lib().mtlp_read_buffer(self._engine._h, self.which, _ptr(out), out.nbytes),
                "mtlp_read_buffer")
        return out

    def write(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        C.check(C.lib().mtlp_write_buffer(self._engine._h, self.which, _ptr(arr), arr.nbytes),
                "mtlp_write_buffer")


class Engine:
    """mtlphys::Engine over libmtlp.so (Engine.hpp:27-100)."""

    kSceneCubeDrop = C.SCENE_CUBE_DROP
    kSceneDamBreak = C.SCENE_DAM_BREAK
    kSceneSphereDrop = C.SCENE_SPHERE_DROP

    def __init__(self, device=0, params=None):
        self._h = None
        p = params if params is not None else C.default_params()
        h = ctypes.c_void_p()
        C.check(C.lib().mtlp_create(int(device), ctypes.byref(p), ctypes.byref(h)), "mtlp_create")
        self._h = h
        self._device = int(device)
        self._scene = C.SCENE_CUBE_DROP

    def close(self):
        if self._h is not None and self._h.value:
            C.check(C.lib().mtlp_destroy(self._h), "mtlp_destroy")
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- Engine.hpp:41-56 -------------------------------------------------
    def reset(self, particle_count, scene=None, seed=0):
        """Engine::reset(n[, scene]) (Engine.cpp:125-134)."""
        if scene is None:
            scene = self._scene
        C.check(C.lib().mtlp_reset(self._h, int(particle_count), int(scene), int(seed)), "mtlp_reset")
        self._scene = scene

    def current_scene(self):
        return self._scene

    def step(self, dt, stream=None):
        """Engine::step(cmd, dt) (Engine.cpp:685-811); asynchronous."""
        C.check(C.lib().mtlp_step(self._h, _stream_handle(stream), ctypes.c_float(dt)), "mtlp_step")

    def integrate(self, dt, stream=None):
        """Engine::integrate (Engine.cpp:578-595): legacy semi-implicit Euler."""
        C.check(C.lib().mtlp_integrate(self._h, _stream_handle(stream), ctypes.c_float(dt)), "mtlp_integrate")

    def save_state(self, path):
        """Write positions, velocities, parameters and the sphere to `path`."""
        C.check(C.lib().mtlp_save_state(self._h, str(path).encode()), "mtlp_save_state")

    def load_state(self, path):
        """Read a state file written by save_state or mtlphys_amd.state.write_state."""
        C.check(C.lib().mtlp_load_state(self._h, str(path).encode()), "mtlp_load_state")

    def apply_mouse_pulse(self, ray_origin, ray_dir, force, radius, stream=None):
        """Engine::applyMousePulse (Engine.cpp:813-841); asynchronous."""
        f3 = ctypes.c_float * 3
        C.check(C.lib().mtlp_apply_mouse_pulse(self._h, _stream_handle(stream), f3(*ray_origin), f3(*ray_dir),
                                               f3(*force), ctypes.c_float(radius)), "mtlp_apply_mouse_pulse")

    def set_sphere(self, enabled=True, center=None, velocity=None, radius=0.0):
        f3 = ctypes.c_float * 3
        C.check(C.lib().mtlp_set_sphere(self._h, 1 if enabled else 0, None if center is None else f3(*center),
                                        None if velocity is None else f3(*velocity), ctypes.c_float(radius)),
                "mtlp_set_sphere")

    def sphere(self):
        """(enabled, center, velocity, radius) of the kinematic sphere."""
        on = ctypes.c_int()
        c, v, r = (ctypes.c_float * 3)(), (ctypes.c_float * 3)(), ctypes.c_float()
        C.check(C.lib().mtlp_get_sphere(self._h, ctypes.byref(on), c, v, ctypes.byref(r)), "mtlp_get_sphere")
        return bool(on.value), list(c), list(v), r.value

    def build_spatial_hash(self, stream=None):
        """Engine::buildSpatialHash(cmd) (Engine.cpp:601-679); asynchronous."""
        C.check(C.lib().mtlp_build_spatial_hash(self._h, _stream_handle(stream)),
                "mtlp_build_spatial_hash")

    def run_exclusive_scan(self, d_in, d_out, d_scratch, n, stream=None):
        """Engine::runExclusiveScan (Engine.cpp:843-882) on device pointers."""
        C.check(C.lib().mtlp_exclusive_scan(self._h, _stream_handle(stream), ctypes.c_void_p(d_in),
                                            ctypes.c_void_p(d_out), ctypes.c_void_p(d_scratch), n),
                "mtlp_exclusive_scan")

    @staticmethod
    def scan_scratch_bytes(n):
        return C.lib().mtlp_scan_scratch_bytes(int(n))

    # ---- parameters and state ----------------------------------------------
    @property
    def params(self):
        p = C.Params()
        C.check(C.lib().mtlp_get_params(self._h, ctypes.byref(p)), "mtlp_get_params")
        return p

    @params.setter
    def params(self, p):
        C.check(C.lib().mtlp_set_params(self._h, ctypes.byref(p)), "mtlp_set_params")

    def upload_state(self, x4=None, v4=None):
        n = self.particle_count
        xa = None if x4 is None else np.ascontiguousarray(x4, np.float32).reshape(n, 4)
        va = None if v4 is None else np.ascontiguousarray(v4, np.float32).reshape(n, 4)
        C.check(C.lib().mtlp_upload_state(self._h, None if xa is None else _ptr(xa),
                                          None if va is None else _ptr(va), n), "mtlp_upload_state")

    # ---- Engine.hpp:78-90 ---------------------------------------------------
    @property
    def particle_count(self):
        return C.lib().mtlp_particle_count(self._h)

    @property
    def total_cells(self):
        return C.lib().mtlp_total_cells(self._h)

    @property
    def cell_size(self):
        return C.lib().mtlp_cell_size(self._h)

    def buffer(self, which):
        return DeviceBuffer(self, which)

    def tensor(self, which):
        """Zero-copy torch tensor over an engine buffer (DLPack): (n, 4)
        float32 for positions/velocities, (n,) otherwise.  It aliases the
        engine's memory, so steps update it in place; it is valid until
        reset() or close()."""
        import torch
        from . import dlpack
        buf = self.buffer(which)
        dtype, comps = _BUFFER_LAYOUT[which]
        count = buf.nbytes // (np.dtype(dtype).itemsize * comps)
        shape = (count, comps) if comps > 1 else (count,)
        cap = dlpack.capsule(buf.ptr, shape, dtype, dlpack.kDLROCM, self._device, owner=self)
        return torch.utils.dlpack.from_dlpack(cap)

    def positions_buffer(self):
        return self.buffer(C.BUF_POSITIONS)

    def velocities_buffer(self):
        return self.buffer(C.BUF_VELOCITIES)

    def prev_positions_buffer(self):
        return self.buffer(C.BUF_PREV_POSITIONS)

    def cell_hashes_buffer(self):
        return self.buffer(C.BUF_CELL_HASHES)

    def cell_start_buffer(self):
        return self.buffer(C.BUF_CELL_START)

    def sorted_index_buffer(self):
        return self.buffer(C.BUF_SORTED_INDEX)

    def neighbor_counts_buffer(self):
        return self.buffer(C.BUF_NEIGHBOR_COUNTS)

    def lambdas_buffer(self):
        return self.buffer(C.BUF_LAMBDAS)

    def sorted_positions_buffer(self):
        return self.buffer(C.BUF_SORTED_POSITIONS)

    def foam_intensity_buffer(self):
        return self.buffer(C.BUF_FOAM_INTENSITY)

    # ---- diagnostics ------------------------------------------------------
    def set_profiling(self, on=True):
        C.check(C.lib().mtlp_set_profiling(self._h, 1 if on else 0), "mtlp_set_profiling")

    def stage_times(self):
        """[(kernel name, ms)] of the last step/build when profiling is on."""
        cap = 1024
        ms = (ctypes.c_float * cap)()
        names = (ctypes.c_char_p * cap)()
        count = ctypes.c_uint32()
        C.check(C.lib().mtlp_stage_times(self._h, ms, names, cap, ctypes.byref(count)),
                "mtlp_stage_times")
        return [(names[i].decode(), ms[i]) for i in range(count.value)]

    def error_flags(self):
        f = ctypes.c_uint32()
        C.check(C.lib().mtlp_error_flags(self._h, ctypes.byref(f)), "mtlp_error_flags")
        return f.value

    # ---- stepping options -------------------------------------------------
    def step_n(self, dt, n_steps, stream=None):
        C.check(C.lib().mtlp_step_n(self._h, _stream_handle(stream), ctypes.c_float(dt), int(n_steps)),
                "mtlp_step_n")

    def set_graph(self, on=True):
        """Capture each step into a HIP graph (once per dt) and replay it."""
        C.check(C.lib().mtlp_set_graph(self._h, 1 if on else 0), "mtlp_set_graph")

    def autotune(self, dt, cache_path=None, stream=None):
        """Pick the stencil workgroup size by timing; returns it."""
        b = ctypes.c_uint32()
        C.check(C.lib().mtlp_autotune(self._h, _stream_handle(stream), ctypes.c_float(dt),
                                      None if cache_path is None else str(cache_path).encode(),
                                      ctypes.byref(b)), "mtlp_autotune")
        return b.value

    def set_stencil_block(self, block):
        C.check(C.lib().mtlp_set_stencil_block(self._h, int(block)), "mtlp_set_stencil_block")

    def set_tile_pull(self, tiles_per_pull):
        """0: one workgroup per tile; k >= 1: persistent workgroups taking k
        tiles at a time from per-XCD counters (same results)."""
        C.check(C.lib().mtlp_set_tile_pull(self._h, int(tiles_per_pull)), "mtlp_set_tile_pull")

    def set_packed_positions(self, on=True):
        """Density pass filters its candidates from a packed (x, y, z) copy of
        each iterate (default on; same results)."""
        C.check(C.lib().mtlp_set_packed_positions(self._h, 1 if on else 0), "mtlp_set_packed_positions")

    def launch_config(self):
        """(stencil workgroup size, tiles per pull, packed positions)."""
        b, t, p = C._U32(), C._U32(), ctypes.c_int()
        C.check(C.lib().mtlp_launch_config(self._h, ctypes.byref(b), ctypes.byref(t), ctypes.byref(p)),
                "mtlp_launch_config")
        return b.value, t.value, bool(p.value)

    def set_neighbor_lists(self, on=True):
        """Density pass records neighbour lists for the delta pass (default on)."""
        C.check(C.lib().mtlp_set_neighbor_lists(self._h, 1 if on else 0), "mtlp_set_neighbor_lists")

    # ---- device-resident state (pointers are device addresses, n float4) ---
    def set_state_device(self, d_x4=None, d_v4=None, stream=None):
        C.check(C.lib().mtlp_set_state_device(self._h, _stream_handle(stream), _dev(d_x4), _dev(d_v4)),
                "mtlp_set_state_device")

    def get_state_device(self, d_x4=None, d_v4=None, stream=None):
        C.check(C.lib().mtlp_get_state_device(self._h, _stream_handle(stream), _dev(d_x4), _dev(d_v4)),
                "mtlp_get_state_device")

    # ---- differentiable rollout ---------------------------------------------
    def rollout(self, dt, steps, ckpt_every, stream=None):
        """`steps` steps keeping (x, v) every ckpt_every steps for the adjoint."""
        C.check(C.lib().mtlp_rollout(self._h, _stream_handle(stream), ctypes.c_float(dt), int(steps),
                                     int(ckpt_every)), "mtlp_rollout")

    def adjoint_rollout(self, dt, steps, ckpt_every, d_xbar_T=None, d_vbar_T=None, d_xbar_0=None,
                        d_vbar_0=None, d_gbar=None, stream=None):
        """Adjoint of the last rollout: (dL/dx_T, dL/dv_T) -> dL/dx_0, dL/dv_0,
        dL/dgravity (device pointers; None = 0 in, not written out)."""
        C.check(C.lib().mtlp_adjoint_rollout(self._h, _stream_handle(stream), ctypes.c_float(dt), int(steps),
                                             int(ckpt_every), _dev(d_xbar_T), _dev(d_vbar_T),
                                             _dev(d_xbar_0), _dev(d_vbar_0), _dev(d_gbar)),
                "mtlp_adjoint_rollout")

    def density(self, d_x4, d_rho, stream=None):
        """rho at the n float4 positions d_x4 (fresh spatial hash) into d_rho."""
        C.check(C.lib().mtlp_density(self._h, _stream_handle(stream), _dev(d_x4), _dev(d_rho)),
                "mtlp_density")

    def density_vjp(self, d_x4, d_rhobar, d_xbar, stream=None):
        C.check(C.lib().mtlp_density_vjp(self._h, _stream_handle(stream), _dev(d_x4), _dev(d_rhobar),
                                         _dev(d_xbar)), "mtlp_density_vjp")

    def tape_record(self, dt, steps, stream=None):
        """`steps` taped steps from the current state, which is left as it is."""
        C.check(C.lib().mtlp_tape_record(self._h, _stream_handle(stream), ctypes.c_float(dt), int(steps)),
                "mtlp_tape_record")

    def tape_read(self, kind, step, iteration=0):
        """Host copy of one tape array (synchronous)."""
        ptr = ctypes.c_void_p()
        nbytes = ctypes.c_size_t()
        C.check(C.lib().mtlp_tape_buffer(self._h, int(kind), int(step), int(iteration), ctypes.byref(ptr),
                                         ctypes.byref(nbytes)), "mtlp_tape_buffer")
        u32 = kind in (C.TAPE_SORTED_INDEX, C.TAPE_CELL_START)
        dtype = np.uint32 if u32 else np.float32
        out = np.empty(nbytes.value // 4, dtype)
        C.check(C.lib().mtlp_tape_read(self._h, int(kind), int(step), int(iteration), _ptr(out), out.nbytes),
                "mtlp_tape_read")
        return out if u32 else out.reshape(-1, 4)


class Batch(Engine):
    """Many independent scenes stepped together, one kernel launch per stage.

    Scene s owns particles [particle_offsets[s], particle_offsets[s+1]) of
    every per-particle buffer, in original and in sorted order.  Each scene
    has its own parameters (bounds/grid, rest density, gravity, ...); all
    share the iteration count."""

    def __init__(self, counts, params, device=0):
        self._h = None
        counts = [int(c) for c in counts]
        if len(params) != len(counts):
            raise ValueError("one parameter set per scene")
        arr_c = (ctypes.c_uint32 * len(counts))(*counts)
        arr_p = (C.Params * len(params))(*params)
        h = ctypes.c_void_p()
        C.check(C.lib().mtlp_batch_create(int(device), len(counts), arr_c, arr_p, ctypes.byref(h)),
                "mtlp_batch_create")
        self._h = h
        self._device = int(device)
        self._scene = C.SCENE_EMPTY
        self.counts = counts
        self.scene_params = [p.copy() for p in params]

    def seed(self, scene, seed=0):
        """Seeds scene s with generate_scene(scene, counts[s], seed + s, params[s])."""
        C.check(C.lib().mtlp_batch_seed(self._h, int(scene), int(seed)), "mtlp_batch_seed")
        self._scene = scene

    def offsets(self):
        ns = ctypes.c_uint32()
        C.check(C.lib().mtlp_batch_info(self._h, ctypes.byref(ns), None, None), "mtlp_batch_info")
        po = (ctypes.c_uint32 * (ns.value + 1))()
        co = (ctypes.c_uint32 * (ns.value + 1))()
        C.check(C.lib().mtlp_batch_info(self._h, None, po, co), "mtlp_batch_info")
        return np.array(po, np.uint32), np.array(co, np.uint32)

    def set_scene_params(self, scene, p):
        C.check(C.lib().mtlp_batch_set_params(self._h, int(scene), ctypes.byref(p)),
                "mtlp_batch_set_params")
        self.scene_params[scene] = p.copy()
