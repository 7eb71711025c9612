"""numpy front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (mtlphys_amd) never
imports it.  See pbf_oracle.h for what the oracle restates and what pins it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcParams(ctypes.Structure):
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
        ("sphere", ctypes.c_float * 4),
        ("sphere_on", ctypes.c_uint32),
    ]


class OrcGrid(ctypes.Structure):
    _fields_ = [
        ("origin", ctypes.c_float * 3),
        ("cell_size", ctypes.c_float),
        ("inv_cell", ctypes.c_float),
        ("dim", ctypes.c_uint32 * 3),
        ("total_cells", ctypes.c_uint32),
    ]


def _trace_struct(real):
    P = ctypes.POINTER
    return type("Trace", (ctypes.Structure,), {"_fields_": [
        ("keys", P(ctypes.c_uint32)), ("cell_start", P(ctypes.c_uint32)),
        ("sorted_index", P(ctypes.c_uint32)), ("nbr_count", P(ctypes.c_uint32)),
        ("xstar4", P(ctypes.c_float)), ("lambda_iter", P(real)), ("sx_iter", P(real)),
        ("lambda_orig", P(real))]})


class OrcTape(ctypes.Structure):
    _fields_ = [("sorted_index", ctypes.c_void_p), ("cell_start", ctypes.c_void_p),
                ("p", ctypes.c_void_p), ("x_in", ctypes.c_void_p)]


TraceF64 = _trace_struct(ctypes.c_double)
TraceF32 = _trace_struct(ctypes.c_float)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `python oracle/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        u32 = ctypes.c_uint32
        PP = ctypes.POINTER(OrcParams)
        PG = ctypes.POINTER(OrcGrid)
        sig = {
            "orc_default_params": (None, [PP]),
            "orc_grid_make": (None, [PP, PG]),
            "orc_predict_f32": (None, [vp, vp, u32, vp, ctypes.c_float, vp]),
            "orc_cell_key": (u32, [vp, PG]),
            "orc_hash": (None, [vp, u32, PG, vp]),
            "orc_count": (None, [vp, u32, u32, vp]),
            "orc_exclusive_scan": (None, [vp, vp, u32]),
            "orc_stable_scatter": (None, [vp, u32, vp, vp]),
            "orc_build_spatial_hash": (None, [vp, u32, PG, vp, vp, vp, vp, ctypes.c_float, vp]),
            "orc_density_lambda_f64": (None, [vp, vp, vp, u32, PG, PP, vp, vp, vp]),
            "orc_density_lambda_f32": (None, [vp, vp, vp, u32, PG, PP, vp, vp, vp]),
            "orc_apply_delta_f64": (None, [vp, vp, vp, u32, PG, PP, vp]),
            "orc_apply_delta_f32": (None, [vp, vp, vp, u32, PG, PP, vp]),
            "orc_finalize_f64": (None, [vp, vp, vp, u32, PP, ctypes.c_float]),
            "orc_finalize_f32": (None, [vp, vp, vp, u32, PP, ctypes.c_float]),
            "orc_step_f64": (ctypes.c_int, [vp, vp, u32, PP, ctypes.c_float, vp, vp,
                                            ctypes.POINTER(TraceF64)]),
            "orc_step_f32": (ctypes.c_int, [vp, vp, u32, PP, ctypes.c_float, vp, vp,
                                            ctypes.POINTER(TraceF32)]),
            "orc_step_d": (ctypes.c_int, [vp, vp, u32, PP, ctypes.c_float, vp, vp,
                                          ctypes.POINTER(OrcTape)]),
            "orc_step_vjp_d": (ctypes.c_int, [ctypes.POINTER(OrcTape), u32, PP, ctypes.c_float,
                                              vp, vp, vp, vp, vp, vp]),
            "orc_step_kinks_d": (None, [ctypes.POINTER(OrcTape), u32, PP, ctypes.c_float, vp]),
            "orc_density_at_d": (None, [vp, u32, PP, vp]),
            "orc_density_vjp_d": (None, [vp, u32, PP, vp, vp]),
            "orc_sphere_advance": (None, [vp, PP, ctypes.c_float]),
            "orc_foam_f64": (None, [vp, vp, u32, ctypes.c_double, vp]),
            "orc_mouse_pulse_f64": (None, [vp, vp, u32, vp, vp, vp, ctypes.c_double]),
            "orc_integrate_f64": (None, [vp, vp, u32, PP, ctypes.c_double]),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_get_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def params(src=None, **overrides):
    """OrcParams from the reference defaults, a ctypes struct with the same
    field names (e.g. mtlphys_amd.Params), or a dict; then overrides."""
    p = OrcParams()
    lib().orc_default_params(ctypes.byref(p))
    items = {}
    if src is not None:
        if isinstance(src, dict):
            items.update(src)
        else:
            items.update({name: getattr(src, name) for name, _ in OrcParams._fields_ if hasattr(src, name)})
    items.update(overrides)
    for k, v in items.items():
        if hasattr(v, "__len__"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def grid(p):
    g = OrcGrid()
    lib().orc_grid_make(ctypes.byref(p), ctypes.byref(g))
    return g


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def f32x4(a):
    a = np.ascontiguousarray(a, np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def predict(x4, v4, p, dt):
    x4, v4 = f32x4(x4), f32x4(v4)
    out = np.empty_like(x4)
    g = (ctypes.c_float * 3)(*p.gravity)
    lib().orc_predict_f32(_p(x4), _p(v4), len(x4), g, ctypes.c_float(dt), _p(out))
    return out


def cell_key(pt, g):
    a = np.ascontiguousarray(pt, np.float32)
    return lib().orc_cell_key(_p(a), ctypes.byref(g))


def hash_keys(p4, g):
    p4 = f32x4(p4)
    out = np.empty(len(p4), np.uint32)
    lib().orc_hash(_p(p4), len(p4), ctypes.byref(g), _p(out))
    return out


def count(keys, total_cells):
    keys = np.ascontiguousarray(keys, np.uint32)
    out = np.empty(total_cells + 1, np.uint32)
    lib().orc_count(_p(keys), len(keys), total_cells, _p(out))
    return out


def exclusive_scan(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib().orc_exclusive_scan(_p(a), _p(out), len(a))
    return out


def stable_scatter(keys, cell_start):
    keys = np.ascontiguousarray(keys, np.uint32)
    cs = np.ascontiguousarray(cell_start, np.uint32)
    out = np.empty(len(keys), np.uint32)
    lib().orc_stable_scatter(_p(keys), len(keys), _p(cs), _p(out))
    return out


def build_spatial_hash(p4, g, r2max):
    """-> dict(keys, cell_start, sorted_index, sorted4, nbr)"""
    p4 = f32x4(p4)
    n = len(p4)
    keys = np.empty(n, np.uint32)
    cs = np.empty(g.total_cells + 1, np.uint32)
    sidx = np.empty(n, np.uint32)
    s4 = np.empty((n, 4), np.float32)
    nbr = np.empty(n, np.uint32)
    lib().orc_build_spatial_hash(_p(p4), n, ctypes.byref(g), _p(keys), _p(cs), _p(sidx), _p(s4),
                                 ctypes.c_float(r2max), _p(nbr))
    return dict(keys=keys, cell_start=cs, sorted_index=sidx, sorted4=s4, nbr=nbr)


def _real(prec):
    return (np.float64, "f64") if prec == 64 else (np.float32, "f32")


def density_lambda(sx4, cell_start, sorted_index, g, p, prec=64):
    dt, sfx = _real(prec)
    sx4 = np.ascontiguousarray(sx4, dt)
    n = len(sx4)
    lam_s = np.empty(n, dt)
    lam_o = np.empty(n, dt)
    nbr = np.empty(n, np.uint32)
    getattr(lib(), "orc_density_lambda_" + sfx)(
        _p(sx4), _p(np.ascontiguousarray(cell_start, np.uint32)),
        _p(np.ascontiguousarray(sorted_index, np.uint32)), n, ctypes.byref(g), ctypes.byref(p),
        _p(lam_s), _p(lam_o), _p(nbr))
    return lam_s, lam_o, nbr


def apply_delta(sx4, lam_sorted, cell_start, g, p, prec=64):
    dt, sfx = _real(prec)
    sx4 = np.ascontiguousarray(sx4, dt)
    out = np.empty_like(sx4)
    getattr(lib(), "orc_apply_delta_" + sfx)(
        _p(sx4), _p(np.ascontiguousarray(lam_sorted, dt)),
        _p(np.ascontiguousarray(cell_start, np.uint32)), len(sx4), ctypes.byref(g),
        ctypes.byref(p), _p(out))
    return out


def finalize(x4, v4, prev4, p, dt, prec=64):
    rdt, sfx = _real(prec)
    x = np.array(x4, rdt, copy=True, order="C")
    v = np.array(v4, rdt, copy=True, order="C")
    prev = np.ascontiguousarray(prev4, rdt)
    getattr(lib(), "orc_finalize_" + sfx)(_p(x), _p(v), _p(prev), len(x), ctypes.byref(p),
                                          ctypes.c_float(dt))
    return x, v


def step(x4, v4, p, dt, prec=64, trace=False):
    """Engine::step on the oracle.  Returns (x4, v4[, trace dict])."""
    rdt, sfx = _real(prec)
    x4, v4 = f32x4(x4), f32x4(v4)
    n = len(x4)
    xo = np.empty((n, 4), rdt)
    vo = np.empty((n, 4), rdt)
    tr = None
    arrays = None
    if trace:
        g = grid(p)
        it = p.iterations
        arrays = dict(keys=np.empty(n, np.uint32), cell_start=np.empty(g.total_cells + 1, np.uint32),
                      sorted_index=np.empty(n, np.uint32), nbr_count=np.empty(n, np.uint32),
                      xstar4=np.empty((n, 4), np.float32), lambda_iter=np.empty((it, n), rdt),
                      sx_iter=np.empty((it + 1, n, 4), rdt), lambda_orig=np.empty(n, rdt))
        T = TraceF64 if prec == 64 else TraceF32
        kw = {}
        for name, arr in arrays.items():
            ctype = dict(T._fields_)[name]._type_
            kw[name] = arr.ctypes.data_as(ctypes.POINTER(ctype))
        tr = T(**kw)
    getattr(lib(), "orc_step_" + sfx)(_p(x4), _p(v4), n, ctypes.byref(p), ctypes.c_float(dt),
                                      _p(xo), _p(vo), ctypes.byref(tr) if tr is not None else None)
    if trace:
        return xo, vo, arrays
    return xo, vo


# ---- FP64 rollout step and adjoint ------------------------------------------

def f64x4(a):
    a = np.ascontiguousarray(a, np.float64)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def step_d(x4, v4, p, dt):
    """FP64 step (predict in FP64, binning at the FP32 rounding of x*).
    Returns (x4, v4, tape); the tape feeds step_vjp_d."""
    x4, v4 = f64x4(x4), f64x4(v4)
    n = len(x4)
    g = grid(p)
    tape = dict(sorted_index=np.empty(n, np.uint32), cell_start=np.empty(g.total_cells + 1, np.uint32),
                p=np.empty((p.iterations + 1, n, 4), np.float64), x_in=np.empty((n, 4), np.float64))
    xo = np.empty((n, 4), np.float64)
    vo = np.empty((n, 4), np.float64)
    t = OrcTape(*(_p(tape[k]) for k in ("sorted_index", "cell_start", "p", "x_in")))
    lib().orc_step_d(_p(x4), _p(v4), n, ctypes.byref(p), ctypes.c_float(dt), _p(xo), _p(vo),
                     ctypes.byref(t))
    return xo, vo, tape


def step_vjp_d(tape, p, dt, xbar_out, vbar_out, sums=None):
    """(xbar_out, vbar_out) -> (xbar_in, vbar_in, gbar[3]) at the tape.
    sums: None (gbar from the direct sum of pbar^0) or a float64 array of
    the 6 running sums (sum xbar, sum vbar), updated in place."""
    xb, vb = f64x4(xbar_out), f64x4(vbar_out)
    n = len(xb)
    xi = np.empty((n, 4), np.float64)
    vi = np.empty((n, 4), np.float64)
    gb = np.zeros(3, np.float64)
    if sums is not None:
        assert sums.dtype == np.float64 and sums.shape == (6,) and sums.flags.c_contiguous
    t = OrcTape(*(_p(tape[k]) for k in ("sorted_index", "cell_start", "p", "x_in")))
    lib().orc_step_vjp_d(ctypes.byref(t), n, ctypes.byref(p), ctypes.c_float(dt), _p(xb), _p(vb),
                         _p(xi), _p(vi), _p(gb), _p(sums))
    return xi, vi, gb


def adjoint_sums(xbar, vbar):
    """Initial running sums (sum xbar.xyz, sum vbar.xyz) for step_vjp_d."""
    return np.concatenate([f64x4(xbar)[:, :3].sum(axis=0), f64x4(vbar)[:, :3].sum(axis=0)])


def step_kinks_d(tape, p, dt):
    """(n, 2 + iterations) uint64 branch signature of a taped step."""
    n = len(tape["sorted_index"])
    out = np.zeros((n, 2 + p.iterations), np.uint64)
    t = OrcTape(*(_p(tape[k]) for k in ("sorted_index", "cell_start", "p", "x_in")))
    lib().orc_step_kinks_d(ctypes.byref(t), n, ctypes.byref(p), ctypes.c_float(dt), _p(out))
    return out


def rollout_kinks(x0, v0, p, dt, steps):
    """Branch signatures of every step of a rollout, stacked: (steps, n, 2 + I)."""
    x, v = f64x4(x0), f64x4(v0)
    sig = []
    for _ in range(steps):
        xn, vn, t = step_d(x, v, p, dt)
        sig.append(step_kinks_d(t, p, dt))
        x, v = xn, vn
    return np.stack(sig)


def density_at_d(x4, p):
    x4 = f64x4(x4)
    rho = np.empty(len(x4), np.float64)
    lib().orc_density_at_d(_p(x4), len(x4), ctypes.byref(p), _p(rho))
    return rho


def density_vjp_d(x4, p, rhobar):
    x4 = f64x4(x4)
    out = np.empty_like(x4)
    lib().orc_density_vjp_d(_p(x4), len(x4), ctypes.byref(p), _p(np.ascontiguousarray(rhobar, np.float64)),
                            _p(out))
    return out


def rollout_loss(x0, v0, p, dt, steps, target, w_density=0.1):
    """Loss L = |centroid(x_T) - target|^2 + w mean((rho(x_T)/rho0 - 1)^2)
    of a `steps`-step FP64 rollout.  Returns (L, tapes, x_T)."""
    x, v = f64x4(x0), f64x4(v0)
    tapes = []
    for _ in range(steps):
        x, v, t = step_d(x, v, p, dt)
        tapes.append(t)
    n = len(x)
    c = x[:, :3].mean(axis=0) - np.asarray(target, np.float64)
    rho = density_at_d(x, p)
    dev = rho / p.rest_density - 1.0
    L = float(c @ c + w_density * np.mean(dev * dev))
    return L, tapes, x


def rollout_grad(x0, v0, p, dt, steps, target, w_density=0.1, exact_sums=True):
    """dL/dx0, dL/dv0 (n x 4, w = 0) and dL/dgravity (3) by the adjoint
    (gravity through the running-sum recursion unless exact_sums=False)."""
    L, tapes, xT = rollout_loss(x0, v0, p, dt, steps, target, w_density)
    n = len(xT)
    c = xT[:, :3].mean(axis=0) - np.asarray(target, np.float64)
    rho = density_at_d(xT, p)
    dev = rho / p.rest_density - 1.0
    xb = np.zeros((n, 4))
    xb[:, :3] = 2.0 * c / n
    xb += density_vjp_d(xT, p, w_density * 2.0 * dev / (n * p.rest_density))
    vb = np.zeros((n, 4))
    gb = np.zeros(3)
    sums = adjoint_sums(xb, vb) if exact_sums else None
    for t in reversed(tapes):
        xb, vb, g1 = step_vjp_d(t, p, dt, xb, vb, sums)
        gb += g1
    return L, xb, vb, gb


# ---- scene extras -------------------------------------------------------------

def sphere_advance(state, p, dt):
    """One step of the sphere kinematics; state = [cx, cy, cz, vx, vy, vz] (float32)."""
    st = np.ascontiguousarray(state, np.float32).copy()
    lib().orc_sphere_advance(_p(st), ctypes.byref(p), ctypes.c_float(dt))
    return st


def foam(v4, nbr, max_nbr=96.0):
    v4 = f64x4(v4)
    out = np.empty(len(v4), np.float64)
    lib().orc_foam_f64(_p(v4), _p(np.ascontiguousarray(nbr, np.uint32)), len(v4), ctypes.c_double(max_nbr), _p(out))
    return out


def mouse_pulse(x4, v4, origin, direction, force, radius):
    x4 = f64x4(x4)
    v = np.array(v4, np.float64, copy=True, order="C")
    d3 = lambda a: np.ascontiguousarray(a, np.float64)
    lib().orc_mouse_pulse_f64(_p(x4), _p(v), len(x4), _p(d3(origin)), _p(d3(direction)), _p(d3(force)),
                              ctypes.c_double(radius))
    return v


def integrate(x4, v4, p, dt):
    """Legacy semi-implicit Euler step; returns new (x4, v4) in FP64."""
    x = np.array(x4, np.float64, copy=True, order="C")
    v = np.array(v4, np.float64, copy=True, order="C")
    lib().orc_integrate_f64(_p(x), _p(v), len(x), ctypes.byref(p), ctypes.c_double(dt))
    return x, v
