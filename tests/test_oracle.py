"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against something that does not restate
it: the reference's own known answers (tests/test_spatial.cpp,
tests/test_scan.cpp, tests/test_pbf.cpp), numpy library routines (cumsum,
bincount, stable argsort), O(N^2) brute force, and closed forms of the
SPH kernels at hand-placed configurations.  A dropped term, a wrong sign, a
wrong index or a transposed operand fails at least one of them.
"""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def spiky_norm(h):
    return 45.0 / (math.pi * h ** 6)


def poly6_norm(h):
    return 315.0 / (64.0 * math.pi * h ** 9)


def W(r2, h):
    return poly6_norm(h) * (h * h - r2) ** 3 if r2 < h * h else 0.0


# ------------------------------------------------------------- grid/hash --
def test_default_params_are_reference_constants(orc):
    p = orc.params()
    assert list(p.bounds_min) == [-3, -3, -3] and list(p.bounds_max) == [3, 6, 3]
    assert np.float32(p.h) == np.float32(0.1) and np.float32(p.cell_size) == np.float32(0.1)
    assert p.rest_density == 8000 and p.epsilon == 800 and p.iterations == 5
    assert np.float32(p.gravity[1]) == np.float32(-9.81)
    g = orc.grid(p)
    assert list(g.dim) == [60, 90, 60] and g.total_cells == 324000  # Engine.cpp:398-403


def _np_hash(pts, origin, inv, dims):
    q = (pts[:, :3].astype(np.float32) - np.asarray(origin, np.float32)) * np.float32(inv)
    q = np.clip(q, np.float32(-1), np.asarray(dims, np.float32))
    c = np.clip(np.floor(q).astype(np.int64), 0, np.asarray(dims) - 1)
    return (c[:, 0] + c[:, 1] * dims[0] + c[:, 2] * dims[0] * dims[1]).astype(np.uint32)


def test_hash_reference_known_answers(orc):
    """tests/test_spatial.cpp:52-70 with cpuCellHash (TestUtils.cpp:30-42)."""
    pts = np.array([[0, 0, 0, 1], [2.5, 4.0, -1.5, 1], [-2.9, 5.5, 2.5, 1], [1.0, -2.0, -2.5, 1]], np.float32)
    g = orc.grid(orc.params())
    keys = orc.hash_keys(pts, g)
    # hand-derived FP32 cells: (30,30,30), (55,70,15), (0,85,55), (40,10,5);
    # x of the third point: (-2.9f + 3) * 10 = 0.99999905 -> cell 0
    expect = [30 + 30 * 60 + 30 * 5400, 55 + 70 * 60 + 15 * 5400, 0 + 85 * 60 + 55 * 5400,
              40 + 10 * 60 + 5 * 5400]
    assert keys.tolist() == expect
    np.testing.assert_array_equal(keys, _np_hash(pts, [-3, -3, -3], np.float32(1) / np.float32(0.1), [60, 90, 60]))


def test_hash_faces_and_clamps(orc):
    p = orc.params(bounds_min=[0, 0, 0], bounds_max=[1, 1, 1], cell_size=0.25, h=0.25)
    g = orc.grid(p)
    assert list(g.dim) == [4, 4, 4]
    pts = np.array([
        [0.25, 0.5, 0.75, 1],    # on faces: floor -> upper cell (1, 2, 3)
        [-5.0, 0.1, 0.1, 1],     # below the grid -> cell 0
        [7.0, 7.0, 7.0, 1],      # above -> last cell
        [np.inf, -np.inf, 0.3, 1],
        [np.nan, 0.3, 0.3, 1],   # NaN -> cell 0 on that axis
        [0.9999999, 0.0, 1.0, 1],
    ], np.float32)
    keys = orc.hash_keys(pts, g)
    assert keys.tolist() == [1 + 2 * 4 + 3 * 16, 0, 63, 3 + 0 + 16, 0 + 4 + 16, 3 + 0 + 3 * 16]


def test_hash_random_vs_numpy(orc):
    rng = np.random.default_rng(7)
    p = orc.params()
    g = orc.grid(p)
    pts = np.concatenate([rng.uniform(-4, 7, (20000, 3)), np.ones((20000, 1))], 1).astype(np.float32)
    np.testing.assert_array_equal(orc.hash_keys(pts, g), _np_hash(pts, g.origin, g.inv_cell, list(g.dim)))


# ------------------------------------------------------- count/scan/perm --
@pytest.mark.parametrize("kind,n,seed,maxv", [("zeros", 100, 0, 0), ("ones", 500, 0, 1), ("rand", 1024, 0xA, 7),
                                              ("rand", 700, 0xB, 7), ("rand", 2048, 0xC, 7),
                                              ("rand", 635626, 0xD, 3)])
def test_scan_reference_cases(orc, kind, n, seed, maxv):
    """tests/test_scan.cpp:59-110 (inputs from numpy's generator)."""
    rng = np.random.default_rng(seed)
    a = {"zeros": np.zeros(n, np.uint32), "ones": np.ones(n, np.uint32)}.get(
        kind, rng.integers(0, maxv, n, endpoint=True).astype(np.uint32))
    out = orc.exclusive_scan(a)
    ref = np.concatenate([[0], np.cumsum(a[:-1], dtype=np.uint64)]).astype(np.uint32)
    np.testing.assert_array_equal(out, ref)
    if kind == "ones":
        np.testing.assert_array_equal(out, np.arange(n))


def test_count_scan_permutation_vs_numpy(orc):
    rng = np.random.default_rng(11)
    C = 5000
    keys = rng.integers(0, C, 100000).astype(np.uint32)
    cnt = orc.count(keys, C)
    np.testing.assert_array_equal(cnt[:C], np.bincount(keys, minlength=C))
    assert cnt[C] == 0
    cs = orc.exclusive_scan(cnt)
    assert cs[C] == len(keys) and (np.diff(cs.astype(np.int64)) >= 0).all()
    np.testing.assert_array_equal(orc.stable_scatter(keys, cs), np.argsort(keys, kind="stable"))


# ---------------------------------------------------------- neighbours ----
def test_neighbor_counts_reference_cases(orc):
    """tests/test_spatial.cpp:102-145."""
    p = orc.params()
    g = orc.grid(p)
    r2 = np.float32(0.1) * np.float32(0.1)
    iso = np.array([[0, 0, 0, 1], [1.5, 1.5, 1.5, 1], [-1.5, -1.5, -1.5, 1]], np.float32)
    assert orc.build_spatial_hash(iso, g, r2)["nbr"].tolist() == [0, 0, 0]
    pair = np.array([[0, 0, 0, 1], [0.05, 0, 0, 1]], np.float32)
    assert orc.build_spatial_hash(pair, g, r2)["nbr"].tolist() == [1, 1]
    s = np.float32(0.1) * np.float32(0.99)
    lat = np.array([[x * s, y * s, z * s, 1] for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)], np.float32)
    nb = orc.build_spatial_hash(lat, g, r2)["nbr"]
    assert nb[13] == 6 and nb[0] < nb[13] and nb[26] < nb[13]


def fmaf32(a, b, c):
    """IEEE fused multiply-add in float32 (one rounding of a*b + c), from
    float64: a*b is exact in float64, the sum is formed exactly as s + e
    (TwoSum), and s is rounded to float32; s + e rounds differently from s
    only when s is itself a float32 midpoint and e != 0, which picks the
    neighbour on e's side."""
    a, b, c = (np.asarray(t, np.float32).astype(np.float64) for t in (a, b, c))
    p = a * b
    s = p + c
    bv = s - p
    e = (p - (s - bv)) + (c - bv)
    f = s.astype(np.float32)
    f64 = f.astype(np.float64)
    other = np.nextafter(f, np.where(s > f64, np.float32(np.inf), np.float32(-np.inf))).astype(np.float32)
    mid = (f64 + other.astype(np.float64)) * 0.5
    tie = (s == mid) & (e != 0) & (s != f64)
    up = np.maximum(f, other)
    down = np.minimum(f, other)
    return np.where(tie, np.where(e > 0, up, down), f).astype(np.float32)


def nbr_key32(d0, d1, d2, h2):
    """The neighbour test's key ((d0 d0 - h^2) + d1 d1) + d2 d2, fused."""
    return fmaf32(d2, d2, fmaf32(d1, d1, fmaf32(d0, d0, -np.float32(h2))))


def _brute_counts(pts, h2):
    d = pts[:, None, :3].astype(np.float32) - pts[None, :, :3].astype(np.float32)
    key = nbr_key32(d[..., 0], d[..., 1], d[..., 2], h2)
    np.fill_diagonal(key, np.inf)
    return (key < 0).sum(1)


def test_fmaf32_emulation():
    """fmaf32 against exact rational arithmetic on random and tie cases."""
    from fractions import Fraction
    rng = np.random.default_rng(12)
    a = rng.standard_normal(400).astype(np.float32)
    b = rng.standard_normal(400).astype(np.float32)
    c = (rng.standard_normal(400) * 1e-3).astype(np.float32)
    # constructed ties: (1 + 2^-12)^2 = 1 + 2^-11 + 2^-24 is a float32
    # midpoint; adding +-2^-80 (below float64's resolution there) must break
    # the tie towards its sign, adding 0 leaves it to round-half-even
    t = np.float32(1.0 + 2.0 ** -12)
    a = np.concatenate([a, [t, t, t, -t]]).astype(np.float32)
    b = np.concatenate([b, [t, t, t, t]]).astype(np.float32)
    c = np.concatenate([c, [2.0 ** -80, -(2.0 ** -80), 0.0, 2.0 ** -80]]).astype(np.float32)
    got = fmaf32(a, b, c)
    for x, y, z, g in zip(a, b, c, got):
        exact = Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z))
        lo = np.float32(float(exact))
        cands = [lo, np.nextafter(lo, np.float32(np.inf)), np.nextafter(lo, np.float32(-np.inf))]
        best = min(cands, key=lambda t: (abs(Fraction(float(t)) - exact),
                                          int(np.float32(t).view(np.uint32)) & 1))
        assert g == best, (x, y, z, g, best)


def test_neighbor_key_matches_exact_arithmetic():
    """The fused key's sign is the exact sign of |r|^2 - h^2 (extended
    precision) for every pair except those within a few float32 ulps of the
    sphere (random pairs, and pairs placed on the sphere)."""
    rng = np.random.default_rng(13)
    h = np.float32(0.02)
    h2 = h * h
    d = (rng.uniform(-1.2, 1.2, (200000, 3)) * h).astype(np.float32)
    u = rng.standard_normal((20000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    ring = (u * float(h) * (1.0 + rng.uniform(-1e-6, 1e-6, (20000, 1)))).astype(np.float32)
    d = np.concatenate([d, ring])
    key = nbr_key32(d[:, 0], d[:, 1], d[:, 2], h2)
    dl = d.astype(np.longdouble)
    exact = (dl[:, 0] * dl[:, 0] + dl[:, 1] * dl[:, 1] + dl[:, 2] * dl[:, 2]) - np.longdouble(h2)
    near = np.abs(exact) <= 4 * np.spacing(h2)
    assert (np.sign(key[~near]) == np.sign(exact[~near]).astype(np.float32)).all()
    assert near.sum() > 0 and (~near).sum() > 200000


def test_neighbor_sets_brute_force(orc):
    """Iteration 0 (binning position == current position, cell = h): the
    stencil finds exactly the all-pairs neighbours."""
    rng = np.random.default_rng(3)
    p = orc.params(bounds_min=[0, 0, 0], bounds_max=[1, 1, 1], cell_size=0.1, h=0.1)
    g = orc.grid(p)
    pts = np.concatenate([rng.uniform(0.0, 1.0, (1500, 3)) ** 2, np.ones((1500, 1))], 1).astype(np.float32)
    h2 = np.float32(0.1) * np.float32(0.1)
    bh = orc.build_spatial_hash(pts, g, h2)
    np.testing.assert_array_equal(bh["nbr"], _brute_counts(pts, h2))
    _, _, nbr = orc.density_lambda(bh["sorted4"].astype(np.float64), bh["cell_start"], bh["sorted_index"], g, p, 64)
    np.testing.assert_array_equal(nbr, _brute_counts(pts, h2))


# ------------------------------------------------------- density/lambda ---
def _lambda_of(orc, pts, p, prec=64):
    g = orc.grid(p)
    h2 = np.float32(p.h) * np.float32(p.h)
    bh = orc.build_spatial_hash(np.asarray(pts, np.float32), g, h2)
    sx = bh["sorted4"].astype(np.float64 if prec == 64 else np.float32)
    lam_s, lam_o, nbr = orc.density_lambda(sx, bh["cell_start"], bh["sorted_index"], g, p, prec)
    return lam_o, bh


def test_lambda_isolated_closed_form(orc):
    """tests/test_pbf.cpp:102-106 with the engine's epsilon (800, not the test's 1500)."""
    p = orc.params()
    lam, _ = _lambda_of(orc, [[2, 2, 2, 1]], p)
    h = float(np.float32(0.1))
    expect = -(poly6_norm(h) * h ** 6 / 8000.0 - 1.0) / 800.0
    assert lam[0] == pytest.approx(expect, rel=1e-12)


def test_lambda_sign_reference_case(orc):
    """tests/test_pbf.cpp:80-100: centre of a 3x3x3 cube at 0.4h is over-dense."""
    p = orc.params()
    s = 0.04
    pts = [[x * s, y * s, z * s, 1] for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)] + [[2, 2, 2, 1]]
    lam, _ = _lambda_of(orc, pts, p)
    assert lam[13] < 0 and lam[27] > 0


def test_lambda_pair_closed_form(orc):
    """Two particles at distance d: rho = W(0) + W(d^2), sum grad = grad,
    sum |grad|^2 = |grad|^2, so lambda = -(rho/rho0 - 1) / (2 |grad|^2 / rho0^2 + eps)."""
    p = orc.params(rest_density=300.0, epsilon=0.5)
    d = 0.0375
    lam, _ = _lambda_of(orc, [[0.5, 0.5, 0.5, 1], [0.5 + d, 0.5, 0.5, 1]], p)
    h = float(np.float32(0.1))
    dd = float(np.float32(0.5 + d) - np.float32(0.5))
    rho = W(0, h) + W(dd * dd, h)
    g2 = (spiky_norm(h) * (h - dd) ** 2) ** 2
    expect = -(rho / 300.0 - 1.0) / (2.0 * g2 / 300.0 ** 2 + 0.5)
    assert lam[0] == pytest.approx(expect, rel=1e-10) and lam[1] == pytest.approx(expect, rel=1e-10)


def test_lambda_lattice_interior_closed_form(orc):
    """Interior of a lattice at spacing h/2: 6 + 12 + 8 neighbours inside h
    (the next shell sits at exactly h and is excluded), so
    rho = W(0) (1 + 6 (3/4)^3 + 12 (1/2)^3 + 8 (1/4)^3) = 5.15625 W(0) and
    sum grad = 0."""
    h = 0.125  # spacing 1/16: every coordinate exact in binary
    p = orc.params(bounds_min=[0, 0, 0], bounds_max=[1, 1, 1], cell_size=h, h=h, rest_density=1000.0, epsilon=1.0)
    s = h / 2
    pts = [[0.25 + x * s, 0.25 + y * s, 0.25 + z * s, 1] for z in range(9) for y in range(9) for x in range(9)]
    lam, bh = _lambda_of(orc, pts, p)
    centre = 4 + 4 * 9 + 4 * 81
    rho = 5.15625 * W(0, h)
    gsq = sum(c * (spiky_norm(h) * (h - r * s) ** 2) ** 2 for c, r in ((6, 1.0), (12, math.sqrt(2)), (8, math.sqrt(3))))
    expect = -(rho / 1000.0 - 1.0) / (gsq / 1000.0 ** 2 + 1.0)
    assert lam[centre] == pytest.approx(expect, rel=1e-12)
    assert bh["nbr"][centre] == 26


# --------------------------------------------------------------- delta ----
def _delta_of(orc, pts, p):
    g = orc.grid(p)
    lam, bh = _lambda_of(orc, pts, p)
    sidx = bh["sorted_index"]
    nxt = orc.apply_delta(bh["sorted4"].astype(np.float64), lam[sidx], bh["cell_start"], g, p, 64)
    out = np.empty_like(nxt)
    out[sidx] = nxt
    return out[:, :3] - np.asarray(pts, np.float32)[:, :3].astype(np.float64), lam


def test_delta_pair_with_scorr_at_dq(orc):
    """Pair at d = dq h: W(d^2)/W(dq^2) = 1 so s_corr = -k; then
    dx_0 = (1/rho0) (2 lambda - k) grad W(r_01), dx_1 = -dx_0."""
    h = float(np.float32(0.1))
    p = orc.params(rest_density=1e7, epsilon=50.0, scorr_k=0.1, scorr_dq=0.25)
    x0 = np.float32(0.5)
    x1 = np.float32(0.5) + np.float32(0.025)
    d = float(x1 - x0)
    pts = [[x0, 0.5, 0.5, 1], [x1, 0.5, 0.5, 1]]
    dx, lam = _delta_of(orc, pts, p)
    dq = float(np.float32(0.25)) * h
    scorr = -0.1 * (W(d * d, h) / W(dq * dq, h)) ** 4
    grad0 = -spiky_norm(h) * (h - d) ** 2 * (-1.0)  # r_01 = x0 - x1 = -d along x
    expect = (1.0 / 1e7) * (2 * lam[0] + scorr) * grad0
    assert abs(expect) < 0.3 * h and abs(2 * lam[0]) > 0.1 * abs(scorr)  # no cap; both terms matter
    assert dx[0, 0] == pytest.approx(expect, rel=1e-6)   # positions round to FP32 only via inputs
    assert dx[1, 0] == pytest.approx(-expect, rel=1e-6)
    assert abs(scorr + 0.1) < 1e-3


def test_delta_step_cap(orc):
    """|dx| > 0.3 h is scaled to 0.3 h along the same direction."""
    p = orc.params(rest_density=300.0, epsilon=1e-6, scorr_k=0.0)
    pts = [[0.5, 0.5, 0.5, 1], [0.5, 0.5 + 0.09, 0.5, 1]]
    dx, _ = _delta_of(orc, pts, p)
    h = float(np.float32(0.1))
    for k in range(2):
        assert np.linalg.norm(dx[k]) == pytest.approx(0.3 * h, rel=1e-6)
        assert abs(dx[k, 0]) < 1e-12 and abs(dx[k, 2]) < 1e-12
    assert dx[0, 1] < 0 < dx[1, 1]


def test_delta_antisymmetry_and_bounds_clamp(orc):
    """Iteration 0 without clamps: the pair term is antisymmetric, so sum_i dx_i = 0."""
    rng = np.random.default_rng(5)
    p = orc.params(bounds_min=[0, 0, 0], bounds_max=[1, 1, 1], cell_size=0.1, h=0.1, rest_density=1e6,
                   epsilon=1000.0, scorr_k=1e-3)
    pts = np.concatenate([rng.uniform(0.3, 0.7, (800, 3)), np.ones((800, 1))], 1).astype(np.float32)
    dx, _ = _delta_of(orc, pts, p)
    assert np.linalg.norm(dx, axis=1).max() < 0.3 * 0.1 * 0.999  # no step clamp engaged
    assert np.abs(dx.sum(0)).max() <= 1e-12 * np.abs(dx).sum()
    # a particle pushed through a wall ends exactly on it
    p2 = orc.params(bounds_min=[0.5, 0, 0], bounds_max=[1, 1, 1], cell_size=0.1, h=0.1, rest_density=300.0,
                    epsilon=1e-6, scorr_k=0.0)
    dx2, _ = _delta_of(orc, [[0.5005, 0.5, 0.5, 1], [0.5015, 0.5, 0.5, 1]], p2)
    assert float(np.float32(0.5005)) + float(dx2[0, 0]) == pytest.approx(0.5, abs=1e-15)
    assert dx2[1, 0] > 0.0


# ------------------------------------------------------ predict/finalize --
def test_free_fall_reference_case(orc):
    """tests/test_pbf.cpp:47-65: one step of an isolated particle."""
    p = orc.params()
    x = np.array([[0, 1, 0, 1]], np.float32)
    v = np.zeros((1, 4), np.float32)
    xs = orc.predict(x, v, p, 0.01)
    assert xs[0, 1] == np.float32(1) + (np.float32(0) + np.float32(-9.81) * np.float32(0.01)) * np.float32(0.01)
    xo, vo = orc.step(x, v, p, 0.01)
    assert xo[0, 1] == pytest.approx(0.99902, rel=1e-3)
    assert xo[0, 1] == pytest.approx(float(xs[0, 1]), abs=1e-12)
    assert vo[0, 1] == pytest.approx((float(xs[0, 1]) - 1.0) / float(np.float32(0.01)) * float(np.float32(0.996)), rel=1e-7)


def _fin(orc, p, x_new, prev, dt=0.01):
    x, v = orc.finalize(np.array([x_new], np.float64), np.zeros((1, 4)), np.array([prev], np.float32), p, dt)
    return x[0], v[0]


def test_finalize_walls(orc):
    """Constraints.metal:315-353 hand cases."""
    p = orc.params(damping=1.0)
    dt = 0.01
    inv = 1.0 / float(np.float32(dt))
    # floor: into-wall component removed, tangential x0.55
    x, v = _fin(orc, p, [0.0, -3.0, 0.0, 1], [-0.001, -2.999, 0.0005, 1])
    assert v[1] == 0.0
    assert v[0] == pytest.approx(0.001 * inv * float(np.float32(0.55)), rel=1e-5)
    assert v[2] == pytest.approx(-0.0005 * inv * float(np.float32(0.55)), rel=1e-4)
    # side wall x max, moving out: clamped position, v.x zeroed, v x0.97
    x, v = _fin(orc, p, [3.2, 0.0, 0.0, 1], [2.999, 0.001, 0.0, 1])
    assert x[0] == 3.0 and v[0] == 0.0
    assert v[1] == pytest.approx(-0.001 * inv * float(np.float32(0.97)), rel=1e-4)
    # interior: only damping
    p2 = orc.params()
    x, v = _fin(orc, p2, [0.1, 0.2, 0.3, 1], [0.1, 0.201, 0.3, 1])
    assert v[1] == pytest.approx(-0.001 * inv * float(np.float32(0.996)), rel=1e-4)
    # speed cap h/dt
    x, v = _fin(orc, p2, [0.5, 0.0, 0.0, 1], [0.0, 0.0, 0.0, 1])
    assert np.linalg.norm(v[:3]) == pytest.approx(float(np.float32(0.1)) * inv, rel=1e-6)


def test_step_zero_dt_is_noop_and_zero_iterations_is_free_flight(orc):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-1, 1, (100, 3)), np.ones((100, 1))], 1).astype(np.float32)
    v = np.concatenate([rng.uniform(-1, 1, (100, 3)), np.zeros((100, 1))], 1).astype(np.float32)
    p = orc.params()
    xo, vo = orc.step(x, v, p, 0.0)
    np.testing.assert_array_equal(xo, x)
    np.testing.assert_array_equal(vo, v)
    p0 = orc.params(iterations=0)
    xo, vo = orc.step(x, v, p0, 0.01)
    xs = orc.predict(x, v, p0, 0.01)
    np.testing.assert_array_equal(xo[:, :3], xs[:, :3].astype(np.float64))


def test_fp32_and_fp64_instantiations_agree(orc):
    import mtlphys_amd as m
    p = m.scene_params(3, 4096)
    x, v = m.generate_scene(3, 4096, 1, p)
    op = orc.params(p)
    x64, v64 = orc.step(x, v, op, 1 / 120, 64)
    x32, v32 = orc.step(x, v, op, 1 / 120, 32)
    assert np.abs(x64[:, :3] - x32[:, :3]).max() < 1e-3 * op.h


# ---------------------------------------------------------- golden file ---
def test_reference_known_answers_fixture(orc):
    """tests/golden/reference_known_answers.json holds the reference tests'
    expected values; the oracle reproduces every one."""
    with open(os.path.join(GOLDEN, "reference_known_answers.json")) as f:
        ka = json.load(f)
    g = orc.grid(orc.params())
    pts = np.array(ka["hash"]["points"], np.float32)
    assert orc.hash_keys(pts, g).tolist() == ka["hash"]["keys"]
    r2 = np.float32(0.1) * np.float32(0.1)
    for case in ka["neighbors"]:
        nb = orc.build_spatial_hash(np.array(case["points"], np.float32), g, r2)["nbr"]
        for i, c in case["expect"].items():
            assert nb[int(i)] == c
    ff = ka["free_fall"]
    xo, _ = orc.step(np.array([ff["x"]], np.float32), np.zeros((1, 4), np.float32), orc.params(), ff["dt"])
    assert xo[0, 1] == pytest.approx(ff["y"], rel=ff["rel"])
