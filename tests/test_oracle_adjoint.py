"""CPU: the oracle's FP64 step adjoint (orc_step_vjp_d), density adjoint and
rollout gradient, pinned by central finite differences of the oracle's FP64
step (orc_step_d).

The reference has no adjoint, so finite differences are the pin.  Two
regimes:
  * the reference's own constants (rho0 8000, eps 800, s_corr k 1e-4) on a
    lattice at spacing h/2 (rho = 8078 ~ rho0): a smooth map, where central
    FD with step 1e-4 along unit directions agrees with the adjoint to ~1e-8;
  * the benchmark constants (rho0 6378, eps 600, s_corr k 0.1): a strongly
    compressed lattice whose steps saturate the 0.3 h cap; the map is
    chaotic (|dL/dv0| grows ~100x per step), so FD at 1e-4 leaves the
    linear range after a step or two.  There the test is that FD converges
    to the adjoint as the step shrinks.
Gravity is float32 in the parameters, so its FD divides by the float32
difference actually applied."""
import numpy as np
import pytest

DT = 1.0 / 120.0


def _calm(orc, mtlp, n=512, iterations=3):
    prm = mtlp.scene_params(3, n)
    x, v = mtlp.generate_scene(3, n, 7, prm)
    p = orc.params(prm, iterations=iterations, rest_density=8000.0, epsilon=800.0, scorr_k=1e-4,
                   gravity=[0.0, -9.81, 0.0])
    return p, x.astype(np.float64), v.astype(np.float64)


def _unit(rng, n):
    d = np.zeros((n, 4))
    d[:, :3] = rng.standard_normal((n, 3))
    return d / np.linalg.norm(d)


def _fd(f, x, d, eps):
    return (f(x + eps * d) - f(x - eps * d)) / (2.0 * eps)


def test_step_vjp_one_step_all_inputs(orc, mtlp):
    """One step, a random linear functional of (x_out, v_out)."""
    p, x, v = _calm(orc, mtlp, 512)
    rng = np.random.default_rng(1)
    wx, wv = _unit(rng, len(x)), _unit(rng, len(x))

    def L(xx, vv, pp=p):
        xo, vo, _ = orc.step_d(xx, vv, pp, DT)
        return float((xo * wx).sum() + (vo * wv).sum())

    _, _, tape = orc.step_d(x, v, p, DT)
    xb, vb, gb = orc.step_vjp_d(tape, p, DT, wx, wv)
    assert np.all(xb[:, 3] == 0) and np.all(vb[:, 3] == 0)
    for k in range(4):
        d = _unit(rng, len(x))
        fx = _fd(lambda y: L(y, v), x, d, 1e-4)
        fv = _fd(lambda y: L(x, y), v, d, 1e-4)
        assert abs(fx - (xb * d).sum()) <= 1e-6 * np.linalg.norm(xb), (k, fx, (xb * d).sum())
        assert abs(fv - (vb * d).sum()) <= 1e-6 * np.linalg.norm(vb), (k, fv, (vb * d).sum())
    gy = np.float32(p.gravity[1])
    gp, gm = np.float32(gy + np.float32(2.0 ** -12)), np.float32(gy - np.float32(2.0 ** -12))
    fg = (L(x, v, orc.params(p, gravity=[0.0, float(gp), 0.0])) -
          L(x, v, orc.params(p, gravity=[0.0, float(gm), 0.0]))) / (float(gp) - float(gm))
    assert abs(fg - gb[1]) <= 1e-5 * abs(gb[1]), (fg, gb[1])


def test_rollout_gradient_fd_32_directions(orc, mtlp):
    """10 steps x 3 iterations; 32 seeded unit directions in v0, step 1e-4."""
    p, x, v = _calm(orc, mtlp, 512)
    target = x[:, :3].mean(axis=0) + np.array([0.2, 0.0, 0.0])
    steps = 10
    L0, xb, vb, gb = orc.rollout_grad(x, v, p, DT, steps, target)
    rng = np.random.default_rng(32)
    errs = []
    for _ in range(32):
        d = _unit(rng, len(x))
        Lp = orc.rollout_loss(x, v + 1e-4 * d, p, DT, steps, target)[0]
        Lm = orc.rollout_loss(x, v - 1e-4 * d, p, DT, steps, target)[0]
        fd = (Lp - Lm) / 2e-4
        errs.append(abs(fd - (vb * d).sum()) / np.linalg.norm(vb))
    assert max(errs) < 1e-5, max(errs)
    # x0 directions too
    for _ in range(4):
        d = _unit(rng, len(x))
        fd = (orc.rollout_loss(x + 1e-4 * d, v, p, DT, steps, target)[0] -
              orc.rollout_loss(x - 1e-4 * d, v, p, DT, steps, target)[0]) / 2e-4
        assert abs(fd - (xb * d).sum()) < 1e-5 * np.linalg.norm(xb)
    # gravity: dL/dg_y through every predict
    gy = np.float32(p.gravity[1])
    gp, gm = np.float32(gy + np.float32(2.0 ** -12)), np.float32(gy - np.float32(2.0 ** -12))
    fg = (orc.rollout_loss(x, v, orc.params(p, gravity=[0.0, float(gp), 0.0]), DT, steps, target)[0] -
          orc.rollout_loss(x, v, orc.params(p, gravity=[0.0, float(gm), 0.0]), DT, steps, target)[0]) / (
        float(gp) - float(gm))
    assert abs(fg - gb[1]) < 1e-4 * abs(gb[1]), (fg, gb[1])
    # the running-sum recursion and the direct sum of pbar^0 agree here
    _, _, _, gb_direct = orc.rollout_grad(x, v, p, DT, steps, target, exact_sums=False)
    assert np.abs(gb - gb_direct).max() < 1e-9 * np.abs(gb).max(), (gb, gb_direct)


def test_gravity_recursion_with_walls(orc, mtlp):
    """The gravity recursion's deviation terms (friction, wall gates, speed
    cap, clamps) against the direct sum on a scene that engages them."""
    n = 2000
    prm = mtlp.scene_params(4, n)
    x, v = mtlp.generate_scene(4, n, 5, prm)
    p = orc.params(prm, iterations=3, rest_density=8000.0 * 125.0, epsilon=800.0, scorr_k=1e-4,
                   gravity=[0.0, -9.81, 0.0])
    x, v = x.astype(np.float64), v.astype(np.float64)
    for _ in range(20):
        x, v, _ = orc.step_d(x, v, p, DT)
    target = x[:, :3].mean(axis=0) + np.array([0.2, 0.0, 0.0])
    _, _, _, ge = orc.rollout_grad(x, v, p, DT, 3, target)
    _, _, _, gd = orc.rollout_grad(x, v, p, DT, 3, target, exact_sums=False)
    assert np.abs(ge - gd).max() < 1e-8 * np.abs(gd).max(), (ge, gd)


def test_walls_and_floor_branches(orc, mtlp):
    """A dam column resting on the floor against the walls (friction, wall
    gates, box clamps active), reference constants scaled to h = 0.02.
    Directions whose +-eps runs change a branch signature (clamp, cap, wall
    flag, gate, bin, stencil cell: a kink crossing) are counted and skipped;
    particles resting exactly on a wall with zero normal velocity sit on the
    clamp's kink, where central FD returns half the one-sided derivative."""
    n = 2000
    prm = mtlp.scene_params(4, n)
    x, v = mtlp.generate_scene(4, n, 5, prm)
    p = orc.params(prm, iterations=3, rest_density=8000.0 * 125.0, epsilon=800.0, scorr_k=1e-4,
                   gravity=[0.0, -9.81, 0.0])
    x, v = x.astype(np.float64), v.astype(np.float64)
    for _ in range(20):  # let it reach the floor and walls
        x, v, _ = orc.step_d(x, v, p, DT)
    lo = np.array(p.bounds_min, np.float64)
    on_floor = (x[:, 1] <= lo[1] + np.float64(p.wall_eps)).sum()
    assert on_floor > 50, on_floor
    rng = np.random.default_rng(4)
    wx, wv = _unit(rng, n), _unit(rng, n)

    def L(xx, vv):
        xo, vo, _ = orc.step_d(xx, vv, p, DT)
        return float((xo * wx).sum() + (vo * wv).sum())

    _, _, tape = orc.step_d(x, v, p, DT)
    xb, vb, _ = orc.step_vjp_d(tape, p, DT, wx, wv)
    # For each direction, particles whose signature differs between the +-eps
    # runs are taken out of the direction (a few rounds), then FD must match.
    eps = 1e-6
    removed = []
    for _ in range(8):
        d = _unit(rng, n)
        dropped = 0
        for _round in range(4):
            sp = orc.rollout_kinks(x, v + eps * d, p, DT, 1)[0]
            sm = orc.rollout_kinks(x, v - eps * d, p, DT, 1)[0]
            flip = np.any(sp != sm, axis=1)
            if not flip.any():
                break
            dropped += int(flip.sum())
            d[flip] = 0.0
            d /= np.linalg.norm(d)
        else:
            pytest.fail("kink crossings did not go away")
        removed.append(dropped)
        fd = _fd(lambda y: L(x, y), v, d, eps)
        assert abs(fd - (vb * d).sum()) < 1e-5 * np.linalg.norm(vb), (fd, (vb * d).sum(), dropped)
    assert sum(removed) > 0  # this scene does put particles on kinks
    assert max(removed) < n // 50, removed


def test_chaotic_regime_fd_converges(orc, mtlp):
    """Benchmark constants (s_corr k = 0.1): FD error falls with the step."""
    n = 512
    prm = mtlp.scene_params(3, n)
    x, v = mtlp.generate_scene(3, n, 7, prm)
    p = orc.params(prm, iterations=3)
    x, v = x.astype(np.float64), v.astype(np.float64)
    target = x[:, :3].mean(axis=0) + np.array([0.2, 0.0, 0.0])
    _, _, vb, _ = orc.rollout_grad(x, v, p, DT, 2, target)
    d = _unit(np.random.default_rng(0), n)
    ad = (vb * d).sum()
    errs = []
    for eps in (1e-4, 1e-5, 1e-6):
        fd = (orc.rollout_loss(x, v + eps * d, p, DT, 2, target)[0] -
              orc.rollout_loss(x, v - eps * d, p, DT, 2, target)[0]) / (2 * eps)
        errs.append(abs(fd - ad) / abs(ad))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-4, errs


def test_density_vjp_fd(orc, mtlp):
    p, x, _ = _calm(orc, mtlp, 729)
    rng = np.random.default_rng(2)
    rb = rng.standard_normal(len(x))
    xb = orc.density_vjp_d(x, p, rb)
    rho = orc.density_at_d(x, p)
    # isolated-particle and lattice closed forms of the density itself
    w0 = 315.0 / (64.0 * np.pi * np.float64(p.h) ** 3)
    assert rho.min() >= w0
    for _ in range(4):
        d = _unit(rng, len(x))
        fd = _fd(lambda y: float(orc.density_at_d(y, p) @ rb), x, d, 1e-6)
        assert abs(fd - (xb * d).sum()) < 1e-7 * np.linalg.norm(xb)


def test_finalize_speed_cap_and_clamp(orc):
    """Single particles: one beyond the speed cap h/dt, one predicted out of
    the box (the clamp zeroes that axis), one sliding on the floor."""
    p = orc.params(None, iterations=1)
    x = np.array([[0.0, 1.0, 0.0, 1.0], [2.99, 0.0, 0.0, 1.0], [1.0, -2.9995, 0.5, 1.0]])
    v = np.array([[30.0, -5.0, 2.0, 0.0], [4.0, 0.5, 0.0, 0.0], [1.0, -0.05, 0.3, 0.0]])
    rng = np.random.default_rng(3)
    wx, wv = _unit(rng, 3), _unit(rng, 3)
    xo, vo, tape = orc.step_d(x, v, p, DT)
    cap = p.vmax * p.h * (1.0 / float(np.float32(DT)))  # dt enters as float32
    assert np.linalg.norm(vo[0, :3]) == pytest.approx(cap, rel=1e-12)
    assert xo[1, 0] == pytest.approx(p.bounds_max[0])
    assert vo[2, 1] == 0.0 and abs(vo[2, 0]) < abs(v[2, 0])  # floor friction
    xb, vb, _ = orc.step_vjp_d(tape, p, DT, wx, wv)
    # particle 1 ends clamped on the x wall with its normal velocity gated to 0:
    # x_out and v_out no longer depend on its x inputs
    assert xb[1, 0] == 0.0 and vb[1, 0] == 0.0 and xb[1, 1] != 0.0
    for arr, grad in ((x, xb), (v, vb)):
        for _ in range(4):
            d = _unit(rng, 3)

            def L(y):
                xx, vv = (y, v) if arr is x else (x, y)
                a, b, _ = orc.step_d(xx, vv, p, DT)
                return float((a * wx).sum() + (b * wv).sum())

            fd = _fd(L, arr, d, 1e-7)
            assert abs(fd - (grad * d).sum()) < 1e-6 * max(np.linalg.norm(grad), 1.0)
