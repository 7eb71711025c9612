"""GPU: the scene extras of Engine::step and Engine::applyMousePulse --
computeFoam (fused into finalize), sphereCollision (fused into the delta
pass) with the sphere's host kinematics, and the mouse pulse -- against the
FP64 oracle fed the GPU's own inputs."""
import numpy as np
import pytest

from helpers import position_error_h

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0
X_TOL_H = 1e-5


def _engine(mtlp, p, x, v):
    e = mtlp.Engine(0, p)
    e.reset(len(x), mtlp.capi.SCENE_EMPTY)
    e.upload_state(x, v)
    return e


def test_foam_matches_oracle(mtlp, orc):
    """Foam from the step's own velocities and neighbour counts; the dam
    break's first steps give speeds across the smoothstep's 1.2..4.5 band."""
    p = mtlp.scene_params(4, 20000)
    x, v = mtlp.generate_scene(4, 20000, 2, p)
    rng = np.random.default_rng(1)
    v[:, :3] = rng.normal(0.0, 2.5, (len(v), 3)).astype(np.float32)  # spread the speeds
    with _engine(mtlp, p, x, v) as e:
        e.step(DT)
        foam = e.foam_intensity_buffer().read()
        vel = e.velocities_buffer().read()
        nbr = e.neighbor_counts_buffer().read()
        assert e.error_flags() == 0
    ref = orc.foam(vel.astype(np.float64), nbr)
    assert ((ref > 0) & (ref < 1)).sum() > 1000
    assert np.abs(foam - ref).max() < 2e-6


def test_mouse_pulse_matches_oracle(mtlp, orc):
    p = mtlp.scene_params(3, 32768)
    x, v = mtlp.generate_scene(3, 32768, 4, p)
    c = x[:, :3].mean(axis=0)
    origin = c + np.array([-2.0, 0.3, 0.1], np.float32)
    d = (c - origin) / np.linalg.norm(c - origin)
    force, radius = [0.5, 3.0, -1.0], 0.25
    with _engine(mtlp, p, x, v) as e:
        e.apply_mouse_pulse(origin.tolist(), d.tolist(), force, radius)
        got = e.velocities_buffer().read()
    ref = orc.mouse_pulse(x.astype(np.float64), v.astype(np.float64), origin.astype(np.float64),
                          d.astype(np.float64), force, radius)
    hit = np.abs(ref[:, :3] - v[:, :3]).max(axis=1) > 0
    assert hit.sum() > 100 and (~hit).sum() > 100
    assert np.abs(got[:, :3] - ref[:, :3]).max() < 1e-5 * np.abs(force).max()
    np.testing.assert_array_equal(got[~hit].view(np.uint32), v[~hit].view(np.uint32))


def _sphere_run(mtlp, p, x, v, sphere, iterations):
    q = p.copy()
    q.iterations = iterations
    with _engine(mtlp, q, x, v) as e:
        e.set_sphere(True, sphere[0], sphere[1], sphere[2])
        e.step(DT)
        on, c, vel, r = e.sphere()
        out = dict(x=e.positions_buffer().read(), v=e.velocities_buffer().read(),
                   sorted=e.sorted_positions_buffer().read(), lam=e.lambdas_buffer().read(),
                   cs=e.cell_start_buffer().read(), sidx=e.sorted_index_buffer().read(), center=c)
        assert e.error_flags() == 0
    return out


def test_sphere_collision_kernel_level(mtlp, orc):
    """A sphere sunk into a settled cube: every delta pass with the sphere
    push against the FP64 oracle fed the GPU's iterates; no final iterate is
    left inside the sphere."""
    p = mtlp.scene_params(3, 8000)
    x, v = mtlp.generate_scene(3, 8000, 3, p)
    with _engine(mtlp, p, x, v) as e:
        for _ in range(10):
            e.step(DT)
        x, v = e.positions_buffer().read(), e.velocities_buffer().read()
    centre = x[:, :3].mean(axis=0).tolist()
    sphere = (centre, [0.0, 0.0, 0.0], 0.3)
    runs = [_sphere_run(mtlp, p, x, v, sphere, k) for k in range(p.iterations + 1)]
    c = np.array(runs[0]["center"], np.float32)
    op = orc.params(p, sphere=[*c.tolist(), 0.3], sphere_on=1)
    g = orc.grid(op)
    cs, sidx = runs[0]["cs"], runs[0]["sidx"]
    inside0 = (np.linalg.norm(runs[0]["sorted"][:, :3] - c, axis=1) < 0.3).sum()
    assert inside0 > 10
    for k in range(p.iterations):
        inp = runs[k]["sorted"].astype(np.float64)
        nxt = orc.apply_delta(inp, runs[k + 1]["lam"][sidx].astype(np.float64), cs, g, op, 64)
        xe = position_error_h(runs[k + 1]["sorted"], nxt, op.h)
        assert xe["max"] <= X_TOL_H, (k, xe)
        r = np.linalg.norm(runs[k + 1]["sorted"][:, :3].astype(np.float64) - c, axis=1)
        assert r.min() >= 0.3 * (1 - 1e-6)


def test_sphere_scene_kinematics_and_stability(mtlp, orc):
    """Scene 2 (the lake) with its sphere: the engine integrates the sphere
    like the oracle's restatement of Engine.cpp:693-712, and the lake stays
    finite and in bounds while the sphere falls into it."""
    n = 20000
    with mtlp.Engine(0) as e:
        e.reset(n, mtlp.capi.SCENE_SPHERE_DROP)
        on, c0, v0, r = e.sphere()
        assert on and c0 == [0.0, 4.5, 0.0] and r == pytest.approx(0.6)
        p = e.params
        op = orc.params(p, sphere=[0.0, 0.0, 0.0, r])
        st = np.array(c0 + v0, np.float32)
        for _ in range(130):
            e.step(DT)
            st = orc.sphere_advance(st, op, DT)
        on, c, vel, _ = e.sphere()
        np.testing.assert_allclose(c, st[:3], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(vel, st[3:], rtol=1e-4, atol=1e-5)
        x = e.positions_buffer().read()
        assert e.error_flags() == 0
    assert np.isfinite(x).all()
    lo, hi = np.array(p.bounds_min), np.array(p.bounds_max)
    assert (x[:, :3] >= lo - 1e-6).all() and (x[:, :3] <= hi + 1e-6).all()
    assert c[1] < 4.5 - 1.0  # it fell


def test_integrate_matches_oracle_and_reference_cases(mtlp, orc):
    """Engine::integrate: the reference's two integrator tests
    (tests/test_integrator.cpp:47-99) and 10k random particles vs the FP64
    oracle."""
    with mtlp.Engine(0) as e:
        e.reset(1, mtlp.capi.SCENE_EMPTY)
        e.upload_state(np.array([[0, 5, 0, 1]], np.float32), np.zeros((1, 4), np.float32))
        x, v = np.array([[0.0, 5.0, 0.0, 1.0]]), np.zeros((1, 4))
        op = orc.params(e.params)
        for _ in range(100):
            e.integrate(0.01)
            x, v = orc.integrate(x, v, op, 0.01)
        gx, gv = e.positions_buffer().read(), e.velocities_buffer().read()
        assert gx[0, 1] == pytest.approx(x[0, 1], rel=1e-4) and gv[0, 1] == pytest.approx(v[0, 1], rel=1e-4)
        e.upload_state(np.array([[0, -2.95, 0, 1]], np.float32), np.array([[0, -10, 0, 0]], np.float32))
        e.integrate(0.05)
        assert e.positions_buffer().read()[0, 1] == pytest.approx(-3.0) and e.velocities_buffer().read()[0, 1] > 0
    rng = np.random.default_rng(9)
    n = 10000
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.uniform(-3.2, 3.2, (n, 3))
    x[:, 3] = 1
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = rng.normal(0, 20, (n, 3))
    with mtlp.Engine(0) as e:
        e.reset(n, mtlp.capi.SCENE_EMPTY)
        e.upload_state(x, v)
        e.integrate(1.0 / 60.0)
        gx, gv = e.positions_buffer().read(), e.velocities_buffer().read()
        op = orc.params(e.params)
    rx, rv = orc.integrate(x.astype(np.float64), v.astype(np.float64), op, np.float32(1.0 / 60.0))
    assert np.abs(gx[:, :3] - rx[:, :3]).max() < 1e-5
    assert np.abs(gv[:, :3] - rv[:, :3]).max() < 1e-4
