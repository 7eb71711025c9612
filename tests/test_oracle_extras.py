"""CPU: the oracle's scene extras -- sphereCollision inside the delta pass,
the sphere's host kinematics, computeFoam and applyMousePulse -- pinned by
hand cases of the formulas in Constraints.metal:222-299 and
Engine.cpp:693-712."""
import numpy as np
import pytest

DT = 1.0 / 120.0


def test_sphere_pushes_iterate_to_surface(orc):
    """An isolated particle (no neighbours, so no delta) inside the sphere
    lands on the surface along the ray from the centre; one outside stays."""
    p = orc.params(None, iterations=1, sphere=[0.5, 0.5, 0.5, 0.6], sphere_on=1)
    g = orc.grid(p)
    pts = np.array([[0.7, 0.5, 0.5, 1.0], [2.0, 2.0, 2.0, 1.0], [0.5, 0.5, 0.5, 1.0]], np.float32)
    h = orc.build_spatial_hash(pts, g, np.float32(p.h) ** 2)
    sx = h["sorted4"].astype(np.float64)
    out = orc.apply_delta(sx, np.zeros(3), h["cell_start"], g, p, 64)
    res = np.empty_like(out)
    res[h["sorted_index"]] = out
    np.testing.assert_allclose(res[0, :3], [0.5 + 0.6, 0.5, 0.5], rtol=0, atol=1e-7)
    np.testing.assert_array_equal(res[1, :3], pts[1, :3].astype(np.float64))
    # at the centre: r = sqrt(max(0, 1e-8)), the direction 0/r = 0 keeps it at the centre
    np.testing.assert_allclose(res[2, :3], [0.5, 0.5, 0.5], atol=1e-7)
    assert np.linalg.norm(res[0, :3] - 0.5) == pytest.approx(0.6, abs=1e-7)


def test_sphere_kinematics(orc):
    p = orc.params(None, sphere=[0.0, 0.0, 0.0, 0.6])
    st = np.array([0.0, 4.5, 0.0, 0.0, 0.0, 0.0], np.float32)
    s1 = orc.sphere_advance(st, p, DT)
    vy = np.float32(np.float32(-9.81) * np.float32(DT)) * np.float32(0.998)
    assert s1[4] == pytest.approx(vy, rel=1e-6)
    assert s1[1] == pytest.approx(4.5 + vy * DT, rel=1e-6)
    # floor bounce: below bmin + r it is put back and the velocity reverses with 0.4
    low = np.array([1.0, -2.45, 0.0, 0.5, -3.0, 0.0], np.float32)
    s2 = orc.sphere_advance(low, p, DT)
    assert s2[1] == pytest.approx(p.bounds_min[1] + 0.6)
    assert s2[4] > 0 and s2[4] == pytest.approx(0.4 * abs((-3.0 - 9.81 * DT) * 0.998), rel=1e-5)
    assert s2[3] == pytest.approx(0.5 * 0.998, rel=1e-6)


def test_foam_formula(orc):
    v = np.zeros((6, 4))
    v[:, 0] = [0.5, 1.2, 2.85, 4.5, 9.0, 2.85]
    nbr = np.array([0, 0, 0, 0, 96, 48], np.uint32)
    f = orc.foam(v, nbr)
    np.testing.assert_allclose(f, [0.0, 0.0, 0.5, 1.0, 0.0, 0.25], atol=1e-12)


def test_mouse_pulse_formula(orc):
    x = np.array([[1.0, 0.0, 0.0, 1.0], [1.0, 0.1, 0.0, 1.0], [-1.0, 0.0, 0.0, 1.0], [1.0, 0.3, 0.0, 1.0]])
    v = np.zeros((4, 4))
    out = orc.mouse_pulse(x, v, [0, 0, 0], [1, 0, 0], [0, 2.0, 0], 0.2)
    assert out[0, 1] == pytest.approx(2.0)                       # on the ray: weight 1
    assert out[1, 1] == pytest.approx(2.0 * np.exp(-0.01 / (0.04 * 0.4)))
    assert out[2, 1] == 0.0                                      # behind the origin
    assert out[3, 1] == 0.0                                      # outside the radius


def test_integrate_reference_cases(orc):
    """tests/test_integrator.cpp:47-99 restated: free fall for 100 steps of
    0.01 s against the closed form of the recurrence, and a floor bounce."""
    p = orc.params(None)
    x = np.array([[0.0, 5.0, 0.0, 1.0]])
    v = np.zeros((1, 4))
    vy, py = 0.0, 5.0
    for _ in range(100):
        x, v = orc.integrate(x, v, p, 0.01)
        vy = (vy - 9.81 * 0.01) * 0.999
        py += vy * 0.01
        if py < -3.0:
            py, vy = -3.0, -vy * 0.4
    assert x[0, 1] == pytest.approx(py, rel=1e-12) and v[0, 1] == pytest.approx(vy, rel=1e-12)
    x, v = orc.integrate(np.array([[0.0, -2.95, 0.0, 1.0]]), np.array([[0.0, -10.0, 0.0, 0.0]]), p, 0.05)
    assert x[0, 1] == -3.0 and v[0, 1] > 0.0
