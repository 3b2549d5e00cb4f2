"""The oracle's camera model and Jacobian, pinned by the reference's own
reader (golden vectors from oracle/_ref), the reference's finite-difference
cases (tests/test_vs_ceres.cpp:34-72) and complex-step derivatives of an
independent numpy residual."""
import os

import numpy as np
import pytest

from conftest import rel

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# (rodrigues, t, f, k1, k2), point, pixel: reference tests/test_vs_ceres.cpp:34-39
REF_CASES = [
    [0.1, -0.5, 0.3, 0.02, 0.01, -1.0, 500.0, -0.2, 0.01, 1.0, 2.0, 30.0, 100.0, -80.0],
    [1e-8, 2e-8, -1e-8, -0.05, 0.2, 0.4, 350.0, 0.1, -0.001, -2.0, 1.0, 25.0, -250.0, 300.0],
]


def np_residual(cam, X, obs):
    """Independent residual (numpy, complex-safe): Rodrigues via the matrix exponential."""
    a = cam[0:3]
    th2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2]
    th = np.sqrt(th2)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]], dtype=complex)
    R = np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th2 * (K @ K)
    P = R @ X + cam[3:6]
    u, v = -P[0] / P[2], -P[1] / P[2]
    s = u * u + v * v
    d = 1 + cam[7] * s + cam[8] * s * s
    return np.array([cam[6] * d * u - obs[0], cam[6] * d * v - obs[1]])


def complex_step_jac(cam, X, obs):
    h = 1e-30
    Jc = np.zeros((2, 9))
    Jp = np.zeros((2, 3))
    for j in range(9):
        c = cam.astype(complex)
        c[j] += 1j * h
        Jc[:, j] = np_residual(c, X.astype(complex), obs).imag / h
    for j in range(3):
        x = X.astype(complex)
        x[j] += 1j * h
        Jp[:, j] = np_residual(cam.astype(complex), x, obs).imag / h
    return Jc, Jp


def test_residual_matches_reference_reader_golden(orc):
    g = np.load(os.path.join(GOLD, "residuals.npz"))
    for i in range(len(g["res"])):
        r = orc.snavely_residual(g["cams"][i], g["pts"][i], g["obs"][i])
        assert np.array_equal(r, g["res"][i]) or rel(r, g["res"][i]) < 1e-15


def test_edge_residual_agrees_with_reader_model(orc):
    g = np.load(os.path.join(GOLD, "residuals.npz"))
    for i in range(len(g["res"])):
        r = orc.edge_residual(g["cams"][i], g["pts"][i], g["obs"][i])
        # first-order (lm.cu) vs zero-order (reader) small-angle branch differ by ~theta*|X|
        assert np.max(np.abs(r - g["res"][i])) <= 1e-9 * max(1.0, np.max(np.abs(g["res"][i])))


def test_total_squared_error_golden(orc):
    g = np.load(os.path.join(GOLD, "problem_sse.npz"))

    class P:
        pass
    p = P()
    p.cameras, p.points = g["cams"], g["pts"]
    p.obs_camera, p.obs_point, p.obs_pixel = g["obs_camera"], g["obs_point"], g["obs_pixel"]
    assert orc.total_squared_error(p) == float(g["sse"])


@pytest.mark.parametrize("case", REF_CASES)
def test_jacobian_finite_difference_reference_cases(orc, case):
    cam, X, obs = np.array(case[:9]), np.array(case[9:12]), np.array(case[12:14])
    r0, Jc, Jp = orc.edge_residual_jacobian(cam, X, obs)
    eps, worst = 1e-6, 0.0
    for params, n, J in ((cam, 9, Jc), (X, 3, Jp)):
        for j in range(n):
            s = params[j]
            params[j] = s + eps
            rp = orc.edge_residual(cam, X, obs)
            params[j] = s - eps
            rm = orc.edge_residual(cam, X, obs)
            params[j] = s
            num = (rp - rm) / (2 * eps)
            worst = max(worst, np.max(np.abs(num - J[:, j]) / np.maximum(1.0, np.abs(J[:, j]))))
    assert worst < 1e-5


def test_jacobian_complex_step(orc):
    rng = np.random.default_rng(3)
    for _ in range(50):
        cam = np.concatenate([rng.normal(0, 1.0, 3), rng.normal(0, 1, 3) + [0, 0, -30],
                              [rng.uniform(300, 700), rng.normal(0, 0.05), rng.normal(0, 0.005)]])
        X, obs = rng.uniform(-5, 5, 3), rng.normal(0, 100, 2)
        _, Jc, Jp = orc.edge_residual_jacobian(cam, X, obs)
        Jc_cs, Jp_cs = complex_step_jac(cam, X, obs)
        assert rel(Jc, Jc_cs) < 1e-11
        assert rel(Jp, Jp_cs) < 1e-11


def test_jacobian_small_angle_branch(orc):
    # theta^2 <= 1e-24: first-order model, dRX/da = -[X]x, R = I (src/lm.cu:141-148);
    # it agrees with the exact derivative to O(theta)
    cam = np.array([1e-13, -2e-13, 3e-13, 0.1, -0.2, -30.0, 500.0, -0.1, 0.01])
    X, obs = np.array([1.0, -2.0, 3.0]), np.array([10.0, -5.0])
    _, Jc, Jp = orc.edge_residual_jacobian(cam, X, obs)
    Jc_cs, Jp_cs = complex_step_jac(cam, X, obs)
    assert rel(Jc, Jc_cs) < 1e-10
    assert rel(Jp, Jp_cs) < 1e-10
