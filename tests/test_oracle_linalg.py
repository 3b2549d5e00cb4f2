"""The oracle's linearisation, batched blocks, dense LLT and Schur step, pinned
by brute force: dense J built from complex-step derivatives of an independent
residual, numpy/LAPACK solves, and the full damped normal equations."""
import numpy as np
import pytest

from conftest import rel
from test_oracle_model import complex_step_jac, np_residual


def tiny(gba, C=3, P=20, O=50, seed=5):
    s = gba.synth_spec(1)
    s.num_cameras, s.num_points, s.num_observations, s.kmax = C, P, O, C
    return gba.generate(spec=s, seed=seed)


def dense_system(prob):
    C, P, O = prob.num_cameras, prob.num_points, prob.num_observations
    J = np.zeros((2 * O, 9 * C + 3 * P))
    r = np.zeros(2 * O)
    for e in range(O):
        c, p = prob.obs_camera[e], prob.obs_point[e]
        Jc, Jp = complex_step_jac(prob.cameras[c], prob.points[p], prob.obs_pixel[e])
        J[2 * e:2 * e + 2, 9 * c:9 * c + 9] = Jc
        J[2 * e:2 * e + 2, 9 * C + 3 * p:9 * C + 3 * p + 3] = Jp
        r[2 * e:2 * e + 2] = np_residual(prob.cameras[c], prob.points[p], prob.obs_pixel[e]).real
    return J, r


def test_linearize_matches_dense_jacobian(gba, orc):
    prob = tiny(gba)
    C, P = prob.num_cameras, prob.num_points
    lin = orc.linearize(prob, nthreads=3)
    J, r = dense_system(prob)
    H = J.T @ J
    b = -J.T @ r
    for c in range(C):
        blk = H[9 * c:9 * c + 9, 9 * c:9 * c + 9]
        assert rel(lin["U"][c], np.tril(blk)) < 1e-10
        assert np.all(np.triu(lin["U"][c], 1) == 0)
    for p in range(P):
        o = 9 * C + 3 * p
        assert rel(lin["V"][p], np.tril(H[o:o + 3, o:o + 3])) < 1e-10
    assert rel(lin["bc"].reshape(-1), b[:9 * C]) < 1e-10
    assert rel(lin["bp"].reshape(-1), b[9 * C:]) < 1e-10
    for e in range(prob.num_observations):
        c, p = prob.obs_camera[e], prob.obs_point[e]
        Jc, Jp = J[2 * e:2 * e + 2, 9 * c:9 * c + 9], J[2 * e:2 * e + 2, 9 * C + 3 * p:9 * C + 3 * p + 3]
        assert rel(lin["W"][e], Jc.T @ Jp) < 1e-10
    assert abs(lin["sse"] - r @ r) <= 1e-12 * (r @ r)
    assert lin["anomalies"] == 0


def test_linearize_thread_count_invariant(gba, orc):
    prob = gba.generate(2, 3)
    a = orc.linearize(prob, nthreads=1)
    b = orc.linearize(prob, nthreads=7)
    for k in ("U", "V", "W", "bc", "bp"):
        assert np.array_equal(a[k], b[k])
    assert a["sse"] == b["sse"]


def test_anomalous_edge_is_skipped(gba, orc):
    prob = tiny(gba)
    prob.obs_pixel[4, 0] = np.nan
    lin = orc.linearize(prob, nthreads=2)
    assert lin["anomalies"] == 1
    assert np.all(lin["W"][4] == 0)
    ref = orc.linearize(prob, nthreads=1)
    assert np.isfinite(lin["sse"]) and lin["sse"] == ref["sse"]
    s, an = orc.cost(prob)
    assert an == 1 and np.isfinite(s)


def spd(rng, n, shift):
    M = rng.uniform(-1, 1, (n, n))
    return M.T @ M + shift * np.eye(n)


def test_batch_chol9_and_inverse3_reference_protocol(orc):
    # include tests/test_schur.cu:26-47,83-93: A = M^T M + 9I (3I), M ~ U(-1,1), 64 blocks, 1e-9
    rng = np.random.default_rng(42)
    A = np.stack([spd(rng, 9, 9.0) for _ in range(64)])
    b = rng.uniform(-1, 1, (64, 9))
    x, ok = orc.batch_chol9_factor_solve(A, b)
    assert ok.all()
    assert np.max(np.abs(x - np.linalg.solve(A, b[..., None])[..., 0])) <= 1e-9
    A3 = np.stack([spd(rng, 3, 3.0) for _ in range(64)])
    Ai, ok = orc.batch_inverse3_spd(A3)
    assert ok.all()
    assert np.max(np.abs(Ai - np.linalg.inv(A3))) <= 1e-9


def test_batch_blocks_reject_non_spd(orc):
    A = np.tile(np.eye(9), (3, 1, 1))
    A[1, 4, 4] = -1.0
    A[2, 7, 7] = np.nan
    x, ok = orc.batch_chol9_factor_solve(A, np.ones((3, 9)))
    assert list(ok) == [1, 0, 0]
    A3 = np.tile(np.eye(3), (2, 1, 1))
    A3[1] = [[1, 2, 0], [2, 1, 0], [0, 0, 1]]
    _, ok = orc.batch_inverse3_spd(A3)
    assert list(ok) == [1, 0]


def test_damp_invert_v(orc):
    rng = np.random.default_rng(1)
    V = np.stack([np.tril(spd(rng, 3, 0.5)) for _ in range(10)])
    V[3] = 0.0  # point with no information: diag'(0) = 1e-12
    lam = 1e-3
    Vi, bad = orc.damp_invert_v(V, lam)
    assert bad == 0
    for p in range(10):
        full = np.tril(V[p]) + np.tril(V[p], -1).T
        d = np.diag(full).copy()
        d[d <= 0] = 1e-12
        assert rel(Vi[p], np.linalg.inv(full + lam * np.diag(d))) < 1e-10


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 441])
def test_dense_cholesky_vs_lapack(orc, n):
    rng = np.random.default_rng(n)
    A = spd(rng, n, n * 0.1 + 1)
    b = rng.normal(size=n)
    x, ok, L = orc.dense_cholesky_solve(A, b)
    assert ok
    assert rel(L, np.linalg.cholesky(A)) < 1e-12
    assert rel(x, np.linalg.solve(A, b)) < 1e-10


def test_dense_cholesky_detects_indefinite(orc):
    A = np.eye(70)
    A[66, 66] = -2.0
    _, ok, _ = orc.dense_cholesky_solve(A, np.ones(70))
    assert not ok


def test_schur_step_equals_full_damped_normal_equations(gba, orc):
    prob = tiny(gba, C=4, P=30, O=90, seed=9)
    C, P = prob.num_cameras, prob.num_points
    lin = orc.linearize(prob)
    lam = 1e-3
    J, r = dense_system(prob)
    H = J.T @ J
    b = -J.T @ r
    d = np.diag(H).copy()
    d[d <= 0] = 1e-12
    delta = np.linalg.solve(H + lam * np.diag(d), b)
    out = orc.schur_step(prob, lin, lam, nthreads=2, want_S=True)
    assert out["status"] == "ok"
    assert rel(out["dc"], delta[:9 * C]) < 1e-8
    assert rel(out["dp"].reshape(-1), delta[9 * C:]) < 1e-8
    assert abs(out["denom"] - delta @ (lam * d * delta + b)) <= 1e-8 * abs(out["denom"])
    # S is the Schur complement of the damped system (lower triangle)
    A = H + lam * np.diag(d)
    nc = 9 * C
    S = A[:nc, :nc] - A[:nc, nc:] @ np.linalg.solve(A[nc:, nc:], A[nc:, :nc])
    assert rel(np.tril(out["S"]), np.tril(S)) < 1e-8


@pytest.mark.parametrize("huber", [0.0, 3.0])
def test_scaled_schur_sums_stay_in_fixed_point_range(gba, orc, huber):
    # The mixed form sums the off-diagonal Schur blocks in 64-bit fixed point
    # of the diagonally scaled products D_a W_e V'^-1 W_f^T D_b, relying on
    # every partial sum lying in [-1, 1]: by Cauchy-Schwarz the sum of the
    # magnitudes is at most sqrt(U_a,ii U_b,jj) D_a,i D_b,j <= 1.  Checked here
    # on the oracle's linearisation (with gross outliers and Huber weights,
    # and at a tiny damping), summing magnitudes, the worst case.
    s = gba.synth_spec(1)
    s.num_cameras, s.num_points, s.num_observations = 12, 500, 2500
    s.kmin, s.kmax = 2, 12
    p = gba.generate(spec=s, seed=9)
    rng = np.random.default_rng(9)
    k = rng.choice(p.num_observations, 60, replace=False)
    p.obs_pixel[k] += rng.choice([-40.0, 40.0], size=(60, 2))
    o = orc.linearize(p, huber_delta=huber)
    U, V, W = o["U"], o["V"], o["W"]
    for lam in (1e-4, 1e-12):
        ud = np.array([np.diag(u) for u in U])
        D = 1.0 / np.sqrt(ud + lam * np.maximum(ud, 1e-12))
        Vf = V + np.transpose(V, (0, 2, 1)) * (1 - np.eye(3))
        vd = np.einsum("pii->pi", Vf)
        Vinv = np.linalg.inv(Vf + lam * np.maximum(vd, 1e-12)[:, :, None] * np.eye(3))
        C = p.num_cameras
        mag = np.zeros((C, C, 9, 9))
        order = np.argsort(p.obs_point, kind="stable")
        bounds = np.flatnonzero(np.diff(p.obs_point[order])) + 1
        for es in np.split(order, bounds):
            q = p.obs_point[es[0]]
            for e in es:
                for f in es:
                    a, b = p.obs_camera[e], p.obs_camera[f]
                    if b >= a:
                        continue
                    mag[a, b] += np.abs(W[e] @ Vinv[q] @ W[f].T) * D[a][:, None] * D[b][None, :]
        assert mag.max() <= 1.0 + 1e-9, mag.max()
        assert mag.max() > 0.05  # the check is not vacuous
