"""Huber robust loss (the reference's Ceres harness, benchmarks/bench_ceres.cpp:120-141:
cost = sum rho(|r|^2) / 2, rho(s) = s for s <= delta^2, else 2 delta sqrt(s) - delta^2).

CPU: the oracle's robust cost against an independent numpy evaluation of the
reference harness's formula, and its IRLS linearisation against dense blocks
built from complex-step Jacobians of an independent numpy residual, scaled by
sqrt(rho'(s)); a huge delta reduces to the plain pass bit for bit; the robust
LM lowers the robust cost monotonically.  GPU: linearisation, cost and the whole
LM trajectory against the oracle."""
import numpy as np
import pytest

from test_oracle_model import complex_step_jac, np_residual


def with_outliers(gba, cfg=1, seed=2, frac=0.03, size=40.0):
    p = gba.generate(cfg, seed)
    rng = np.random.default_rng(seed)
    k = rng.choice(p.num_observations, int(frac * p.num_observations), replace=False)
    p.obs_pixel[k] += rng.choice([-size, size], size=(len(k), 2))
    return p


def np_rho(s, d):
    return np.where(s <= d * d, s, 2 * d * np.sqrt(s) - d * d)


def test_oracle_cost_matches_harness_formula(gba, orc):
    p = with_outliers(gba)
    d = 3.0
    r = np.array([np_residual(p.cameras[c], p.points[q], p.obs_pixel[e]).real
                  for e, (c, q) in enumerate(zip(p.obs_camera, p.obs_point))])
    s = np.sum(r * r, axis=1)
    assert np.mean(s > d * d) > 0.02  # the outliers are in the linear part of the loss
    ref = float(np.sum(np_rho(s, d)))
    got, an = orc.cost(p, huber_delta=d)
    assert an == 0
    assert abs(got - ref) <= 1e-12 * ref
    plain, _ = orc.cost(p)
    assert got < plain


def test_oracle_huge_delta_is_plain(gba, orc):
    p = with_outliers(gba)
    a = orc.linearize(p)
    b = orc.linearize(p, huber_delta=1e30)
    for k in ("U", "V", "W", "bc", "bp"):
        assert np.array_equal(a[k], b[k])
    assert a["sse"] == b["sse"]
    assert orc.cost(p)[0] == orc.cost(p, huber_delta=1e30)[0]


def test_oracle_irls_linearisation_brute_force(gba, orc):
    # a small problem so the complex-step Jacobians stay cheap
    s = gba.synth_spec(1)
    s.num_cameras, s.num_points, s.num_observations = 5, 60, 240
    s.kmin, s.kmax = 2, 5
    p = gba.generate(spec=s, seed=3)
    rng = np.random.default_rng(1)
    k = rng.choice(p.num_observations, 20, replace=False)
    p.obs_pixel[k] += 25.0
    d = 2.0
    C, P = p.num_cameras, p.num_points
    U = np.zeros((C, 9, 9)); V = np.zeros((P, 3, 3)); bc = np.zeros((C, 9)); bp = np.zeros((P, 3))
    W = np.zeros((p.num_observations, 9, 3)); rho = 0.0
    for e, (c, q) in enumerate(zip(p.obs_camera, p.obs_point)):
        r = np_residual(p.cameras[c], p.points[q], p.obs_pixel[e]).real
        Jc, Jp = complex_step_jac(p.cameras[c], p.points[q], p.obs_pixel[e])
        sq = float(r @ r)
        w = np.sqrt(d / np.sqrt(sq)) if sq > d * d else 1.0   # sqrt(rho'(s))
        r, Jc, Jp = w * r, w * Jc, w * Jp
        rho += float(np_rho(sq, d))
        U[c] += Jc.T @ Jc; V[q] += Jp.T @ Jp; W[e] = Jc.T @ Jp
        bc[c] -= Jc.T @ r; bp[q] -= Jp.T @ r
    o = orc.linearize(p, huber_delta=d)
    low9, low3 = np.tril(np.ones((9, 9))), np.tril(np.ones((3, 3)))
    for name, ref, got in (("U", U * low9, o["U"]), ("V", V * low3, o["V"]), ("W", W, o["W"]),
                           ("bc", bc, o["bc"]), ("bp", bp, o["bp"])):
        assert np.max(np.abs(got - ref)) <= 1e-7 * max(1.0, np.max(np.abs(ref))), name
    assert abs(o["sse"] - rho) <= 1e-12 * rho


def test_oracle_robust_lm_monotone(gba, orc):
    # reweighting keeps the relative decrease above 1e-6 for many iterations
    # with 3% gross outliers, so the solve is capped
    p = with_outliers(gba, cfg=2)
    o = orc.lm_solve(p, initial_lambda=1e-4, huber_delta=3.0, max_iterations=20)
    assert o["termination"] == "MAX_ITERATIONS" and o["iterations"] == 20
    acc = o["trace"][o["trace"][:, 4] == 0][:, 2]
    assert np.all(np.diff(acc) < 0) and acc[-1] == o["final_cost"] < o["initial_cost"]
    assert abs(2 * o["final_cost"] - orc.cost(p, huber_delta=3.0)[0]) <= 1e-12 * o["final_cost"]


@pytest.mark.gpu
def test_gpu_robust_linearisation_and_cost(gba, orc):
    p = with_outliers(gba)
    d = 3.0
    g = gba.linearize(p, huber_delta=d)
    o = orc.linearize(p, huber_delta=d)
    for k in ("U", "V", "W", "bc", "bp"):
        assert np.max(np.abs(g[k] - o[k])) <= 1e-12 * max(1.0, np.max(np.abs(o[k]))), k
    assert abs(g["sse"] - o["sse"]) <= 1e-12 * o["sse"]
    gc, _ = gba.cost(p, huber_delta=d)
    oc, _ = orc.cost(p, huber_delta=d)
    assert abs(gc - oc) <= 1e-12 * oc


# The robust problems below are still far from converged after a few
# iterations (every accepted step changes the weights) and their later steps
# are ill-conditioned: the GPU's FP64 Schur sums run in a different order
# than the oracle's, and those rounding differences grow to ~1e-6 of the cost
# by iteration 12.  The trajectory is compared over the first five
# iterations, where it agrees to 1e-9.
@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [1, 2])
def test_gpu_robust_lm_matches_oracle(gba, orc, cfg):
    d = 3.0
    p_gpu, p_cpu = with_outliers(gba, cfg), with_outliers(gba, cfg)
    r = gba.solve_bal_gpu(p_gpu, gba.LMOptions(initial_lambda=1e-4, max_iterations=5), huber_delta=d)
    o = orc.lm_solve(p_cpu, initial_lambda=1e-4, huber_delta=d, max_iterations=5)
    assert r.termination == o["termination"]
    assert r.iterations == o["iterations"]
    assert abs(r.initial_cost - o["initial_cost"]) <= 1e-12 * o["initial_cost"]
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]
    ga = r.trace[r.trace[:, 4] == 0][:, 2]
    oa = o["trace"][o["trace"][:, 4] == 0][:, 2]
    assert len(ga) == len(oa) and np.max(np.abs(ga - oa) / oa) < 1e-9


@pytest.mark.gpu
def test_gpu_robust_mixed_reaches_fp64_cost(gba):
    d = 3.0
    p64, p32 = with_outliers(gba, 2), with_outliers(gba, 2)
    opts = gba.LMOptions(initial_lambda=1e-4, max_iterations=5)
    r64 = gba.solve_bal_gpu(p64, opts, huber_delta=d)
    r32 = gba.solve_bal_gpu(p32, opts, gba.PRECISION_MIXED, huber_delta=d)
    assert r32.termination == r64.termination
    assert abs(r32.final_cost - r64.final_cost) <= 1e-6 * r64.final_cost


@pytest.mark.gpu
def test_gpu_robust_mixed_refinement_residual_matches_oracle(gba, orc):
    """The mixed form's FP64 refinement with the Huber loss: its residual uses
    the linearisation's robust weights (the square of the IRLS weight on
    W = Jc^T Jp).  The first trial's step (the change of the cameras after a
    one-iteration solve) against the oracle's robust S and rhs (src/lm.cu's
    system with bench_ceres.cpp's HuberLoss) recomputed on the host: the
    residual the solve reports, to the first digits.  A wrong weight would
    make the refinement's residual that of another operator."""
    d, lam = 3.0, 1e-4
    p = with_outliers(gba, 2)
    s = gba.Solver(p, gba.PRECISION_MIXED, huber_delta=d)
    r = s.solve(gba.LMOptions(initial_lambda=lam, max_iterations=1))
    assert r.trials == 1 and r.iterations == 1
    cams, _ = s.params()
    dc = (cams - p.cameras).reshape(-1)
    lin = orc.linearize(p, huber_delta=d)
    o = orc.schur_step(p, lin, lam, want_S=True)
    S = np.tril(o["S"]) + np.tril(o["S"], -1).T
    dd = np.concatenate([np.diag(u) for u in lin["U"]])
    D = 1.0 / np.sqrt(dd + lam * np.maximum(dd, 1e-12))
    ref = np.linalg.norm(D * (o["rhs"] - S @ dc)) / np.linalg.norm(D * o["rhs"])
    print(f"Huber: reported FP64 residual {r.max_refine_residual:.3e}, recomputed {ref:.3e}")
    assert ref < 1e-8
    assert abs(r.max_refine_residual - ref) <= 0.05 * ref + 1e-12
