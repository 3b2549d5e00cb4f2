"""Parity of the HIP path (through the C ABI of libgba_amd.so) with the CPU
oracle on the same seeded inputs.  Tolerances: FP64 pieces agree to 1e-12
relative (summation order only), the FP64 step / final cost to 1e-9; the mixed
form's final cost to 1e-6 of the FP64 path."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def heavy(gba, seed=3):
    # points with more than 64 observations exercise the looping work units
    s = gba.synth_spec(3)
    s.num_cameras, s.num_points, s.num_observations, s.kmax = 120, 600, 6000, 120
    return gba.generate(spec=s, seed=seed)


def ragged(gba):
    # a camera without observations (zero U block: its S diagonal is only the
    # damping floor lambda 1e-12) and 300 points seen once (rank-2 V blocks)
    p = gba.generate(2, 1)
    keep = np.ones(p.num_observations, bool)
    seen = set()
    for e, q in enumerate(p.obs_point):
        if q < 300:
            keep[e] = q not in seen
            seen.add(q)
    cams = np.vstack([p.cameras, p.cameras[:1] * 1.01])
    return gba.BALProblem(cams, p.points.copy(), p.obs_camera[keep].copy(), p.obs_point[keep].copy(),
                          p.obs_pixel[keep].copy())


def arel(a, b):
    """max |a-b| relative to max |b| (per array)."""
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1e-300, np.max(np.abs(b))))


@pytest.mark.parametrize("which", ["cfg2", "heavy"])
def test_linearize_fp64_matches_oracle(gba, orc, which):
    prob = gba.generate(2, 1) if which == "cfg2" else heavy(gba)
    g = gba.linearize(prob)
    o = orc.linearize(prob)
    for k in ("U", "V", "W", "bc", "bp"):
        assert arel(g[k], o[k]) < 1e-12, k
    assert abs(g["sse"] - o["sse"]) <= 1e-12 * o["sse"]
    assert g["anomalies"] == 0


def test_linearize_mixed_stores_fp32_w(gba, orc):
    prob = gba.generate(2, 2)
    g = gba.linearize(prob, gba.PRECISION_MIXED)
    o = orc.linearize(prob)
    assert arel(g["W"], o["W"]) < 1e-6
    for k in ("U", "V", "bc", "bp"):
        assert arel(g[k], o[k]) < 1e-12, k


def test_linearize_anomaly_skipped(gba, orc):
    prob = gba.generate(2, 1)
    prob.obs_pixel[17, 0] = np.inf
    prob.obs_pixel[3000, 1] = np.nan
    g = gba.linearize(prob)
    o = orc.linearize(prob)
    assert g["anomalies"] == o["anomalies"] == 2
    for k in ("U", "V", "W", "bc", "bp"):
        assert arel(g[k], o[k]) < 1e-12, k
    s, an = gba.cost(prob)
    so, ano = orc.cost(prob)
    assert an == ano == 2 and abs(s - so) <= 1e-12 * so


@pytest.mark.parametrize("which,lam", [("cfg2", 1e-4), ("heavy", 1e-3), ("cfg1", 1e-2)])
def test_schur_step_fp64_matches_oracle(gba, orc, which, lam):
    prob = {"cfg2": lambda: gba.generate(2, 1), "heavy": lambda: heavy(gba),
            "cfg1": lambda: gba.generate(1, 1)}[which]()
    g = gba.schur_step(prob, lam, want_S=True)
    o = orc.schur_step(prob, orc.linearize(prob), lam, want_S=True)
    assert g["status"] == "accept" and o["status"] == "ok"
    assert arel(np.tril(g["S"]), np.tril(o["S"])) < 1e-12
    assert arel(g["rhs"], o["rhs"]) < 1e-12
    assert arel(g["dc"], o["dc"]) < 1e-9
    assert arel(g["dp"], o["dp"]) < 1e-9
    assert abs(g["denom"] - o["denom"]) <= 1e-9 * abs(o["denom"])


def shuffled(gba, seed):
    """cfg2 with its observations in a random order and one observation
    repeated (same point, same camera): the device sort must order by
    (point, camera) with ties in the caller's order."""
    p = gba.generate(2, 1)
    rng = np.random.default_rng(seed)
    order = rng.permutation(p.num_observations)
    order = np.concatenate([order, order[:1]])
    return gba.BALProblem(p.cameras.copy(), p.points.copy(), p.obs_camera[order].copy(),
                          p.obs_point[order].copy(), p.obs_pixel[order].copy())


@pytest.mark.parametrize("seed", [3, 4])
def test_layout_any_observation_order_matches_oracle(gba, orc, seed):
    prob = shuffled(gba, seed)
    g = gba.linearize(prob)
    o = orc.linearize(prob)
    for k in ("U", "V", "W", "bc", "bp"):
        assert arel(g[k], o[k]) < 1e-12, k
    gs = gba.schur_step(prob, 1e-4, want_S=True)
    os_ = orc.schur_step(prob, o, 1e-4, want_S=True)
    assert gs["status"] == "accept" and os_["status"] == "ok"
    assert arel(np.tril(gs["S"]), np.tril(os_["S"])) < 1e-12
    assert arel(gs["dc"], os_["dc"]) < 1e-9
    assert arel(gs["dp"], os_["dp"]) < 1e-9


@pytest.mark.parametrize("C", [1, 2])
def test_one_and_two_camera_problems_match_oracle(gba, orc, C):
    """The smallest layouts: one camera (S is one 9x9 block, no pair tasks;
    every point seen once, so the cost falls to rounding level) and two
    cameras (one pair task).  Costs compared to the initial cost's scale."""
    p = gba.generate(2, 1)
    keep = p.obs_camera < C
    used = np.unique(p.obs_point[keep])
    remap = -np.ones(p.num_points, np.int32)
    remap[used] = np.arange(used.size, dtype=np.int32)
    q = gba.BALProblem(p.cameras[:C].copy(), p.points[used].copy(), p.obs_camera[keep].copy(),
                       remap[p.obs_point[keep]].copy(), p.obs_pixel[keep].copy())
    opts = dict(initial_lambda=1e-4, max_iterations=5)
    r = gba.solve_bal_gpu(q.copy(), gba.LMOptions(**opts))
    o = orc.lm_solve(q.copy(), **opts)
    assert r.termination == o["termination"]
    assert abs(r.iterations - o["iterations"]) <= 1
    assert abs(r.initial_cost - o["initial_cost"]) <= 1e-12 * o["initial_cost"]
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["initial_cost"]


@pytest.mark.parametrize("cfg,lam", [(2, 1e-4), (1, 1e-3)])
def test_schur_step_mixed_close_to_fp64(gba, orc, cfg, lam):
    """The mixed step: the reduced camera system assembled from FP32 W and
    stored in FP32 is solved in FP64, then one FP64 iterative-refinement
    correction solves the residual of the FP64 system (evaluated from the
    Jacobians, without W).  Against the FP64 system (the oracle's S and rhs,
    the reference's src/lm.cu:515-604) the step's residual is then far below
    the FP32 system's distance from it (about 2e-8 here): this test measures
    it and bounds it."""
    prob = gba.generate(cfg, 1)
    g = gba.schur_step(prob, lam, gba.PRECISION_MIXED, want_S=True)
    o = orc.schur_step(prob, orc.linearize(prob), lam, want_S=True)
    assert g["status"] == "accept"
    assert arel(np.tril(g["S"]), np.tril(o["S"])) < 1e-5
    assert arel(g["rhs"], o["rhs"]) < 1e-6  # FP64 arithmetic over FP32-stored W
    S64 = np.tril(o["S"]) + np.tril(o["S"], -1).T
    res64 = np.linalg.norm(S64 @ g["dc"] - o["rhs"]) / np.linalg.norm(o["rhs"])
    gap = arel(g["dc"], o["dc"])
    print(f"cfg {cfg}: mixed step vs FP64 system: relative residual {res64:.2e}, dc gap {gap:.2e}")
    # measured on MI355X: residual 1.1e-10 / 6.4e-10 (cfg 2 / 1); 2.4e-8 for
    # the unrefined FP32-system solution
    assert res64 < 2e-9
    assert gap < 1e-5
    assert abs(g["denom"] - o["denom"]) <= 1e-4 * abs(o["denom"])


def test_schur_step_reports_non_spd_v(gba, orc):
    # a point without observations has V = 0; at lambda = 0 its damped block
    # is not SPD and the trial is abandoned (src/lm.cu:477-481)
    prob = gba.generate(2, 1)
    prob.obs_point[prob.obs_point == 5] = 6
    g = gba.schur_step(prob, 0.0)
    _, bad = orc.damp_invert_v(orc.linearize(prob)["V"], 0.0)
    assert bad >= 1
    assert g["status"] == "nonspd_v"


@pytest.mark.parametrize("n", [1, 63, 64, 65, 441, 1000])
def test_dense_spd_solve_fp64(gba, n):
    rng = np.random.default_rng(n)
    M = rng.uniform(-1, 1, (n, n))
    A = M.T @ M + n * 0.05 * np.eye(n) + np.eye(n)
    b = rng.normal(size=n)
    x, ok, _ = gba.dense_spd_solve(A, b)
    assert ok
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b) * np.linalg.cond(A)
    assert arel(x, np.linalg.solve(A, b)) < 1e-10


# 1, 4, 7, 16 and 32 tile columns: the explicit inverse X = L^-1 that the
# refinement applies is formed inside the factor's launches, row by row, and
# the last 32 columns take their TRSM inside the trailing update
@pytest.mark.parametrize("n", [64, 200, 441, 1000, 2000])
def test_dense_spd_solve_mixed_refines_to_fp64(gba, n):
    rng = np.random.default_rng(n + 1)
    M = rng.uniform(-1, 1, (n, n))
    A = (M.T @ M + n * 0.05 * np.eye(n) + np.eye(n)).astype(np.float32).astype(np.float64)
    b = rng.normal(size=n)
    x, ok, it = gba.dense_spd_solve(A, b, gba.PRECISION_MIXED)
    assert ok and 1 <= it <= 6  # a wrong factor (e.g. a permuted MFMA layout) would not refine this fast
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)


# Grouped factor (large matrices): a group of panel columns takes its TRSMs
# and in-group updates column by column, then one update of the rest of the
# trailing matrix by the whole group.  Forced on small matrices here
# (GBA_FACTOR_GROUP_TAIL=0), every group width, both precisions; and at its
# default threshold on a matrix large enough to use it (n = 5,000: 79 tile
# columns, three groups of 4, then column by column).
@pytest.mark.parametrize("group", [2, 3, 4])
@pytest.mark.parametrize("n", [441, 1000, 2000])
def test_dense_spd_solve_grouped_factor(gba, monkeypatch, group, n):
    monkeypatch.setenv("GBA_FACTOR_GROUP", str(group))
    monkeypatch.setenv("GBA_FACTOR_GROUP_TAIL", "0")
    rng = np.random.default_rng(n + group)
    M = rng.uniform(-1, 1, (n, n))
    A = M.T @ M + n * 0.05 * np.eye(n) + np.eye(n)
    b = rng.normal(size=n)
    x, ok, _ = gba.dense_spd_solve(A, b)
    assert ok
    assert arel(x, np.linalg.solve(A, b)) < 1e-10
    A32 = A.astype(np.float32).astype(np.float64)
    x, ok, it = gba.dense_spd_solve(A32, b, gba.PRECISION_MIXED)
    assert ok and 1 <= it <= 6
    assert np.linalg.norm(A32 @ x - b) <= 1e-10 * np.linalg.norm(b)


def test_dense_spd_solve_grouped_at_default_threshold(gba, monkeypatch):
    n = 5000
    rng = np.random.default_rng(11)
    M = rng.uniform(-1, 1, (n, n))
    A = M.T @ M + n * 0.05 * np.eye(n) + np.eye(n)
    b = rng.normal(size=n)
    x, ok, _ = gba.dense_spd_solve(A, b)
    assert ok
    assert arel(x, np.linalg.solve(A, b)) < 1e-10
    monkeypatch.setenv("GBA_FACTOR_GROUP", "1")
    x1, ok1, _ = gba.dense_spd_solve(A, b)
    assert ok1 and arel(x, x1) < 1e-11


def test_dense_spd_solve_rejects_indefinite(gba):
    A = np.eye(100)
    A[70, 70] = -1.0
    _, ok, _ = gba.dense_spd_solve(A, np.ones(100))
    assert not ok


def test_batch_kernels_reference_protocol(gba):
    import torch
    rng = np.random.default_rng(42)
    n = 200
    M = rng.uniform(-1, 1, (n, 9, 9))
    A = M.transpose(0, 2, 1) @ M + 9 * np.eye(9)
    A[7, 3, 3] = -5.0  # not SPD
    b = rng.uniform(-1, 1, (n, 9))
    x, ok = gba.batch_chol9_factor_solve(torch.tensor(A, device="cuda"), torch.tensor(b, device="cuda"))
    ok = ok.cpu().numpy()
    assert ok[7] == 0 and ok.sum() == n - 1
    good = ok == 1
    ref = np.linalg.solve(A[good], b[good][..., None])[..., 0]
    assert np.max(np.abs(x.cpu().numpy()[good] - ref)) <= 1e-9
    M3 = rng.uniform(-1, 1, (n, 3, 3))
    A3 = M3.transpose(0, 2, 1) @ M3 + 3 * np.eye(3)
    Ai, ok3 = gba.batch_inverse3_spd(torch.tensor(A3, device="cuda"))
    assert ok3.cpu().numpy().all()
    assert np.max(np.abs(Ai.cpu().numpy() - np.linalg.inv(A3))) <= 1e-9


@pytest.mark.parametrize("cfg,lam", [(1, 1e-3), (2, 1e-4)])
def test_lm_fp64_matches_oracle(gba, orc, cfg, lam):
    p_gpu, p_cpu = gba.generate(cfg, 1), gba.generate(cfg, 1)
    r = gba.solve_bal_gpu(p_gpu, gba.LMOptions(initial_lambda=lam))
    o = orc.lm_solve(p_cpu, initial_lambda=lam)
    assert r.termination == o["termination"] == "CONVERGENCE"
    assert r.iterations == o["iterations"]
    assert r.anomalies == o["anomalies"] == 0
    assert abs(r.initial_cost - o["initial_cost"]) <= 1e-12 * o["initial_cost"]
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]
    assert arel(p_gpu.cameras, p_cpu.cameras) < 1e-6
    # accepted-cost trajectory
    ga = r.trace[r.trace[:, 4] == 0][:, 2]
    oa = o["trace"][o["trace"][:, 4] == 0][:, 2]
    assert len(ga) == len(oa) and np.max(np.abs(ga - oa) / oa) < 1e-9
    # every trial: same iteration, status and lambda as the oracle (whose
    # schedule is pinned against the reference's rules on the CPU)
    assert r.trace.shape[0] == o["trace"].shape[0] == r.trials
    assert np.array_equal(r.trace[:, [0, 4]], o["trace"][:, [0, 4]])
    assert np.allclose(r.trace[:, 1], o["trace"][:, 1], rtol=1e-9, atol=0)


def test_lm_schedule_with_rejections_matches_oracle(gba, orc):
    """Rejection runs inside iterations (lambda *= nu, nu *= 2) and accepts on
    both branches of max(1/3, .): the GPU's trial-by-trial lambda equals the
    oracle's."""
    s = gba.synth_spec(1, 1)
    s.num_cameras, s.num_points, s.num_observations, s.kmin, s.kmax = 10, 400, 1600, 2, 8
    s.perturb = 0.2
    prob = gba.generate(spec=s, seed=2)
    opts = dict(initial_lambda=1e-6, max_iterations=30)
    r = gba.solve_bal_gpu(prob.copy(), gba.LMOptions(**opts))
    o = orc.lm_solve(prob.copy(), **opts)
    assert np.sum(o["trace"][:, 4] != 0) >= 4  # the scenario does reject
    assert r.termination == o["termination"] and r.iterations == o["iterations"]
    assert np.array_equal(r.trace[:, [0, 4]], o["trace"][:, [0, 4]])
    # 20% parameter noise: rho carries the rounding of a far-from-quadratic
    # step, and accepts on the 1 - (2 rho - 1)^3 branch pass it into lambda
    # (d factor / d rho = -6 (2 rho - 1)^2), so lambda agrees to 1e-7 here
    # (1.4e-9 measured) while every decision is the same
    assert np.allclose(r.trace[:, 1], o["trace"][:, 1], rtol=1e-7, atol=0)
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]


def with_unobserved_point(gba):
    p = gba.generate(2, 3)
    p.points = np.vstack([p.points[:100], [[1.25, -2.5, 0.75]], p.points[100:]])
    p.obs_point = np.where(p.obs_point >= 100, p.obs_point + 1, p.obs_point).astype(np.int32)
    return p.normalize()


def test_unobserved_point_keeps_its_value(gba, orc):
    """A point no camera sees has V = 0, bp = 0, so dp = 0 at any lambda > 0
    (src/lm.cu:289-301,645-646): it must come back unchanged, single and
    batched, after any number of accepted steps."""
    p = with_unobserved_point(gba)
    opts = dict(initial_lambda=1e-4)
    o = orc.lm_solve(p.copy(), **opts)
    assert o["iterations"] >= 2
    for prec in (gba.PRECISION_FP64, gba.PRECISION_MIXED):
        q = p.copy()
        r = gba.solve_bal_gpu(q, gba.LMOptions(**opts), prec)
        assert r.termination == "CONVERGENCE" and r.iterations >= 2
        assert np.array_equal(q.points[100], p.points[100])
    q = p.copy()
    r = gba.solve_bal_gpu(q, gba.LMOptions(**opts))
    assert r.iterations == o["iterations"]
    ref = p.copy()
    orc.lm_solve(ref, **opts)
    assert arel(q.points, ref.points) < 1e-6
    b = gba.Batch([p, gba.generate(5, 2), p])
    res = b.solve(gba.LMOptions(**opts))
    for i in (0, 2):
        assert res[i].iterations == o["iterations"]
        cams_b, pts_b = b.params(i)
        assert np.array_equal(pts_b[100], p.points[100])
        assert arel(pts_b, ref.points) < 1e-6


def test_batch_zero_inner_trials_is_failure(gba):
    """max_inner_trials = 0: no trial runs, every problem fails at its initial
    cost (the reference's inner loop runs zero times, src/lm.cu:462,704-707)."""
    probs = [gba.generate(5, 4), gba.generate(5, 5)]
    res = gba.Batch(probs).solve(gba.LMOptions(max_inner_trials=0))
    for r in res:
        assert r.termination == "FAILURE" and r.iterations == 0 and r.trials == 0
        assert r.final_cost == r.initial_cost


def test_lm_mixed_reaches_fp64_cost(gba):
    p64, p32 = gba.generate(2, 1), gba.generate(2, 1)
    opts = gba.LMOptions(initial_lambda=1e-4)
    r64 = gba.solve_bal_gpu(p64, opts)
    r32 = gba.solve_bal_gpu(p32, opts, gba.PRECISION_MIXED)
    assert r32.termination == "CONVERGENCE"
    assert abs(r32.final_cost - r64.final_cost) <= 1e-6 * r64.final_cost
    # the FP64 system's residual after one correction per trial; every trial
    # evaluated it at least once
    assert r32.max_refine_residual <= 1e-7
    assert r32.refine_outer_steps >= r32.trials


def test_lm_ragged_problem_matches_oracle(gba, orc):
    q = ragged(gba)
    assert np.bincount(q.obs_camera, minlength=q.num_cameras)[-1] == 0
    opts = dict(initial_lambda=1e-4)
    r = gba.solve_bal_gpu(q.copy(), gba.LMOptions(**opts))
    o = orc.lm_solve(q.copy(), **opts)
    assert r.termination == o["termination"] == "CONVERGENCE"
    assert r.iterations == o["iterations"] and r.anomalies == 0
    assert abs(r.initial_cost - o["initial_cost"]) <= 1e-12 * o["initial_cost"]
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]
    r32 = gba.solve_bal_gpu(q.copy(), gba.LMOptions(**opts), gba.PRECISION_MIXED)
    assert r32.termination == "CONVERGENCE"
    assert abs(r32.final_cost - r.final_cost) <= 1e-6 * r.final_cost


def test_lm_failure_and_max_iterations_paths(gba):
    p = gba.generate(2, 1)
    r = gba.solve_bal_gpu(p, gba.LMOptions(max_inner_trials=0))
    assert r.termination == "FAILURE" and r.iterations == 0 and r.final_cost == r.initial_cost
    r = gba.solve_bal_gpu(gba.generate(2, 1), gba.LMOptions(max_iterations=2, rel_decrease_tol=0.0))
    assert r.termination == "MAX_ITERATIONS" and r.iterations == 2


def test_lm_anomalies_match_oracle(gba, orc):
    p_gpu, p_cpu = gba.generate(2, 1), gba.generate(2, 1)
    for p in (p_gpu, p_cpu):
        p.obs_pixel[10, 1] = np.nan
    r = gba.solve_bal_gpu(p_gpu, gba.LMOptions(max_iterations=3))
    o = orc.lm_solve(p_cpu, max_iterations=3)
    assert r.anomalies == o["anomalies"] > 0
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]


def test_launch_bounds_gate_on_gpu(gba):
    # reference tests/test_launch_bounds.cpp:82-95 (P = 400000 > 1024 * 256)
    prob = gba.generate_launch_bounds(8, 400000, 7)
    r = gba.solve_bal_gpu(prob)
    assert r.initial_cost > 1.0
    assert r.termination == "CONVERGENCE"
    assert r.anomalies == 0
    assert r.final_cost < 1e-4 * r.initial_cost


def test_mixed_refinement_preconditioners_agree(gba, monkeypatch):
    """The mixed refinement preconditioned by the factor's triangular solves
    (what large matrices use) against the explicit inverse (small ones): both
    solve each trial's FP32-stored system and refine against the FP64 one to
    the same tolerances, so the LM trajectory is the same."""
    prob = heavy(gba)
    opts = gba.LMOptions(initial_lambda=1e-3)
    monkeypatch.setenv("GBA_PRECOND", "inverse")
    inv = gba.Solver(prob, gba.PRECISION_MIXED).solve(opts)
    s_inv = gba.schur_step(prob, 1e-3, gba.PRECISION_MIXED)
    monkeypatch.setenv("GBA_PRECOND", "trsv")
    tri = gba.Solver(prob, gba.PRECISION_MIXED).solve(opts)
    s_tri = gba.schur_step(prob, 1e-3, gba.PRECISION_MIXED)
    assert s_tri["status"] == s_inv["status"] == "accept"
    assert arel(s_tri["dc"], s_inv["dc"]) < 1e-8
    assert tri.termination == inv.termination and tri.iterations == inv.iterations
    assert tri.max_refine_residual <= 1e-7 and inv.max_refine_residual <= 1e-7
    assert abs(tri.final_cost - inv.final_cost) <= 1e-9 * inv.final_cost


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_one_launch_factor_matches_launch_per_column(gba, monkeypatch, prec):
    """GBA_FACTOR=persistent (one launch: a critical-chain workgroup plus
    ticketed workers, protocol checked on the CPU by test_persist_protocol)
    against the default launch per tile column: the same tile products, so
    the same Schur step and the same LM trajectory."""
    P = gba.PRECISION_FP64 if prec == "fp64" else gba.PRECISION_MIXED
    prob = heavy(gba)
    opts = gba.LMOptions(initial_lambda=1e-3)
    ref = gba.Solver(prob, P).solve(opts)
    s_ref = gba.schur_step(prob, 1e-3, P)
    monkeypatch.setenv("GBA_FACTOR", "persistent")
    one = gba.Solver(prob, P).solve(opts)
    s_one = gba.schur_step(prob, 1e-3, P)
    assert s_one["status"] == s_ref["status"] == "accept"
    assert arel(s_one["dc"], s_ref["dc"]) < (1e-9 if prec == "fp64" else 1e-6)
    assert one.termination == ref.termination and one.iterations == ref.iterations
    assert abs(one.final_cost - ref.final_cost) <= 1e-9 * ref.final_cost
