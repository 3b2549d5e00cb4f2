"""Block-Jacobi conjugate gradients (k_cg.hip), the factor-free inner solver
of the mixed form: the same FP32-stored reduced camera system as the factor
preconditioner, solved to the same tolerances and refined against the same
FP64 residual, so the step agrees with the factor-preconditioned one, and the
LM trajectory and final cost follow.  Against the FP64 system (the oracle,
the reference's src/lm.cu:515-604) the refined step's residual is far below
the FP32 system's distance from it."""
import numpy as np
import pytest

from test_gpu_path import arel, heavy, ragged

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,lam", [(2, 1e-4), (1, 1e-3)])
def test_cg_step_matches_factor_step(gba, orc, monkeypatch, cfg, lam):
    prob = gba.generate(cfg, 1)
    monkeypatch.setenv("GBA_MIXED_PRECOND", "factor")
    f = gba.schur_step(prob, lam, gba.PRECISION_MIXED, want_S=True)
    monkeypatch.setenv("GBA_MIXED_PRECOND", "jacobi")
    j = gba.schur_step(prob, lam, gba.PRECISION_MIXED, want_S=True)
    assert f["status"] == j["status"] == "accept"
    # the same FP32 assembly (written to the lower copy alone in the Jacobi form)
    assert np.array_equal(np.tril(j["S"]), np.tril(f["S"]))
    assert np.array_equal(j["rhs"], f["rhs"])
    gap = arel(j["dc"], f["dc"])
    S = np.tril(j["S"]) + np.tril(j["S"], -1).T
    res32 = np.linalg.norm(S @ j["dc"] - j["rhs"]) / np.linalg.norm(j["rhs"])
    o = orc.schur_step(prob, orc.linearize(prob), lam, want_S=True)
    S64 = np.tril(o["S"]) + np.tril(o["S"], -1).T
    res64 = np.linalg.norm(S64 @ j["dc"] - o["rhs"]) / np.linalg.norm(o["rhs"])
    print(f"cfg {cfg}: Jacobi vs factor step {gap:.2e}; residual of the FP32-stored system {res32:.2e}, "
          f"of the FP64 system {res64:.2e}")
    assert gap < 1e-6
    # refined against the FP64 system: measured 1.1e-10 / 6.4e-10 there, and
    # 2.4e-8 against the FP32 system it started from
    assert res64 < 2e-9
    assert res64 < 0.1 * res32
    assert abs(j["denom"] - f["denom"]) <= 1e-8 * abs(f["denom"])


@pytest.mark.parametrize("which", ["cfg2", "heavy", "ragged"])
def test_cg_lm_matches_factor_lm(gba, which):
    prob = {"cfg2": lambda: gba.generate(2, 1), "heavy": lambda: heavy(gba), "ragged": lambda: ragged(gba)}[which]()
    opts = gba.LMOptions(initial_lambda=1e-4)
    fs = gba.Solver(prob, gba.PRECISION_MIXED, precond=gba.PRECOND_FACTOR)
    js = gba.Solver(prob, gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI)
    assert fs.precond() == gba.PRECOND_FACTOR and js.precond() == gba.PRECOND_JACOBI
    f, j = fs.solve(opts), js.solve(opts)
    print(f"{which}: Jacobi {j.refine_iterations} steps over {j.trials} trials, factor {f.refine_iterations}; "
          f"final costs {j.final_cost:.12g} / {f.final_cost:.12g}")
    assert j.termination == f.termination == "CONVERGENCE"
    assert j.iterations == f.iterations and j.trials == f.trials
    assert np.array_equal(j.trace[:, 4], f.trace[:, 4])
    assert abs(j.final_cost - f.final_cost) <= 1e-9 * f.final_cost
    assert j.refine_factor_trials == 0 and f.refine_factor_trials == f.trials
    # the FP64 system's residual after one correction per trial (measured
    # 4e-9 to 7e-9, the late trials: lambda near 1e-6)
    assert 0.0 < j.max_refine_residual <= 1e-7 and 0.0 < f.max_refine_residual <= 1e-7
    assert j.refine_outer_steps >= j.trials
    # repeatable: fixed-order sums throughout
    j2 = js.solve(opts)
    assert j2.final_cost == j.final_cost and j2.refine_iterations == j.refine_iterations


def test_cg_step_cap_falls_back_to_factor(gba):
    """A trial that misses the tolerance within jacobi_max_iters is refined
    with the FP32 factor instead, and so is the rest of the solve: the result
    is the factor-preconditioned solve's."""
    prob = gba.generate(2, 1)
    opts = gba.LMOptions(initial_lambda=1e-4)
    f = gba.Solver(prob, gba.PRECISION_MIXED, precond=gba.PRECOND_FACTOR).solve(opts)
    c = gba.Solver(prob, gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI, jacobi_max_iters=3).solve(opts)
    assert c.refine_factor_trials == c.trials == f.trials
    assert c.iterations == f.iterations
    assert abs(c.final_cost - f.final_cost) <= 1e-12 * f.final_cost


def test_cg_camera_block_not_spd_is_a_factor_failure(gba, orc):
    """lambda = 0 with an unobserved camera: its S block is zero, not SPD.  The
    block inverse raises the failed-factorisation status, as the reference's
    LLT failure does (src/lm.cu:597-601): the oracle's trial statuses."""
    p = gba.generate(2, 1)
    q = gba.BALProblem(np.vstack([p.cameras, p.cameras[:1] * 1.01]), p.points.copy(), p.obs_camera.copy(),
                       p.obs_point.copy(), p.obs_pixel.copy())
    opts = dict(initial_lambda=0.0, max_iterations=2)
    r = gba.Solver(q, gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI).solve(gba.LMOptions(**opts))
    o = orc.lm_solve(q.copy(), **opts)
    assert np.array_equal(r.trace[:, 4], o["trace"][:, 4])
    assert r.trace[0, 4] == 3  # GBA_TRIAL_CHOL_FAIL
    assert r.termination == o["termination"]


def test_cg_kernel_timers(gba):
    s = gba.Solver(gba.generate(2, 1), gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI)
    assert s.time_kernel("cg_step", 1e-4, 5) > 0.0
    assert s.time_kernel("cg_setup", 1e-4, 5) > 0.0
    assert s.time_kernel("factor", 1e-4, 2) > 0.0  # the alternative, assembled with its copy
    with pytest.raises(gba.GBAError):
        s.time_kernel("solve", 1e-4, 1)
    f = gba.Solver(gba.generate(2, 1), gba.PRECISION_MIXED, precond=gba.PRECOND_FACTOR)
    with pytest.raises(gba.GBAError):
        f.time_kernel("cg_step", 1e-4, 1)


def test_fp64_refinement_corrections_lower_the_fp64_residual(gba):
    """refine_corrections < 0: the FP64 system's residual is evaluated once per
    trial and reported, no correction; the default (one correction per trial)
    reports a smaller worst residual for the same final cost within 1e-6, and
    evaluates it once more per corrected trial."""
    prob = gba.generate(2, 1)
    opts = gba.LMOptions(initial_lambda=1e-4)
    none = gba.Solver(prob, gba.PRECISION_MIXED, refine_corrections=-1).solve(opts)
    one = gba.Solver(prob, gba.PRECISION_MIXED).solve(opts)
    print(f"worst FP64 residual: no correction {none.max_refine_residual:.2e}, one {one.max_refine_residual:.2e}")
    assert none.refine_outer_steps == none.trials
    assert one.refine_outer_steps > one.trials
    assert 0.0 < one.max_refine_residual < 0.1 * none.max_refine_residual
    assert abs(one.final_cost - none.final_cost) <= 1e-6 * none.final_cost


@pytest.mark.parametrize("cfg,lam", [(2, 1e-4), (1, 1e-3)])
def test_reported_fp64_residual_matches_oracle_system(gba, orc, cfg, lam):
    """max_refine_residual is |D (rhs - S dc)| / |D rhs| of the FP64 system,
    evaluated on the device without W.  Recomputed here from the oracle's S
    and rhs (the reference's src/lm.cu:515-604) and D = (diag(U) + lambda
    diag'(U))^-1/2 for the first trial's step (the step schur_step returns),
    it matches what a one-trial solve reports."""
    prob = gba.generate(cfg, 1)
    g = gba.schur_step(prob, lam, gba.PRECISION_MIXED)
    assert g["status"] == "accept"
    lin = orc.linearize(prob)
    o = orc.schur_step(prob, lin, lam, want_S=True)
    S = np.tril(o["S"]) + np.tril(o["S"], -1).T
    d = np.concatenate([np.diag(u) for u in lin["U"]])
    D = 1.0 / np.sqrt(d + lam * np.maximum(d, 1e-12))
    ref = np.linalg.norm(D * (o["rhs"] - S @ g["dc"])) / np.linalg.norm(D * o["rhs"])
    r = gba.Solver(prob, gba.PRECISION_MIXED).solve(gba.LMOptions(initial_lambda=lam, max_iterations=1))
    assert r.trials == 1 and r.iterations == 1
    print(f"cfg {cfg}: reported FP64 residual {r.max_refine_residual:.3e}, recomputed {ref:.3e}")
    assert ref < 1e-9
    assert abs(r.max_refine_residual - ref) <= 0.05 * ref + 1e-13


@pytest.mark.parametrize("which", ["cfg2", "heavy", "ragged"])
def test_cg_gauge_deflation(gba, which):
    """Deflating CG by the seven similarity-gauge directions solves the same
    systems to the same tolerances: the LM trajectory and final cost are those
    of plain block-Jacobi CG, in far fewer steps (the preconditioned operator's
    seven eigenvalues of the order of lambda are the gauge; measured on the
    oracle's S at C=49: 33 -> 14 steps to 1e-8 at lambda 1e-4, 42 -> 15 at
    3.7e-6).  A wrong gauge basis would still solve correctly but no faster,
    so the step count is what pins the kernels' basis."""
    prob = {"cfg2": lambda: gba.generate(2, 1), "heavy": lambda: heavy(gba), "ragged": lambda: ragged(gba)}[which]()
    opts = gba.LMOptions(initial_lambda=1e-4)
    ps = gba.Solver(prob, gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI, deflation=False)
    ds = gba.Solver(prob, gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI)
    p, d = ps.solve(opts), ds.solve(opts)
    print(f"{which}: deflated {d.refine_iterations} steps, plain {p.refine_iterations} over {d.trials} trials; "
          f"final costs {d.final_cost:.12g} / {p.final_cost:.12g}; FP64 residuals {d.max_refine_residual:.2e} / "
          f"{p.max_refine_residual:.2e}")
    assert d.termination == p.termination == "CONVERGENCE"
    assert d.iterations == p.iterations and d.trials == p.trials
    assert np.array_equal(d.trace[:, 4], p.trace[:, 4])
    assert abs(d.final_cost - p.final_cost) <= 1e-9 * p.final_cost
    assert d.refine_factor_trials == 0
    assert 0.0 < d.max_refine_residual <= 1e-7
    assert d.refine_iterations < 0.7 * p.refine_iterations
    d2 = ds.solve(opts)  # fixed-order sums: repeatable bit for bit
    assert d2.final_cost == d.final_cost and d2.refine_iterations == d.refine_iterations


def test_cg_deflation_rank_deficient_basis_is_plain_cg(gba):
    """One camera: its scale column lies in the span of its translation
    columns, so Z^T A Z is singular, fails the pivot test and leaves
    E^-1 = 0 -- the deflated kernels then compute exactly the plain block-
    Jacobi iterates (x_0 = 0, r_0 = b, mu = 0), bit for bit."""
    p = gba.generate(2, 1)
    keep = p.obs_camera < 1
    used = np.unique(p.obs_point[keep])
    remap = -np.ones(p.num_points, np.int32)
    remap[used] = np.arange(used.size, dtype=np.int32)
    q = gba.BALProblem(p.cameras[:1].copy(), p.points[used].copy(), p.obs_camera[keep].copy(),
                       remap[p.obs_point[keep]].copy(), p.obs_pixel[keep].copy())
    opts = gba.LMOptions(initial_lambda=1e-4, max_iterations=5)
    d = gba.Solver(q, gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI).solve(opts)
    pl = gba.Solver(q, gba.PRECISION_MIXED, precond=gba.PRECOND_JACOBI, deflation=False).solve(opts)
    assert d.refine_iterations == pl.refine_iterations > 0
    assert d.final_cost == pl.final_cost and d.iterations == pl.iterations
    assert np.array_equal(d.trace[:, [1, 2, 4]], pl.trace[:, [1, 2, 4]])
