"""The heavy points' share of the reduced camera system as one dense MFMA
product (k_heavy.hip): points seen at least heavy_k times leave the pair
lists and the diagonal-block sums, and S -= Y Y^T with Y = D W L over them.
The assembled S, rhs and step must be the oracle's (the restated
src/lm.cu:515-604, which sums every observation pair) to the same bounds as
the sparse kernels alone; GBA_SCHUR_HEAVY_K forces the threshold down so
that small problems (and a point a camera observes twice) take the path."""
import numpy as np
import pytest

from test_gpu_path import arel, heavy, shuffled

pytestmark = pytest.mark.gpu


def pick(gba, which):
    return {"cfg2": lambda: gba.generate(2, 1), "heavy": lambda: heavy(gba),
            "repeated": lambda: shuffled(gba, 3)}[which]()


@pytest.mark.parametrize("which,kh", [("cfg2", 8), ("cfg2", 20), ("heavy", 16), ("repeated", 8)])
def test_heavy_schur_fp64_matches_oracle(gba, orc, monkeypatch, which, kh):
    prob = pick(gba, which)
    k = np.bincount(prob.obs_point, minlength=prob.num_points)
    assert (k >= kh).sum() > 0  # the dense product is in use
    monkeypatch.setenv("GBA_SCHUR_HEAVY_K", str(kh))
    lam = 1e-4
    g = gba.schur_step(prob, lam, want_S=True)
    o = orc.schur_step(prob, orc.linearize(prob), lam, want_S=True)
    assert g["status"] == "accept" and o["status"] == "ok"
    print(f"{which} k >= {kh}: {(k >= kh).sum()} heavy points; S {arel(np.tril(g['S']), np.tril(o['S'])):.1e}, "
          f"dc {arel(g['dc'], o['dc']):.1e}")
    assert arel(np.tril(g["S"]), np.tril(o["S"])) < 1e-12
    assert arel(g["rhs"], o["rhs"]) < 1e-12
    assert arel(g["dc"], o["dc"]) < 1e-9
    assert arel(g["dp"], o["dp"]) < 1e-9
    assert abs(g["denom"] - o["denom"]) <= 1e-9 * abs(o["denom"])


@pytest.mark.parametrize("which,kh", [("cfg2", 8), ("repeated", 8)])
def test_heavy_schur_mixed_matches_sparse(gba, orc, monkeypatch, which, kh):
    """Mixed form: Y is FP32 and the product accumulates in FP32, so S moves by
    FP32 rounding of the scaled entries (unit diagonal) against the sparse
    kernels' fixed-point sums; the refined step still meets the FP64 system."""
    prob = pick(gba, which)
    lam = 1e-4
    monkeypatch.setenv("GBA_SCHUR_HEAVY_K", "0")
    s = gba.schur_step(prob, lam, gba.PRECISION_MIXED, want_S=True)
    monkeypatch.setenv("GBA_SCHUR_HEAVY_K", str(kh))
    h = gba.schur_step(prob, lam, gba.PRECISION_MIXED, want_S=True)
    assert s["status"] == h["status"] == "accept"
    o = orc.schur_step(prob, orc.linearize(prob), lam, want_S=True)
    S64 = np.tril(o["S"]) + np.tril(o["S"], -1).T
    d = 1.0 / np.sqrt(np.abs(np.diag(S64)))
    # compare in the scaled units the FP32 copy is stored in
    diff = np.abs(np.tril(h["S"]) - np.tril(s["S"])) * np.outer(d, d)
    res64 = np.linalg.norm(S64 @ h["dc"] - o["rhs"]) / np.linalg.norm(o["rhs"])
    print(f"{which}: max scaled |S_heavy - S_sparse| {diff.max():.1e}; FP64 residual of the step {res64:.1e}")
    assert diff.max() < 2e-5
    assert res64 < 2e-9


def test_heavy_lm_matches_oracle(gba, orc, monkeypatch):
    prob = gba.generate(2, 1)
    monkeypatch.setenv("GBA_SCHUR_HEAVY_K", "8")
    opts = gba.LMOptions(initial_lambda=1e-4)
    r = gba.solve_bal_gpu(prob.copy(), opts)
    o = orc.lm_solve(prob.copy(), initial_lambda=1e-4)
    assert r.termination == o["termination"] == "CONVERGENCE"
    assert r.iterations == o["iterations"]
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]
    m = gba.solve_bal_gpu(prob.copy(), opts, precision=gba.PRECISION_MIXED)
    assert abs(m.final_cost - o["final_cost"]) <= 1e-6 * o["final_cost"]


def test_heavy_handle_query_and_timer(gba, monkeypatch):
    prob = gba.generate(2, 1)
    monkeypatch.setenv("GBA_SCHUR_HEAVY_K", "8")
    s = gba.Solver(prob, gba.PRECISION_MIXED)
    k, n, w = s.heavy()
    assert k == 8 and n == int((np.bincount(prob.obs_point) >= 8).sum()) and w % 16 == 0 and 3 * n <= w < 3 * n + 16
    assert s.time_kernel("heavy_syrk", 1e-4, 3) > 0.0
    monkeypatch.setenv("GBA_SCHUR_HEAVY_K", "0")
    s0 = gba.Solver(prob, gba.PRECISION_FP64)
    assert s0.heavy() == (0, 0, 0)
    with pytest.raises(RuntimeError):
        s0.time_kernel("heavy_syrk", 1e-4, 1)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_heavy_huber_lm_matches_oracle(gba, orc, monkeypatch, prec):
    """With the Huber loss the heavy points' Y is formed from the IRLS-weighted
    W and V' like the pair sums: the robust LM trajectory is the oracle's."""
    prob = gba.generate(2, 1)
    monkeypatch.setenv("GBA_SCHUR_HEAVY_K", "8")
    opts = gba.LMOptions(initial_lambda=1e-4)
    p = gba.PRECISION_FP64 if prec == "fp64" else gba.PRECISION_MIXED
    r = gba.solve_bal_gpu(prob.copy(), opts, precision=p, huber_delta=1.0)
    o = orc.lm_solve(prob.copy(), initial_lambda=1e-4, huber_delta=1.0)
    assert r.termination == o["termination"]
    assert abs(r.final_cost - o["final_cost"]) <= (1e-9 if prec == "fp64" else 1e-6) * o["final_cost"]
    if prec == "fp64":
        assert r.iterations == o["iterations"]
