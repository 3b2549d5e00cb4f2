"""Mixed-precision batched solver (FP32 W, diagonally scaled reduced camera
system from 64-bit fixed-point sums, FP64 factor, one FP64 refinement step
whose residual re-evaluates the Jacobians; gba_batch_create_precision with
GBA_PRECISION_MIXED) against the FP64 batch and the oracle.  The bar is the
mixed form's: final cost within 1e-6 relative of the FP64 result.  Also
covered: both batched factor forms, anomalies and the
iteration limit, a ragged problem, graph replay against eager launches, and
the per-problem solve times the device clock records."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-6


def close(rm, rf):
    assert rm.termination == rf.termination
    assert abs(rm.initial_cost - rf.initial_cost) <= 1e-6 * rf.initial_cost
    assert abs(rm.final_cost - rf.final_cost) <= TOL * rf.final_cost, (rm.final_cost, rf.final_cost)
    assert abs(rm.iterations - rf.iterations) <= 2
    assert rm.anomalies == rf.anomalies


def test_mixed_batch_matches_fp64_batch_and_oracle(gba, orc):
    probs = [gba.generate(5, seed) for seed in range(1, 17)]
    bf = gba.Batch(probs)
    rf = bf.solve()
    bm = gba.Batch(probs, precision=gba.PRECISION_MIXED)
    rm = bm.solve()
    worst = 0.0
    for a, b in zip(rm, rf):
        close(a, b)
        assert a.termination == "CONVERGENCE"
        worst = max(worst, abs(a.final_cost - b.final_cost) / b.final_cost)
    # parameters of a converged problem agree with the FP64 batch's
    for i in (0, 7, 15):
        cm, pm = bm.params(i)
        cf, pf = bf.params(i)
        assert np.max(np.abs(cm - cf)) <= 1e-3 * max(1.0, np.max(np.abs(cf)))
        assert np.max(np.abs(pm - pf)) <= 1e-3 * max(1.0, np.max(np.abs(pf)))
    # against the oracle (the reference loop, src/lm.cu:458-702) on two problems
    for i in (2, 11):
        o = orc.lm_solve(probs[i].copy())
        assert rm[i].termination == o["termination"]
        assert abs(rm[i].final_cost - o["final_cost"]) <= TOL * o["final_cost"]
    # the same handle solves again: the Schur sums are order-independent
    rm2 = bm.solve()
    for a, b in zip(rm, rm2):
        assert abs(a.final_cost - b.final_cost) <= 1e-9 * a.final_cost
    print(f"mixed batch of 16: worst final-cost gap to the FP64 batch {worst:.1e}")


@pytest.mark.parametrize("factor", ["tiles", "workgroup"])
def test_mixed_batch_factor_forms(gba, monkeypatch, factor):
    monkeypatch.setenv("GBA_BATCH_FACTOR", factor)
    probs = [gba.generate(5, seed) for seed in (21, 22, 23)]
    rf = gba.Batch(probs).solve()
    b = gba.Batch(probs, precision=gba.PRECISION_MIXED)
    assert b.info()["factor"] == factor
    for a, c in zip(b.solve(), rf):
        close(a, c)


def test_mixed_batch_anomalies_limits_and_ragged(gba):
    from test_gpu_path import ragged
    p_nan = gba.generate(2, 1)
    p_nan.obs_pixel[10, 1] = np.nan
    probs = [p_nan, gba.generate(5, 3), ragged(gba)]
    opts = gba.LMOptions(max_iterations=3)
    rf = gba.Batch(probs).solve(opts)
    rm = gba.Batch(probs, precision=gba.PRECISION_MIXED).solve(opts)
    assert rm[0].anomalies == rf[0].anomalies > 0
    assert rm[1].termination == "MAX_ITERATIONS" and rm[1].iterations == 3
    for a, c in zip(rm, rf):
        assert a.termination == c.termination
        assert abs(a.final_cost - c.final_cost) <= 1e-5 * c.final_cost


def test_mixed_batch_graph_replay_matches_eager(gba, monkeypatch):
    probs = [gba.generate(5, seed) for seed in (31, 32, 33, 34)]
    res_g = gba.Batch(probs, precision=gba.PRECISION_MIXED).solve()
    monkeypatch.setenv("GBA_BATCH_GRAPH", "0")
    b = gba.Batch(probs, precision=gba.PRECISION_MIXED)
    assert not b.info()["graphs"]
    res_e = b.solve()
    for a, c in zip(res_g, res_e):
        assert a.termination == c.termination and a.iterations == c.iterations
        assert abs(a.final_cost - c.final_cost) <= 1e-9 * c.final_cost


def test_batch_per_problem_solve_times(gba):
    """Each problem's solve_time_s runs from the start of the batch solve to
    the round that finished it (device clock): positive, no longer than the
    whole batch, and not decreasing in the number of trials a problem took
    beyond the round granularity."""
    probs = [gba.generate(5, seed) for seed in range(41, 49)]
    b = gba.Batch(probs, precision=gba.PRECISION_MIXED)
    b.solve()
    res = b.solve()
    total = max(r.total_time_s for r in res)
    t = np.array([r.solve_time_s for r in res])
    assert np.all(t > 0) and np.all(t <= total * 1.001 + 1e-4)
    trials = np.array([r.trials for r in res])
    # the problem with the most trials finished no earlier than the one with the fewest
    assert t[np.argmax(trials)] >= t[np.argmin(trials)] * 0.999


def test_mixed_batch_512_preset5_matches_fp64_batch(gba):
    """The full batched workload in the mixed form against the FP64 batch:
    every problem's final cost within 1e-6 and the same iteration count.
    The FP64 refinement is what holds this: without it (refinement steps 0)
    the FP32 rounding of W moves the step by percents where lambda is small
    (2.3% on seed 237 at lambda 4e-7, measured against the oracle's FP64
    step), enough to flip a convergence test that sits near its tolerance."""
    probs = [gba.generate(5, seed) for seed in range(1, 513)]
    rf = gba.Batch(probs).solve()
    rm = gba.Batch(probs, precision=gba.PRECISION_MIXED).solve()
    gaps = [abs(a.final_cost - b.final_cost) / b.final_cost for a, b in zip(rm, rf)]
    assert max(gaps) <= TOL, max(gaps)
    assert all(a.iterations == b.iterations and a.termination == b.termination for a, b in zip(rm, rf))
    print(f"512-problem mixed batch: worst final-cost gap to the FP64 batch {max(gaps):.1e}")
