"""The path at the benchmark's full sizes (C=356 / O=1.26M and C=1,936 /
O=3.4M), checked through properties that do not need the oracle to finish a
whole solve: the first LM iteration against the oracle (the restated
reference, which needs a few seconds for one iteration at C=356), the mixed
form's final cost against the plain FP64 solve (the 1e-6 relative target),
its refinement residual (1e-10), and the FP16-pixel solve's final-cost error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg3(gba):
    return gba.generate(3, 1)


def test_cfg3_first_iteration_matches_oracle(gba, orc, cfg3):
    opts = dict(initial_lambda=1e-4, max_iterations=1)
    r = gba.solve_bal_gpu(cfg3.copy(), gba.LMOptions(**opts))
    o = orc.lm_solve(cfg3.copy(), **opts)
    assert r.termination == o["termination"] and r.iterations == o["iterations"] == 1
    assert r.trials == len(o["trace"])
    assert abs(r.initial_cost - o["initial_cost"]) <= 1e-12 * o["initial_cost"]
    # one accepted step: the device's Schur sums run in another order than
    # the oracle's, so the step agrees to rounding, amplified by S's condition
    assert abs(r.final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]


def test_cfg3_mixed_reaches_fp64_cost(gba, cfg3):
    opts = gba.LMOptions(initial_lambda=1e-4)
    r64 = gba.Solver(cfg3, gba.PRECISION_FP64).solve(opts)
    r32 = gba.Solver(cfg3, gba.PRECISION_MIXED).solve(opts)
    assert r64.termination == r32.termination == "CONVERGENCE"
    assert abs(r32.final_cost - r64.final_cost) <= 1e-6 * r64.final_cost
    assert 0.0 < r32.max_refine_residual <= 1e-10
    assert r32.refine_iterations <= 6 * r32.trials
    # the accepted costs decrease monotonically in both forms
    for r in (r64, r32):
        acc = r.trace[r.trace[:, 4] == 0][:, 2]
        assert np.all(np.diff(acc) < 0) and acc[-1] == r.final_cost


def test_cfg4_mixed_and_fp16_pixels(gba):
    p = gba.generate(4, 1)
    assert p.num_points > 262144  # past the reference's launch-size boundary
    opts = gba.LMOptions(initial_lambda=1e-4)
    r64 = gba.Solver(p, gba.PRECISION_FP64).solve(opts)
    r32 = gba.Solver(p, gba.PRECISION_MIXED).solve(opts)
    r16 = gba.Solver(p, gba.PRECISION_MIXED, obs_storage=gba.OBS_FP16).solve(opts)
    assert r64.termination == r32.termination == r16.termination == "CONVERGENCE"
    assert r64.anomalies == r32.anomalies == r16.anomalies == 0
    assert abs(r32.final_cost - r64.final_cost) <= 1e-6 * r64.final_cost
    err16 = (r16.final_cost - r64.final_cost) / r64.final_cost
    print(f"C=1936 FP16 pixels: final-cost error vs FP64 pixels {err16:+.3e}")
    assert abs(err16) <= 1e-3
