"""The oracle's LM loop (restating src/lm.cu:371-724).  Its exact trajectory
against the reference binary is parity unpinned (CUDA + Eigen cannot build
here); it is pinned by invariants: the chi-square expectation of the final
cost, monotone accepted costs, the 2x gain-ratio convention, the failure and
anomaly paths, and the reference's own data-free end-to-end gate
(tests/test_launch_bounds.cpp:82-95)."""
import numpy as np
import pytest


def expected_cost(p):
    return (2 * p.num_observations - 9 * p.num_cameras - 3 * p.num_points + 7) / 2.0


@pytest.mark.parametrize("cfg", [1, 2])
def test_lm_converges_to_noise_floor(gba, orc, cfg):
    prob = gba.generate(cfg, 1)
    init = orc.total_squared_error(prob) / 2
    r = orc.lm_solve(prob, initial_lambda=1e-4 if cfg == 2 else 1e-3)
    assert r["termination"] == "CONVERGENCE"
    assert r["anomalies"] == 0
    assert abs(r["initial_cost"] - init) <= 1e-9 * init  # reader-cost gate (bench_bal.cpp:87-94)
    exp = expected_cost(prob)
    sd = np.sqrt(2 * (2 * prob.num_observations - 9 * prob.num_cameras - 3 * prob.num_points)) / 2
    assert abs(r["final_cost"] - exp) < 5 * sd
    acc = r["trace"][r["trace"][:, 4] == 0]
    assert np.all(np.diff(acc[:, 2]) < 0)
    # final parameters reproduce the final cost
    assert abs(orc.total_squared_error(prob) / 2 - r["final_cost"]) <= 1e-8 * r["final_cost"]


def test_gain_ratio_is_twice_textbook_near_quadratic(gba, orc):
    # rho = (sse - sse_trial) / (0.5 * denom) (src/lm.cu:672) -> about 2 on a
    # nearly quadratic model; the textbook ratio would be about 1
    prob = gba.generate(1, 2)
    r = orc.lm_solve(prob)
    acc = r["trace"][r["trace"][:, 4] == 0]
    assert np.all(np.abs(acc[1:3, 3] - 2.0) < 0.05)


def test_zero_inner_trials_is_failure(gba, orc):
    prob = gba.generate(2, 1)
    r = orc.lm_solve(prob, max_inner_trials=0)
    assert r["termination"] == "FAILURE" and r["iterations"] == 0
    assert r["final_cost"] == r["initial_cost"]


def test_zero_iterations_is_max_iterations(gba, orc):
    prob = gba.generate(2, 1)
    r = orc.lm_solve(prob, max_iterations=0)
    assert r["termination"] == "MAX_ITERATIONS" and r["final_cost"] == r["initial_cost"]


def test_nan_pixel_counts_anomalies(gba, orc):
    prob = gba.generate(2, 1)
    prob.obs_pixel[10, 1] = np.nan
    r = orc.lm_solve(prob, max_iterations=3)
    # one skipped edge per linearisation and per trial-cost evaluation
    assert r["anomalies"] >= 2
    assert np.isfinite(r["final_cost"])


def test_launch_bounds_problem_converges(gba, orc):
    # reference tests/test_launch_bounds.cpp: C=8, P=400000, seed 7, noise-free
    prob = gba.generate_launch_bounds(8, 400000, 7)
    r = orc.lm_solve(prob)
    assert r["initial_cost"] > 1.0
    assert r["termination"] == "CONVERGENCE"
    assert r["anomalies"] == 0
    assert r["final_cost"] < 1e-4 * r["initial_cost"]
