"""Pins the oracle's LM control schedule: the lambda / nu sequence, the accept
rule and the termination tests of the reference loop (src/lm.cu:455-714).

Two independent checks stand against the oracle's trace (one row per trial:
outer iteration, lambda used, trial cost, rho, status, nu, gain denominator):

* check_schedule() re-derives every transition from the reference's text:
  nu = 2 at the start of each outer iteration (lm.cu:461); after any trial
  that is not accepted -- non-SPD V (:477-481), LLT failure (:599-603), bad
  denominator (:639-643), non-finite trial (:663-669), plain rejection
  (:687-688) -- lambda *= nu, nu *= 2; after an accept
  lambda *= max(1/3, 1 - (2 rho - 1)^3) (:683); rho = (sse - sse_t) /
  (0.5 denom) (:672); accept iff rho > 0 and sse_t < sse (:678); convergence
  once (sse - sse_t) / max(sse, 1) < tol after an accept (:681,696-699);
  failure after max_inner_trials trials without an accept (:704-707).
* Closed forms on problems built to take one path on every trial: a NaN pixel
  (every trial non-finite), lambda0 = 0 with an unobserved point (non-SPD V)
  or an unobserved camera (LLT failure), a zero-residual problem (denominator
  0).  There lambda_k = lambda0 2^(k(k+1)/2) and nu_k = 2^(k+1) for the
  k-th trial, and the solve fails after exactly 16 trials.

The checks are themselves checked: each of thirteen copies of the oracle with
one rule deliberately broken (gba_oracle.c, ORC_MUTATE = 1..13) must fail at
least one of them.
"""
import concurrent.futures
import math
import os
import subprocess

import numpy as np
import pytest

ACCEPT, REJECT, NONSPD, LLT, DENOM, NONFINITE = range(6)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def small_spec(gba, C=10, P=400, O=1600, noise=1.0, perturb=0.01):
    s = gba.synth_spec(1, 1)
    s.num_cameras, s.num_points, s.num_observations = C, P, O
    s.kmin, s.kmax = 2, min(C, 8)
    s.pixel_noise, s.perturb = noise, perturb
    return s


def check_schedule(r, lambda0, max_iterations, max_inner_trials, tol):
    """Raises AssertionError at the first trial that breaks a rule of the
    reference loop; r is an oracle lm_solve result."""
    tr = r["trace"]
    assert len(tr) == r["trials"]
    cost = r["initial_cost"]  # current cost (sse / 2)
    lam, nu = lambda0, 2.0
    it, inner, accepts = 0, 0, 0
    converged = False
    for k, row in enumerate(tr):
        assert not converged, f"trial {k} after convergence"
        t_it, t_lam, t_cost, t_rho, code, t_nu, t_den = row
        code = int(code)
        assert int(t_it) == it, f"trial {k}: iteration {t_it}, expected {it}"
        assert t_nu == nu, f"trial {k}: nu {t_nu}, expected {nu}"
        assert t_lam == pytest.approx(lam, rel=1e-14, abs=0.0), f"trial {k}: lambda {t_lam}, expected {lam}"
        inner += 1
        assert inner <= max_inner_trials, f"trial {k}: more than {max_inner_trials} trials in iteration {it}"
        if code in (NONSPD, LLT):
            assert math.isnan(t_cost) and math.isnan(t_rho)
        elif code == DENOM:
            assert not (t_den > 0.0) and math.isnan(t_cost)
        elif code == NONFINITE:
            assert t_den > 0.0
        else:
            assert t_den > 0.0 and np.isfinite(t_cost)
            rho = (2 * cost - 2 * t_cost) / (0.5 * t_den)
            assert t_rho == pytest.approx(rho, rel=1e-12, abs=1e-300), f"trial {k}: rho {t_rho} vs {rho}"
            should = rho > 0.0 and t_cost < cost
            assert (code == ACCEPT) == should, f"trial {k}: status {code}, accept rule says {should}"
        if code == ACCEPT:
            rel_dec = (2 * cost - 2 * t_cost) / max(2 * cost, 1.0)
            cost = t_cost
            lam = lam * max(1.0 / 3.0, 1.0 - (2.0 * t_rho - 1.0) ** 3)
            nu = 2.0
            accepts += 1
            it += 1
            inner = 0
            if rel_dec < tol:
                converged = True
        else:
            lam = lam * nu
            nu = 2.0 * nu
    if converged:
        assert r["termination"] == "CONVERGENCE"
    elif accepts == max_iterations:
        assert r["termination"] == "MAX_ITERATIONS"
    else:
        assert inner == max_inner_trials, f"stopped after {inner} trials without an accept"
        assert r["termination"] == "FAILURE"
    assert r["iterations"] == accepts
    assert r["final_cost"] == cost


def closed_form(r, lambda0, code, trials=16):
    """Every trial takes the same failing path: lambda_k = lambda0 2^(k(k+1)/2),
    nu_k = 2^(k+1), FAILURE after exactly `trials` trials."""
    tr = r["trace"]
    assert r["termination"] == "FAILURE" and r["iterations"] == 0
    assert len(tr) == trials, f"{len(tr)} trials"
    k = np.arange(trials)
    assert np.all(tr[:, 4] == code)
    assert np.all(tr[:, 0] == 0)
    assert np.array_equal(tr[:, 5], 2.0 ** (k + 1))
    assert np.allclose(tr[:, 1], lambda0 * 2.0 ** (k * (k + 1) / 2), rtol=1e-15, atol=0)
    assert r["final_cost"] == r["initial_cost"]


# ---- scenarios: (name, problem builder, options, extra check) ---------------

def _with_unobserved_point(p):
    p = p.copy()
    p.points = np.vstack([p.points, [[0.5, -0.25, 1.0]]])
    return p.normalize()


def _with_unobserved_camera(p):
    p = p.copy()
    p.cameras = np.vstack([p.cameras, p.cameras[:1]])
    return p.normalize()


def _zero_residual(orc, p):
    p = p.copy()
    for e in range(p.num_observations):
        p.obs_pixel[e] = orc.edge_residual(p.cameras[p.obs_camera[e]], p.points[p.obs_point[e]], np.zeros(2))
    return p


def scenarios(gba, orc):
    base = gba.generate(spec=small_spec(gba), seed=1)
    rough = gba.generate(spec=small_spec(gba, perturb=0.2), seed=2)
    clean = gba.generate(spec=small_spec(gba, noise=0.0), seed=3)
    nanpx = base.copy()
    nanpx.obs_pixel[7, 0] = np.nan
    out = [
        # accepted steps with rho near 2 (the 1/3 clamp)
        ("converge", base, dict(initial_lambda=1e-3), None),
        # runs of rejections inside an iteration, accepts with rho < 0.937
        ("rejections", rough, dict(initial_lambda=1e-6), "runs"),
        # noise-free: the cost falls below 1, where max(sse, 1) decides convergence
        ("noise_free", clean, dict(initial_lambda=1e-3), "below_one"),
        ("nan_pixel", nanpx, dict(initial_lambda=1e-3), ("closed", NONFINITE)),
        ("nonspd_v", _with_unobserved_point(base), dict(initial_lambda=0.0), ("closed", NONSPD)),
        ("llt_fail", _with_unobserved_camera(base), dict(initial_lambda=0.0), ("closed", LLT)),
        ("denominator", _zero_residual(orc, base), dict(initial_lambda=1e-3), ("closed", DENOM)),
    ]
    return out


def run_checks(orc, gba, library=None):
    """Runs every scenario through the given oracle build; returns the list
    of (scenario, message) failures."""
    fails = []
    for name, prob, kw, extra in scenarios(gba, orc):
        opts = dict(max_iterations=30, rel_decrease_tol=1e-6, max_inner_trials=16)
        opts.update(kw)
        try:
            r = orc.lm_solve(prob.copy(), library=library, **opts)
            check_schedule(r, opts["initial_lambda"], opts["max_iterations"], opts["max_inner_trials"],
                           opts["rel_decrease_tol"])
            if isinstance(extra, tuple):
                closed_form(r, opts["initial_lambda"], extra[1])
                if extra[1] == NONFINITE:
                    # one skipped edge in the linearisation and in each trial cost
                    assert r["anomalies"] == 1 + 16
        except AssertionError as e:
            fails.append((name, str(e).splitlines()[0] if str(e) else "assertion"))
    return fails


def test_scenarios_take_their_paths(gba, orc):
    """Each scenario exercises what it is there for (so a mutant of that rule
    has something to break)."""
    sc = {name: (prob, kw, extra) for name, prob, kw, extra in scenarios(gba, orc)}
    r = orc.lm_solve(sc["rejections"][0].copy(), initial_lambda=1e-6, max_iterations=30)
    tr = r["trace"]
    runs = [it for it in np.unique(tr[:, 0]) if np.sum(tr[:, 0] == it) >= 3 and tr[tr[:, 0] == it][-1, 4] == ACCEPT]
    assert len(runs) >= 2
    acc = tr[tr[:, 4] == ACCEPT]
    assert np.any(acc[:, 3] < 0.937) and np.any(acc[:, 3] > 0.937)  # both branches of max(1/3, .)
    r = orc.lm_solve(sc["noise_free"][0].copy(), initial_lambda=1e-3, max_iterations=30)
    assert r["termination"] == "CONVERGENCE" and r["final_cost"] < 0.5
    z = sc["denominator"][0]
    assert orc.total_squared_error(z) == 0.0  # b = 0 exactly, so the step and denominator are 0


def test_oracle_follows_the_reference_schedule(gba, orc):
    assert run_checks(orc, gba) == []


def _build_variant(k, outdir):
    so = os.path.join(outdir, f"liboracle_m{k}.so")
    subprocess.run(["gcc", "-std=c11", "-O1", "-fPIC", "-shared", f"-DORC_MUTATE={k}", "-o", so,
                    os.path.join(ROOT, "oracle", "gba_oracle.c"), "-lpthread", "-lm"], check=True)
    return so


def test_every_planted_rule_change_is_caught(gba, orc, tmp_path):
    ks = list(range(1, 14))
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(lambda k: _build_variant(k, str(tmp_path)), ks))
    survivors = []
    for k, so in zip(ks, libs):
        fails = run_checks(orc, gba, library=orc.variant(so))
        if not fails:
            survivors.append(k)
    assert survivors == [], f"mutants not caught: {survivors}"
