"""FP16 observation storage with a per-camera FP32 scale.

CPU: the library's encoder (its own binary16 rounding, through the C ABI)
against the numpy restatement in oracle/obs_fp16.py bit for bit, and the
restatement against what IEEE binary16 fixes (exact round trips of
representable values, ties to even, the 2^-11 relative error bound, range).
GPU: a solve with FP16 storage is the FP64 solve of the decoded problem (the
device decodes each pixel exactly), and its final cost stays within a few
percent of the FP64-pixel solve (reported; the reference's offline FP16
study saw +0.2-1.5%, docs/RUNGS.md:44,50)."""
import numpy as np
import pytest

from oracle import obs_fp16


def encode_lib(gba, prob):
    scale, px16 = gba.encode_obs_fp16(prob)
    return scale, px16


def same_bits(a, b):
    a16, b16 = np.asarray(a, np.float16), np.asarray(b, np.float16)
    nan = np.isnan(a16)
    assert np.array_equal(nan, np.isnan(b16))
    return np.array_equal(a16[~nan].view(np.uint16), b16[~nan].view(np.uint16))


def test_encode_matches_oracle_on_generated_problem(gba):
    p = gba.generate(1, 3)
    scale, px16 = encode_lib(gba, p)
    s_o, px_o = obs_fp16.encode(p.obs_camera, p.obs_pixel, p.num_cameras)
    assert np.array_equal(scale.view(np.uint32), s_o.view(np.uint32))
    assert same_bits(px16, px_o)
    dec = obs_fp16.decode(p.obs_camera, scale, px16)
    err = np.abs(dec - p.obs_pixel) / np.maximum(np.abs(p.obs_pixel), 1e-300)
    assert np.max(err) <= 2.0 ** -11


def test_half_rounding_ties_range_and_specials(gba):
    # one camera whose largest finite |pixel| is 32768, so its scale is
    # exactly 1 and the stored value is binary16(pixel) itself
    rng = np.random.default_rng(5)
    half = np.arange(0, 0x7801, dtype=np.uint16).view(np.float16).astype(np.float64)  # 0 .. 2^15
    mids = (half[:-1] + half[1:]) / 2            # exact ties (binary16 steps are powers of 2)
    vals = np.concatenate([
        [32768.0, 0.0, -0.0, 32767.99, 2.0 ** -24, 2.0 ** -25,
         3 * 2.0 ** -26, 2.0 ** -26, 1e-300, np.inf, -np.inf, np.nan],
        mids, -mids[::7], half[::3],
        np.clip(rng.standard_normal(20000) * 10.0 ** rng.uniform(-9, 4.5, 20000), -32768, 32768),
    ])
    n = len(vals) + (len(vals) & 1)
    vals = np.resize(vals, n)
    vals[0] = 32768.0
    prob = gba.BALProblem.empty(1, 1, n // 2)
    prob.obs_pixel[:] = vals.reshape(-1, 2)
    prob.obs_camera[:] = 0
    prob.obs_point[:] = 0
    scale, px16 = encode_lib(gba, prob)
    assert scale[0] == 1.0
    got = px16.reshape(-1)
    assert same_bits(got, vals.astype(np.float16))
    # what the standard fixes, independently of numpy: representable values
    # are exact, ties go to the even neighbour, the error bounds hold
    g = got.astype(np.float64)
    fin = np.isfinite(vals)
    rep = np.isin(vals, half) | np.isin(-vals, half)
    assert np.array_equal(g[rep & fin], vals[rep & fin])
    tie = np.isin(np.abs(vals), mids)
    assert np.all(got[tie].view(np.uint16) % 2 == 0)
    assert np.array_equal(np.isinf(g), np.isinf(vals))
    nrm = fin & (np.abs(vals) >= 2.0 ** -14)
    assert np.all(np.abs(g[nrm] - vals[nrm]) <= 2.0 ** -11 * np.abs(vals[nrm]))
    sub = fin & (np.abs(vals) < 2.0 ** -14)
    assert np.all(np.abs(g[sub] - vals[sub]) <= 2.0 ** -25)


def test_scales_per_camera_edge_cases(gba):
    prob = gba.BALProblem.empty(4, 1, 5)
    prob.obs_camera[:] = [0, 0, 1, 2, 2]
    prob.obs_point[:] = 0
    prob.obs_pixel[:] = [[3.0, -700.25], [12.5, 0.5], [0.0, 0.0], [np.nan, 1e-3], [np.inf, -2e-3]]
    scale, px16 = encode_lib(gba, prob)
    s_o, px_o = obs_fp16.encode(prob.obs_camera, prob.obs_pixel, 4)
    assert np.array_equal(scale.view(np.uint32), s_o.view(np.uint32))
    assert same_bits(px16, px_o)
    assert scale[0] == np.float32(700.25 / 32768.0)
    assert scale[1] == 1.0 and scale[3] == 1.0          # all-zero camera, camera without observations
    assert scale[2] == np.float32(2e-3 / 32768.0)      # non-finite pixels do not set the scale
    assert np.isnan(px16[3, 0]) and np.isinf(px16[4, 0])
    # the largest magnitude lands near 2^15, inside binary16's range
    assert abs(float(px16[0, 1])) <= 32768.0 * (1 + 2.0 ** -10)


def decoded_problem(gba, p):
    scale, px16 = gba.encode_obs_fp16(p)
    q = p.copy()
    q.obs_pixel[:] = obs_fp16.decode(p.obs_camera, scale, px16)
    return q


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fp16_solve_is_solve_of_decoded_problem(gba, precision):
    prec = gba.PRECISION_FP64 if precision == "fp64" else gba.PRECISION_MIXED
    p = gba.generate(1, 4)
    q = decoded_problem(gba, p)
    r16 = gba.Solver(p, precision=prec, obs_storage=gba.OBS_FP16).solve()
    rdq = gba.Solver(q, precision=prec).solve()
    assert r16.trials == rdq.trials and r16.termination == rdq.termination
    assert abs(r16.initial_cost - rdq.initial_cost) <= 1e-12 * rdq.initial_cost
    assert abs(r16.final_cost - rdq.final_cost) <= 1e-9 * rdq.final_cost


@pytest.mark.gpu
def test_fp16_initial_cost_matches_oracle_on_decoded_pixels(gba, orc):
    p = gba.generate(2, 6)
    q = decoded_problem(gba, p)
    r16 = gba.Solver(p, obs_storage=gba.OBS_FP16).solve(gba.LMOptions(max_iterations=1))
    ref = 0.5 * orc.total_squared_error(q)  # cost = SSE / 2
    assert abs(r16.initial_cost - ref) <= 1e-12 * ref


@pytest.mark.gpu
def test_fp16_final_cost_error_vs_fp64(gba):
    p = gba.generate(1, 7)
    r64 = gba.Solver(p, precision=gba.PRECISION_MIXED).solve()
    r16 = gba.Solver(p, precision=gba.PRECISION_MIXED, obs_storage=gba.OBS_FP16).solve()
    err = (r16.final_cost - r64.final_cost) / r64.final_cost
    print(f"fp16 storage final-cost error vs FP64 pixels: {err:+.3e}")
    assert r16.termination == "CONVERGENCE"
    assert abs(err) <= 0.05
