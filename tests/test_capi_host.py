"""C ABI of libgba_amd.so on the CPU: every symbol declared in include/gba_c.h
is exported; host-side entry points (camera model, Jacobian, BAL I/O,
generators) agree with the reference reader's golden vectors and with the
oracle.  No GPU work is launched here."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, rel
from test_oracle_model import complex_step_jac

GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gba_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gba_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(gba):
    L = gba.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert "gfx950" in gba.version()


def test_binding_covers_every_declared_symbol(gba):
    assert set(declared_symbols()) <= set(gba._PROTOS)


def test_host_residual_matches_reference_reader(gba):
    g = np.load(os.path.join(GOLD, "residuals.npz"))
    for i in range(len(g["res"])):
        r = gba.snavely_residual(g["cams"][i], g["pts"][i], g["obs"][i])
        assert rel(r, g["res"][i]) < 1e-14


def test_host_total_squared_error_golden(gba):
    g = np.load(os.path.join(GOLD, "problem_sse.npz"))
    prob = gba.BALProblem(g["cams"], g["pts"], g["obs_camera"], g["obs_point"], g["obs_pixel"])
    assert gba.total_squared_error(prob) == float(g["sse"])


def test_kernel_jacobian_host_entry(gba, orc):
    """The kernels' own derivation (SO(3) left Jacobian) vs complex step and vs the oracle."""
    rng = np.random.default_rng(11)
    for i in range(100):
        th = [1.0, 1e-3, 1e-7, 1e-13, 3.0][i % 5]
        a = rng.normal(size=3)
        a *= th / np.linalg.norm(a)
        cam = np.concatenate([a, rng.normal(0, 1, 3) + [0, 0, -30],
                              [rng.uniform(300, 700), rng.normal(0, 0.05), rng.normal(0, 0.005)]])
        X, obs = rng.uniform(-5, 5, 3), rng.normal(0, 100, 2)
        r, Jc, Jp = gba.edge_residual_jacobian_host(cam, X, obs)
        r0, Jc0, Jp0 = orc.edge_residual_jacobian(cam, X, obs)
        assert rel(r, r0) < 1e-12
        assert rel(Jc, Jc0) < 1e-10 and rel(Jp, Jp0) < 1e-10
        if th > 1e-12:
            Jc_cs, Jp_cs = complex_step_jac(cam, X, obs)
            assert rel(Jc, Jc_cs) < 1e-10 and rel(Jp, Jp_cs) < 1e-10


def test_load_bal_matches_reference_loader(gba):
    g = np.load(os.path.join(GOLD, "bal_roundtrip.npz"))
    p = gba.load_bal(os.path.join(GOLD, "bal_roundtrip.txt"))
    assert np.array_equal(p.cameras.reshape(-1), g["cameras"])
    assert np.array_equal(p.points.reshape(-1), g["points"])
    assert np.array_equal(p.obs_camera, g["obs_camera"])
    assert np.array_equal(p.obs_point, g["obs_point"])
    assert np.array_equal(p.obs_pixel.reshape(-1), g["obs_pixel"])


def test_bal_write_load_roundtrip_and_bz2(gba, tmp_path):
    prob = gba.generate(2, 4)
    path = tmp_path / "p.txt"
    gba.write_bal(path, prob)
    q = gba.load_bal(path)
    for a, b in ((prob.cameras, q.cameras), (prob.points, q.points), (prob.obs_pixel, q.obs_pixel)):
        assert np.array_equal(a, b)
    assert np.array_equal(prob.obs_camera, q.obs_camera)
    os.system(f"bzip2 -k {path}")
    r = gba.load_bal(str(path) + ".bz2")
    assert np.array_equal(r.obs_pixel, prob.obs_pixel)


@pytest.mark.parametrize("text,msg", [
    ("0 1 1\n", "positive counts"), ("1 1 1\n0 5 1 2\n", "does not exist"),
    ("1 1 1\n0 0 1 x\n", "not a valid pixel"), ("1 1 1\n0 0 1 0x1p3\n", "not a valid pixel"),
    ("1 1 1\n0 +0 1 2\n", "not a valid point index"),
    ("1 1 1\n0 0 1 2\n" + "0\n" * 11, "ends before its point coordinate"),
    ("1 1 1\n0 0 1 2\n" + "0\n" * 12 + "7\n", "after the last point")])
def test_load_bal_rejects_malformed(gba, tmp_path, text, msg):
    path = tmp_path / "bad.txt"
    path.write_text(text)
    with pytest.raises(gba.GBAError, match=msg):
        gba.load_bal(path)


def test_load_bal_bz2_and_number_forms(gba, tmp_path):
    """A .bz2 file is read through bzcat; signed exponents, a leading '+',
    nan and inf parse as numbers."""
    import bz2
    import shutil
    text = "1 1 1\n0 0 +1.5e+2 -2.5E-1\n" + "0.5\n" * 8 + "nan\n" + "1\n-inf\n3.\n"
    plain = tmp_path / "p.txt"
    plain.write_text(text)
    a = gba.load_bal(plain)
    assert a.obs_pixel.tolist() == [[150.0, -0.25]]
    assert np.isnan(a.cameras[0, 8]) and a.points.tolist()[0][1:] == [-np.inf, 3.0]
    if shutil.which("bzcat"):
        packed = tmp_path / "p.txt.bz2"
        packed.write_bytes(bz2.compress(text.encode()))
        b = gba.load_bal(packed)
        assert np.array_equal(a.obs_pixel, b.obs_pixel) and np.array_equal(a.points, b.points)


def test_generator_shapes_and_determinism(gba):
    for cfg, (C, P, O) in {1: (16, 22106, 83718), 2: (49, 7776, 31843)}.items():
        a, b = gba.generate(cfg, 1), gba.generate(cfg, 1)
        assert (a.num_cameras, a.num_points, a.num_observations) == (C, P, O)
        assert np.array_equal(a.obs_pixel, b.obs_pixel) and np.array_equal(a.cameras, b.cameras)
        k = np.bincount(a.obs_point, minlength=P)
        assert k.min() >= 2 and k.max() <= C
        # distinct cameras per point, camera-major order
        pairs = a.obs_camera.astype(np.int64) * P + a.obs_point
        assert len(np.unique(pairs)) == O and np.all(np.diff(pairs) > 0)
    c = gba.generate(1, 2)
    assert not np.array_equal(c.obs_pixel, gba.generate(1, 1).obs_pixel)


def test_generator_batched_sizes(gba):
    for seed in range(20):
        s = gba.synth_spec(5, seed)
        assert 8 <= s.num_cameras <= 64 and 2000 <= s.num_points <= 20000
        assert s.num_observations == int(np.floor(4.5 * s.num_points + 0.5))


def test_launch_bounds_generator_matches_reference_recipe(gba):
    # tests/test_launch_bounds.cpp:65-80: observations are the truth's predictions
    prob = gba.generate_launch_bounds(8, 1000, 7)
    assert prob.num_observations == 2000
    assert list(prob.obs_camera[:4]) == [0, 1, 1, 2]
    assert np.all(prob.obs_point[::2] == np.arange(1000))


def test_entry_points_reject_null_arguments(gba):
    # checked before any device call: an error return and a message, no crash
    lib = gba.lib()
    calls = [
        ("gba_solve_bal", (None, None, None, None)),
        ("gba_solver_create", (None, None, None)),
        ("gba_solver_solve", (None, None, None)),
        ("gba_solver_params", (None, None, None)),
        ("gba_batch_create", (None, 0, None)),
        ("gba_batch_solve", (None, None, None, None)),
        ("gba_solve_bal_batched", (None, 0, None, None)),
        ("gba_linearize", (None, 0, None, None, None, None, None, None, None)),
        ("gba_cost", (None, None, None)),
        ("gba_dense_spd_solve", (4, 0, None, None, None, None)),
        ("gba_encode_obs_fp16", (None, None, None)),
    ]
    for name, args in calls:
        assert getattr(lib, name)(*args) != 0, name
        msg = lib.gba_last_error().decode()
        assert name.replace("_bal_batched", "") in msg or "null" in msg or "no problems" in msg, (name, msg)
