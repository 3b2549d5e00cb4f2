"""The C++ drop-in path: a program written against include/gba (the
reference's API: gba::load_bal, gba::total_squared_error, gba::solve_bal_gpu)
is compiled with the system compiler, linked with libgba_amd.so alone (no
torch, no Python in the process) and run as a child process.  Its JSON must
agree with the Python binding on the same BAL file; its exit-3 gate fires on
non-finite observations, as the reference's driver does
(benchmarks/bench_bal.cpp:85-99)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "gpu-bundle-adjustment_amd")
SRC = os.path.join(ROOT, "tests", "cpp", "bal_cli.cpp")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_cli(out_dir):
    exe = os.path.join(out_dir, "bal_cli")
    cmd = ["g++", "-std=c++20", "-O2", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include", f"-I{ROCM}/include",
           SRC, "-o", exe, f"-L{PKG}", "-lgba_amd", f"-L{ROCM}/lib", "-lamdhip64",
           f"-Wl,-rpath,{PKG}:{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.fixture(scope="module")
def cli(tmp_path_factory):
    return build_cli(str(tmp_path_factory.mktemp("cli")))


def test_cli_builds_and_links_against_the_library(cli):
    """Compile and link only (CPU): every C++ symbol the program uses is
    defined by libgba_amd.so."""
    r = subprocess.run(["nm", "-D", "--undefined-only", cli], capture_output=True, text=True)
    assert "gba" in r.stdout  # the API comes from the library
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.gpu
def test_cli_solve_matches_python_binding(gba, cli, tmp_path):
    prob = gba.generate(1, 1)
    path = tmp_path / "cfg1.txt"
    gba.write_bal(path, prob)
    out = tmp_path / "out.json"
    r = subprocess.run([cli, str(path), "--lambda=1e-3", f"--json={out}"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    j = json.loads(r.stdout)
    assert j == json.loads(out.read_text())
    ref = gba.solve_bal_gpu(gba.load_bal(path), gba.LMOptions(initial_lambda=1e-3))
    assert (j["num_cameras"], j["num_points"], j["num_observations"]) == (16, 22106, 83718)
    assert j["termination"] == ref.termination == "CONVERGENCE"
    assert j["iterations"] == ref.iterations
    assert abs(j["initial_cost"] - ref.initial_cost) <= 1e-12 * ref.initial_cost
    assert abs(j["final_cost"] - ref.final_cost) <= 1e-12 * ref.final_cost
    assert j["solver"] == "gba-gpu-lm" and j["device"]


@pytest.mark.gpu
def test_cli_refuses_anomalies(gba, cli, tmp_path):
    prob = gba.generate(2, 1)
    prob.obs_pixel[5, 0] = np.nan
    path = tmp_path / "nan.txt"
    gba.write_bal(path, prob)
    r = subprocess.run([cli, str(path), "--iterations=2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, (r.returncode, r.stderr)
    assert "non-finite" in r.stderr
