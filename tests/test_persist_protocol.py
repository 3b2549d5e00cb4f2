"""CPU check of the one-launch tile Cholesky's protocol (k_chol_persistent):
tests/cpp/persist_emulator.cpp runs the task order and the dependency rules of
gpu-bundle-adjustment_amd/csrc/persist_tasks.hpp -- the header the kernel
uses -- on host threads, with the kernel's arithmetic on small tiles.

* No deadlock for any worker count, one worker included: with a single
  worker every task must be runnable once the tasks before it in ticket order
  are done, so a rule that waited on a later ticket would hang here (each
  wait has a deadline; a missed deadline is exit status 3).
* The factor is the Cholesky factor of the input (LAPACK) and X = L^-1, so
  every read happened after the write it depends on.
* A ThreadSanitizer build finds no data race: the published flags order
  every tile write before every read of it."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "persist_emulator.cpp")
INC = os.path.join(ROOT, "gpu-bundle-adjustment_amd", "csrc")


def build(out, tsan=False):
    cmd = ["g++", "-std=c++17", "-O1" if tsan else "-O2", "-pthread", f"-I{INC}", SRC, "-o", out]
    if tsan:
        cmd[3:3] = ["-fsanitize=thread", "-g"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    return build(str(tmp_path_factory.mktemp("emu") / "persist_emulator"))


def run(exe, tmp, nblk, tb, workers, seed, inverse, timeout=120):
    out = os.path.join(tmp, f"o_{nblk}_{tb}_{workers}_{seed}_{inverse}.bin")
    r = subprocess.run([exe, str(nblk), str(tb), str(workers), str(seed), str(int(inverse)), out],
                       capture_output=True, text=True, timeout=timeout)
    return r, out


def check(out, n, inverse):
    raw = np.fromfile(out, dtype=np.float64).reshape(3, n, n)
    A0, L, X = raw
    A = np.tril(A0) + np.tril(A0, -1).T
    Lref = np.linalg.cholesky(A)
    assert np.max(np.abs(L - Lref)) <= 1e-10 * np.max(np.abs(Lref))
    if inverse:
        Xref = np.linalg.inv(Lref)
        assert np.max(np.abs(np.tril(X) - Xref)) <= 1e-9 * np.max(np.abs(Xref))


@pytest.mark.parametrize("nblk,tb,workers,inverse", [
    (1, 4, 1, True), (2, 3, 1, True), (6, 4, 1, True), (7, 3, 1, False), (9, 4, 2, True),
    (12, 2, 3, False), (12, 3, 8, True), (16, 2, 31, True), (20, 2, 64, False)])
def test_protocol_completes_and_factors(emu, tmp_path, nblk, tb, workers, inverse):
    for seed in (1, 2, 3):
        r, out = run(emu, str(tmp_path), nblk, tb, workers, seed, inverse)
        assert r.returncode == 0, (r.returncode, r.stderr)
        check(out, nblk * tb, inverse)


def test_protocol_task_counts(emu, tmp_path):
    # factor tasks (TRSM + trailing updates, the critical chain excluded) and
    # inverse roles: sum over columns of (m - 1) + (m (m + 1) / 2 - 1) and
    # (k + 1) + (nblk - 1 - k) k
    nblk = 10
    r, _ = run(emu, str(tmp_path), nblk, 2, 4, 1, True)
    want = 0
    for k in range(nblk):
        m = nblk - 1 - k
        if m >= 1:
            want += (m - 1) + (m * (m + 1) // 2 - 1)
        want += (k + 1) + (nblk - 1 - k) * k
    assert r.stdout.strip() == f"tasks {want}"


def test_protocol_has_no_data_race(tmp_path):
    exe = build(str(tmp_path / "persist_emulator_tsan"), tsan=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    for nblk, workers in ((11, 7), (20, 64)):
        out = str(tmp_path / f"t{nblk}.bin")
        r = subprocess.run([exe, str(nblk), "2", str(workers), "5", "1", out], capture_output=True, text=True,
                           timeout=300, env=env)
        assert r.returncode == 0 and "ThreadSanitizer" not in r.stderr, r.stderr[-3000:]
        check(out, nblk * 2, True)
