"""gba_dense_spd_solve_noalloc (device buffers, static scratch, no
allocation) against gba_dense_spd_solve (host arrays, allocating) on the same
matrices: the same tile factor and substitutions run on the same numbers, so
the solutions agree to 1e-12 relative (identical in practice).  Two calls in a
row, the second smaller, so a stale first call in the shared scratch would
show; then the error paths: a matrix that is not positive definite (ok = 0,
b untouched) and n above the scratch limit (error return)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spd(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.uniform(-1, 1, (n, n))
    return M.T @ M / n + 0.5 * np.eye(n), rng.uniform(-1, 1, n)


def device_solve(gba, A, b):
    import torch
    At = torch.tensor(np.asfortranarray(A).T.copy(), device="cuda")  # column-major element order
    bt = torch.tensor(b, device="cuda")
    ok = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    gba.dense_spd_solve_noalloc(At, bt, ok)
    return bt.cpu().numpy(), int(ok.cpu().item())


def test_noalloc_matches_allocating_solve_twice(gba):
    for n, seed in ((900, 1), (300, 2)):
        A, b = spd(n, seed)
        x_ref, ok_ref, _ = gba.dense_spd_solve(A, b, gba.PRECISION_FP64)
        x, ok = device_solve(gba, A, b)
        assert ok_ref and ok == 1
        assert np.max(np.abs(x - x_ref)) <= 1e-12 * np.max(np.abs(x_ref))
        assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)


def test_noalloc_not_positive_definite_and_size_limit(gba):
    import torch
    A, b = spd(200, 3)
    A[50, 50] = -1.0
    x, ok = device_solve(gba, A, b)
    assert ok == 0 and np.array_equal(x, b)
    n = gba.NOALLOC_MAX_N + 1
    At = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    bt = torch.zeros(n, dtype=torch.float64, device="cuda")
    okt = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(gba.GBAError, match="static scratch"):
        gba.dense_spd_solve_noalloc(At, bt, okt)
    # the next call still works
    A, b = spd(64, 4)
    x, ok = device_solve(gba, A, b)
    assert ok == 1 and np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
