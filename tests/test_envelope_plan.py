"""The envelope (profile) layout of the reduced camera system, on the CPU
(gba_plan_envelope, the plan the solver uses; camera_graph.cpp):

* the camera graph and its block density, recomputed here from the
  observations independently;
* the order is a permutation and the tile profile is the one the renumbered
  graph gives (first[i] = the tile column of the lowest camera any camera of
  tile row i is joined to);
* fill containment: the LAPACK Cholesky factor of an SPD matrix with the
  renumbered camera graph's block pattern is exactly zero left of the profile
  of every tile row -- the property the envelope factor rests on;
* the reverse Cuthill-McKee order recovers the locality of a randomly numbered
  camera set (no worse than the ring order), and a uniform-visibility problem is refused as
  too dense."""
import numpy as np
import pytest

TILE = 64


def camera_graph(p):
    C = p.num_cameras
    A = np.eye(C, dtype=bool)
    order = np.argsort(p.obs_point, kind="stable")
    oc, op = p.obs_camera[order], p.obs_point[order]
    cut = np.r_[0, np.flatnonzero(np.diff(op)) + 1, len(op)]
    for a, b in zip(cut[:-1], cut[1:]):
        cs = oc[a:b]
        A[np.ix_(cs, cs)] = True
    return A


def profile(A, order):
    """first tile column per tile row of the renumbered 9C x 9C matrix."""
    C = A.shape[0]
    pos = np.empty(C, int)
    pos[order] = np.arange(C)
    lo = np.array([min(pos[np.flatnonzero(A[order[i]])].min(), i) for i in range(C)])
    n = 9 * C
    nblk = (n + TILE - 1) // TILE
    first = np.arange(nblk)
    for t in range(nblk):
        rows = range(t * TILE, min(n, (t + 1) * TILE))
        if len(rows):
            first[t] = min(first[t], min((9 * lo[r // 9]) // TILE for r in rows))
    return first


def work(first):
    nblk = len(first)
    env = dense = 0
    for k in range(nblk - 1):
        m = nblk - 1 - k
        dense += m + m * (m + 1) // 2
        for i in range(k + 1, nblk):
            if first[i] <= k:
                env += 1 + sum(1 for j in range(k + 1, i + 1) if first[j] <= k)
    return env / dense


def small_window(gba, C=120, P=6000, O=30000, window=6, seed=1):
    s = gba.synth_spec(6, 1)
    s.num_cameras, s.num_points, s.num_observations, s.view_window = C, P, O, window
    return gba.generate(spec=s, seed=seed)


def test_plan_matches_independent_graph_and_profile(gba):
    for prob in (gba.generate(6, 1), small_window(gba)):
        A = camera_graph(prob)
        C = prob.num_cameras
        plan = gba.plan_envelope(prob, 1.0)
        assert plan is not None
        assert sorted(plan["order"]) == list(range(C))
        off = A.copy()
        np.fill_diagonal(off, False)
        assert plan["density"] == pytest.approx(off.sum() / C ** 2, abs=1e-12)
        first = profile(A, plan["order"])
        assert np.array_equal(plan["first"], first)
        assert np.all(plan["first"] <= np.arange(len(first)))
        assert plan["work_ratio"] == pytest.approx(work(first), rel=1e-12)
        assert plan["work_ratio"] < 0.5


def test_fill_stays_inside_the_profile(gba):
    prob = small_window(gba)
    A = camera_graph(prob)
    plan = gba.plan_envelope(prob, 1.0)
    order = plan["order"]
    C = prob.num_cameras
    n = 9 * C
    rng = np.random.default_rng(3)
    # SPD with the renumbered block pattern: random symmetric blocks where the
    # graph joins two cameras, then diagonally dominant
    B = A[np.ix_(order, order)]
    M = np.kron(B, np.ones((9, 9))) * rng.uniform(-1, 1, (n, n))
    M = np.tril(M) + np.tril(M, -1).T
    M += np.diag(np.abs(M).sum(axis=1) + 1.0)
    L = np.linalg.cholesky(M)
    for t, f in enumerate(plan["first"]):
        r0, r1 = t * TILE, min(n, (t + 1) * TILE)
        if r0 < n and f > 0:
            assert np.all(L[r0:r1, : f * TILE] == 0.0), t
    # and the profile is not vacuous: it leaves most of the triangle out
    assert plan["work_ratio"] < 0.5


def test_rcm_recovers_locality_of_a_shuffled_numbering(gba):
    prob = small_window(gba, C=160, window=5)
    C = prob.num_cameras
    ring = work(profile(camera_graph(prob), np.arange(C)))
    # the same problem with its cameras numbered at random
    shuffle = np.random.default_rng(7).permutation(C)
    q = prob.copy()
    q.cameras = prob.cameras[np.argsort(shuffle)]
    q.obs_camera = shuffle[prob.obs_camera].astype(np.int32)
    A = camera_graph(q)
    shuffled = work(profile(A, np.arange(C)))
    plan = gba.plan_envelope(q, 1.0)
    assert np.array_equal(plan["first"], profile(A, plan["order"]))
    assert shuffled > 0.8  # a random numbering fills the triangle
    assert plan["work_ratio"] <= ring * 1.05 and plan["work_ratio"] < 0.2, (plan["work_ratio"], ring)


def test_uniform_visibility_is_refused(gba):
    assert gba.plan_envelope(gba.generate(2, 1), 0.5) is None
    assert gba.plan_envelope(gba.generate(3, 1), 0.5) is None
    # forced: a complete graph plans, at no saving
    plan = gba.plan_envelope(gba.generate(2, 1), 1.0)
    assert plan is not None and plan["work_ratio"] == pytest.approx(1.0)
