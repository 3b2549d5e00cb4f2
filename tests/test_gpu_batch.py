"""Batched solver (many independent problems per launch) against each
problem's own sequential FP64 solve through the C ABI.  The batch solves each
factored S with a different kernel (one workgroup per problem), so the steps
agree to rounding; the tolerances are the LM convergence tolerance on the
final cost (1e-6 relative) and at most one iteration of difference.  Both
batched factors are covered: the tile-parallel one and the
one-workgroup-per-problem one (GBA_BATCH_FACTOR=tiles / workgroup; by
default the handle's autotuner picks one), and so are the rounds replayed
from captured HIP graphs (default) against eager launches
(GBA_BATCH_GRAPH=0)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small(gba, C, P, O, seed):
    s = gba.synth_spec(1)
    s.num_cameras, s.num_points, s.num_observations = C, P, O
    s.kmin, s.kmax = 2, min(C, 16)
    return gba.generate(spec=s, seed=seed)


def check_pair(rb, rs):
    assert rb.termination == rs.termination
    assert abs(rb.initial_cost - rs.initial_cost) <= 1e-12 * rs.initial_cost
    assert abs(rb.final_cost - rs.final_cost) <= 1e-6 * rs.final_cost
    assert abs(rb.iterations - rs.iterations) <= 1
    assert rb.anomalies == rs.anomalies


def test_batch_matches_sequential_config5(gba):
    probs = [gba.generate(5, seed) for seed in range(1, 7)]
    seq = [gba.Solver(p).solve() for p in probs]
    b = gba.Batch(probs)
    res = b.solve()
    assert b.rounds >= max(r.trials for r in res)
    for rb, rs in zip(res, seq):
        check_pair(rb, rs)
        assert rb.termination == "CONVERGENCE"
    # the same handle solves again from the uploaded initial parameters
    res2 = b.solve()
    for r1, r2 in zip(res, res2):
        assert abs(r1.final_cost - r2.final_cost) <= 1e-9 * r1.final_cost


def test_batch_512_preset5_against_sequential_and_oracle(gba, orc):
    """The full batched workload: 512 problems of preset 5 (seeds 1..512, C
    8..64, P 2,000..20,000, O = 4.5 P) in one handle.  Every problem against
    its own sequential FP64 solve; six of them, the extremes of C and P and
    two from the middle, against the oracle (the reference loop,
    src/lm.cu:458-702)."""
    probs = [gba.generate(5, seed) for seed in range(1, 513)]
    b = gba.Batch(probs)
    res = b.solve()
    worst = 0.0
    for p, rb in zip(probs, res):
        s = gba.Solver(p)
        rs = s.solve()
        s.close()
        assert rb.termination == rs.termination and rb.anomalies == rs.anomalies == 0
        assert abs(rb.iterations - rs.iterations) <= 1
        assert abs(rb.initial_cost - rs.initial_cost) <= 1e-12 * rs.initial_cost
        d = abs(rb.final_cost - rs.final_cost) / rs.final_cost
        worst = max(worst, d)
        assert d <= 1e-9
    C = np.array([p.num_cameras for p in probs])
    P = np.array([p.num_points for p in probs])
    pick = sorted({int(np.argmin(C)), int(np.argmax(C)), int(np.argmin(P)), int(np.argmax(P)), 100, 400})
    for i in pick:
        o = orc.lm_solve(probs[i].copy())
        assert res[i].termination == o["termination"]
        assert res[i].iterations == o["iterations"], (i, res[i].iterations, o["iterations"])
        assert abs(res[i].final_cost - o["final_cost"]) <= 1e-9 * o["final_cost"]
    print(f"512-problem batch: {b.rounds} rounds, worst final-cost gap to sequential {worst:.1e}")
    b.close()


@pytest.mark.parametrize("factor", ["tiles", "workgroup"])
def test_batch_mixed_sizes_and_params(gba, monkeypatch, factor):
    monkeypatch.setenv("GBA_BATCH_FACTOR", factor)
    probs = [small(gba, 8, 800, 3600, 11), small(gba, 64, 3000, 13500, 12), small(gba, 23, 1500, 6500, 13)]
    seq_solvers = [gba.Solver(p) for p in probs]
    seq = [s.solve() for s in seq_solvers]
    b = gba.Batch(probs)
    res = b.solve()
    for i, (rb, rs) in enumerate(zip(res, seq)):
        check_pair(rb, rs)
        cams_b, pts_b = b.params(i)
        cams_s, pts_s = seq_solvers[i].params()
        assert np.max(np.abs(cams_b - cams_s)) <= 1e-3 * max(1.0, np.max(np.abs(cams_s)))
        assert np.max(np.abs(pts_b - pts_s)) <= 1e-3 * max(1.0, np.max(np.abs(pts_s)))


def test_batch_max_iterations_and_anomalies(gba):
    p_nan = gba.generate(2, 1)
    p_nan.obs_pixel[10, 1] = np.nan
    probs = [p_nan, gba.generate(5, 3)]
    opts = gba.LMOptions(max_iterations=3)
    seq = [gba.solve_bal_gpu(p.copy(), opts) for p in probs]
    res = gba.Batch(probs).solve(opts)
    for rb, rs in zip(res, seq):
        check_pair(rb, rs)
    assert res[0].anomalies > 0
    assert res[1].termination == "MAX_ITERATIONS" and res[1].iterations == 3


def test_batch_of_one_and_one_shot(gba):
    p = gba.generate(5, 9)
    rs = gba.Solver(p).solve()
    q = p.copy()
    (rb,) = gba.solve_bal_batched([q])
    check_pair(rb, rs)
    # the one-shot form writes the accepted parameters back
    assert not np.allclose(q.cameras, p.cameras)


def test_batch_ragged_problem(gba):
    # a camera without observations and points seen once, beside a regular problem
    from test_gpu_path import ragged
    probs = [ragged(gba), gba.generate(5, 4)]
    seq = [gba.Solver(p).solve() for p in probs]
    res = gba.Batch(probs).solve()
    for rb, rs in zip(res, seq):
        check_pair(rb, rs)
        assert rb.termination == "CONVERGENCE"


def test_batch_graph_replay_matches_eager(gba, monkeypatch):
    """Rounds replayed from the captured HIP graphs against the same rounds
    launched one by one: every replay (all rounds of two solves of one
    handle, with the host's uploads, accept launches and synchronisations
    between replays) gives the eager results.  The FP64 Schur sums are LDS
    atomics whose order may differ between runs, hence 1e-9 rather than
    bit equality."""
    monkeypatch.setenv("GBA_BATCH_FACTOR", "tiles")
    probs = [gba.generate(5, seed) for seed in (21, 22, 23)]
    monkeypatch.setenv("GBA_BATCH_GRAPH", "0")
    eager = gba.Batch(probs)
    assert not eager.info()["graphs"]
    ref = eager.solve()
    monkeypatch.delenv("GBA_BATCH_GRAPH")
    g = gba.Batch(probs)
    assert g.info()["graphs"]
    for _ in range(2):
        res = g.solve()
        assert g.rounds == eager.rounds
        for i, (a, r) in enumerate(zip(res, ref)):
            assert a.termination == r.termination
            assert a.iterations == r.iterations and a.trials == r.trials
            assert abs(a.final_cost - r.final_cost) <= 1e-9 * r.final_cost
            ca, pa = g.params(i)
            cr, pr = eager.params(i)
            assert np.max(np.abs(ca - cr)) <= 1e-7 * max(1.0, np.max(np.abs(cr)))
            assert np.max(np.abs(pa - pr)) <= 1e-7 * max(1.0, np.max(np.abs(pr)))


def test_batch_autotuner_picks_the_faster_form(gba, monkeypatch):
    """The handle times one trial of the batch with each factor form and
    keeps the faster; a GBA_BATCH_FACTOR setting skips the timing."""
    monkeypatch.delenv("GBA_BATCH_FACTOR", raising=False)
    probs = [small(gba, 8, 800, 3600, 31), small(gba, 40, 2000, 9000, 32)]
    b = gba.Batch(probs)
    inf = b.info()
    t = inf["tune_ms"]
    assert t["tiles"] > 0.0 and t["workgroup"] > 0.0
    assert inf["factor"] == ("tiles" if t["tiles"] <= t["workgroup"] else "workgroup")
    seq = [gba.Solver(p).solve() for p in probs]
    for rb, rs in zip(b.solve(), seq):
        check_pair(rb, rs)
    monkeypatch.setenv("GBA_BATCH_FACTOR", "workgroup")
    inf2 = gba.Batch(probs).info()
    assert inf2["factor"] == "workgroup" and inf2["tune_ms"]["tiles"] == 0.0


def test_batch_any_order_repeated_observation_and_one_camera(gba):
    """Problems whose observations come in a random order with one repeated
    (same point, same camera), and a one-camera problem, batched with a
    regular one: each against its own sequential solve."""
    p = gba.generate(2, 1)
    rng = np.random.default_rng(5)
    order = rng.permutation(p.num_observations)
    order = np.concatenate([order, order[:1]])
    shuf = gba.BALProblem(p.cameras.copy(), p.points.copy(), p.obs_camera[order].copy(),
                          p.obs_point[order].copy(), p.obs_pixel[order].copy())
    keep = p.obs_camera < 1
    used = np.unique(p.obs_point[keep])
    remap = -np.ones(p.num_points, np.int32)
    remap[used] = np.arange(used.size, dtype=np.int32)
    one = gba.BALProblem(p.cameras[:1].copy(), p.points[used].copy(), p.obs_camera[keep].copy(),
                         remap[p.obs_point[keep]].copy(), p.obs_pixel[keep].copy())
    probs = [shuf, gba.generate(5, 4), one]
    opts = gba.LMOptions(max_iterations=8)
    seq = [gba.Solver(q).solve(opts) for q in probs]
    res = gba.Batch(probs).solve(opts)
    for rb, rs in zip(res[:2], seq[:2]):
        check_pair(rb, rs)
    # the one-camera problem fits exactly (every point seen once): its cost
    # falls to rounding level, compared on the scale of the initial cost
    rb, rs = res[2], seq[2]
    assert rb.termination == rs.termination
    assert abs(rb.initial_cost - rs.initial_cost) <= 1e-12 * rs.initial_cost
    assert abs(rb.final_cost - rs.final_cost) <= 1e-9 * rs.initial_cost
