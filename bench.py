"""Headline benchmark: time-to-final-cost of the LM bundle-adjustment solve at
C=356 / P=226,730 / O=1,255,268 (synthetic Snavely problem, seed 1), mixed
precision vs plain FP64 on one MI355X.

One "step" is one complete LM solve with the observations already resident in
HBM (the solver handle is created, sorted and uploaded before the timed
region).  The plain FP64 solve defines the target cost C64; the time-to-final-
cost of a solve is the elapsed time from solve entry until the first accepted
LM iterate whose cost is <= C64 * (1 + 1e-6).  `value` is the median over the
timed mixed-precision steps (the library's defaults for the mixed form: its
inner solver on the FP32-stored system is the block-Jacobi CG at this size,
then one FP64 iterative-refinement correction per trial against the FP64
system's residual); the mixed form with its FP32 MFMA factor as the inner
solver (`mixed_factor`) and without the correction (`mixed_no_correction`, the
FP64 residual only evaluated) are timed beside it.  Prints one JSON line.  `including_setup`
gives the same measure for one-shot solves that allocate, sort and upload
inside the clock (the reference's total_time_s convention).

Also measured live: per-kernel average durations with HIP events on the
solver's own stream (roofline of the dominant kernel, plus the edge kernel),
and the CPU oracle (restated reference LM, reference threading rules) on a
bounded sample of the same problem on the host cores.
"""
import argparse
import importlib.util
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

PEAK = {"hbm_gbs": 8000.0, "fp64_tflops": 78.6, "fp32_tflops": 157.3}
REL_TOL = 1e-6


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "gba_amd", os.path.join(ROOT, "gpu-bundle-adjustment_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gba_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def time_to_cost(res, target):
    """Seconds from solve entry to the first accepted trial with cost <= target."""
    tr = res.trace
    for row in tr:
        if int(row[4]) == 0 and row[2] <= target:
            return float(row[5])
    return float("nan")


def run_batch(g, args):
    """Preset 5: --batch independent problems (per-problem seeds seed..seed+B-1,
    C ~ U{8..64}, P ~ U{2000..20000}, O = 4.5 P), solved by one batched handle
    per precision (inputs resident): the mixed form (FP32 W, diagonally scaled
    reduced system from fixed-point sums, FP64 factor; the headline value) and
    the plain FP64 form, problems/s over --steps solves each, with each
    problem's own solve time (device clock, start of the batch solve to the
    round that finished it) as p50 / p99; against problems/s of sequential
    FP64 solves (one solver handle per problem, inputs resident) on the first
    --seq-sample problems, whose final costs the batched results are compared
    with."""
    import torch
    probs = [g.generate(5, args.seed + i) for i in range(args.batch)]
    opts = g.LMOptions(initial_lambda=args.lambda0)
    forms = {}
    for name, prec in (("mixed", g.PRECISION_MIXED), ("fp64", g.PRECISION_FP64)):
        t_create = time.perf_counter()
        b = g.Batch(probs, precision=prec)
        t_create = time.perf_counter() - t_create
        for _ in range(args.warmup):
            b.solve(opts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = b.solve(opts)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        pt = sorted(r.solve_time_s for r in res)
        forms[name] = {"res": res, "wall": wall, "rounds": b.rounds, "handle": b.info(), "create_s": t_create,
                       "p50_s": statistics.median(pt), "p99_s": pt[min(len(pt) - 1, int(0.99 * len(pt)))]}
        b.close()
    k = min(args.seq_sample, args.batch)
    solvers = [g.Solver(p) for p in probs[:k]]
    for s in solvers[:2]:
        s.solve(opts)
    times, seq = [], []
    for s in solvers:
        t = time.perf_counter()
        seq.append(s.solve(opts))
        times.append(time.perf_counter() - t)
    seq_rate = k / sum(times)

    def rel(a, b):
        return float(max(abs(x.final_cost - y.final_cost) / y.final_cost for x, y in zip(a, b)))

    m, f = forms["mixed"], forms["fp64"]
    rate = args.batch / m["wall"]
    out = {
        "metric": "batched LM solves, preset 5 (problems/s)", "value": round(rate, 2), "unit": "problems/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * m["wall"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 W / f64",
        "data": "synthetic (seeded Snavely generator, per-problem seeds %d..%d)" % (args.seed, args.seed + args.batch - 1),
        "config": {"workload": f"{args.batch} problems of preset 5, mixed precision", "problems": args.batch,
                   "cameras": sum(p.num_cameras for p in probs), "points": sum(p.num_points for p in probs),
                   "observations": sum(p.num_observations for p in probs)},
        "per_problem_solve_s": {"p50": round(m["p50_s"], 5), "p99": round(m["p99_s"], 5)},
        "sequential_fp64": {"problems": k, "problems_per_s": round(seq_rate, 2),
                            "p50_s": round(statistics.median(times), 5),
                            "p99_s": round(sorted(times)[min(k - 1, int(0.99 * k))], 5)},
        "speedup_vs_sequential": round(rate / seq_rate, 2),
        "rounds": m["rounds"],
        "handle": m["handle"],
        "create_s": round(m["create_s"], 4),
        "max_final_cost_rel_diff_vs_fp64_batch": rel(m["res"], f["res"]),
        "max_final_cost_rel_diff_vs_sequential": rel(m["res"][:k], seq),
        "terminations": {t: sum(1 for r in m["res"] if r.termination == t)
                         for t in set(r.termination for r in m["res"])},
        "fp64_batch": {"problems_per_s": round(args.batch / f["wall"], 2), "ms_per_step": round(1e3 * f["wall"], 3),
                       "speedup_vs_sequential": round(args.batch / f["wall"] / seq_rate, 2),
                       "per_problem_solve_s": {"p50": round(f["p50_s"], 5), "p99": round(f["p99_s"], 5)},
                       "rounds": f["rounds"], "handle": f["handle"],
                       "max_final_cost_rel_diff_vs_sequential": rel(f["res"][:k], seq)},
    }
    print(json.dumps(out))


def run_fp16(g, args):
    """Preset 4 (C=1,936, P=525,000, O=3.4M): the mixed-precision solve with
    FP16 pixel storage (per-camera FP32 scale), timed over --steps solves with
    the observations resident, against the same solve with FP64 pixels and the
    plain FP64 solve, whose final cost is the reference for the reported
    final-cost error."""
    import torch
    prob = g.generate(4, args.seed)
    opts = g.LMOptions(initial_lambda=args.lambda0)
    runs = {}
    for name, prec, obs in (("mixed_fp16_px", g.PRECISION_MIXED, g.OBS_FP16),
                            ("mixed_fp64_px", g.PRECISION_MIXED, g.OBS_FP64),
                            ("fp64", g.PRECISION_FP64, g.OBS_FP64)):
        s = g.Solver(prob, precision=prec, obs_storage=obs)
        steps = args.steps if name == "mixed_fp16_px" else 1
        for _ in range(args.warmup if name == "mixed_fp16_px" else 0):
            s.solve(opts)
        torch.cuda.synchronize()
        times = []
        for _ in range(steps):
            r = s.solve(opts)
            times.append(r.solve_time_s)
        runs[name] = (r, statistics.median(times), len(times))
        print(f"[fp16] {name}: {r.termination} cost {r.final_cost:.9e} trials {r.trials} "
              f"solve {statistics.median(times):.3f} s", file=sys.stderr, flush=True)
        s.close()
        torch.cuda.synchronize()
    c64 = runs["fp64"][0].final_cost
    r16, t16, n16 = runs["mixed_fp16_px"]
    out = {
        "metric": "LM solve time, preset 4, mixed + FP16 pixel storage (s)", "value": round(t16, 4),
        "unit": "s", "n_gpus": 1, "steps": n16, "warmup": args.warmup, "ms_per_step": round(1e3 * t16, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64 (fp16 pixels)",
        "data": "synthetic (seeded Snavely generator, seed %d)" % args.seed,
        "config": {"workload": f"LM solve C={prob.num_cameras} P={prob.num_points} O={prob.num_observations} "
                               "(preset 4)"},
        "final_cost_rel_err_vs_fp64": {k: (v[0].final_cost - c64) / c64 for k, v in runs.items()},
        "solve_s": {k: round(v[1], 4) for k, v in runs.items()},
        "trials": {k: v[0].trials for k, v in runs.items()},
        "terminations": {k: v[0].termination for k, v in runs.items()},
    }
    print(json.dumps(out))


def run_sparse(g, args):
    """Sparse camera graph (preset --config, default 7: C=2,000, P=400,000,
    O=2M, each point seen within a ring window of 12 cameras): the plain FP64
    solve with the envelope factor layout against the dense layout on the same
    problem -- solve time (inputs resident), factor time per trial (HIP events
    on the solver stream), final costs."""
    import torch
    cfg = args.config if args.config in (6, 7) else 7
    prob = g.generate(cfg, args.seed)
    opts = g.LMOptions(initial_lambda=args.lambda0)
    out = {}
    for name, lay in (("envelope", g.LAYOUT_ENVELOPE), ("dense", g.LAYOUT_DENSE)):
        t = time.perf_counter()
        s = g.Solver(prob, g.PRECISION_FP64, layout=lay)
        setup = time.perf_counter() - t
        lay_id, ratio, dens = s.layout()
        for _ in range(args.warmup):
            s.solve(opts)
        torch.cuda.synchronize()
        steps = args.steps if name == "envelope" else max(1, min(args.steps, 3))
        rs = [s.solve(opts) for _ in range(steps)]
        fac = s.time_kernel("factor", args.lambda0, 3)
        sol = s.time_kernel("solve", args.lambda0, 3)
        r = rs[0]
        out[name] = {"setup_s": round(setup, 4), "solve_s": round(statistics.median(x.solve_time_s for x in rs), 5),
                     "samples": steps, "factor_ms": round(fac, 4), "triangular_solves_ms": round(sol, 4),
                     "final_cost": r.final_cost, "iterations": r.iterations, "trials": r.trials,
                     "termination": r.termination, "factor_work_vs_dense": round(ratio, 4),
                     "graph_density": round(dens, 4)}
        print(f"[sparse] {name}: solve {out[name]['solve_s']} s, factor {fac:.3f} ms", file=sys.stderr, flush=True)
        s.close()
        torch.cuda.synchronize()
    e, d = out["envelope"], out["dense"]
    res = {
        "metric": f"LM solve time, preset {cfg} (sparse camera graph), envelope vs dense factor (s)",
        "value": e["solve_s"], "unit": "s", "n_gpus": 1, "steps": e["samples"], "warmup": args.warmup,
        "ms_per_step": round(1e3 * e["solve_s"], 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Snavely generator, view window, seed %d)" % args.seed,
        "config": {"workload": f"LM solve C={prob.num_cameras} P={prob.num_points} O={prob.num_observations} "
                               f"(preset {cfg}, view window 12)"},
        "envelope": e, "dense": d,
        "speedup_solve": round(d["solve_s"] / e["solve_s"], 2),
        "speedup_factor": round(d["factor_ms"] / e["factor_ms"], 2),
        "final_cost_rel_diff": abs(e["final_cost"] - d["final_cost"]) / d["final_cost"],
    }
    print(json.dumps(res))


def run_bal(g, args):
    """The reference's bench_bal harness (benchmarks/bench_bal.cpp:63-139) on a
    BAL file: the reader-vs-solver initial-cost check (exit 2 on a mismatch)
    and the anomaly check (exit 3), then --pairs interleaved FP64 / mixed
    solves (the reference times its Ceres / GPU columns in interleaved pairs,
    docs/RUNGS.md:136-142), each printed as one JSON object in bench_bal's
    schema plus best / median / worst seconds.  Every solve starts from the
    file's parameters; total_time_s includes upload and sort, as the
    reference's does."""
    import torch
    base = g.load_bal(args.file)
    delta = args.delta if args.loss == "huber" else 0.0
    if args.loss == "huber":
        own = 0.5 * g.cost(base, huber_delta=delta)[0]
    else:
        own = g.total_squared_error(base) / 2.0
    device = torch.cuda.get_device_name(0)
    opts = g.LMOptions(max_iterations=args.iterations, initial_lambda=args.lambda0)
    precs = {"fp64": g.PRECISION_FP64, "mixed": g.PRECISION_MIXED}
    runs = {k: [] for k in precs}
    for _ in range(args.pairs):
        for name, prec in precs.items():
            p = base.copy()
            r = g.solve_bal_gpu(p, opts, prec, huber_delta=delta)
            if abs(r.initial_cost - own) > 1e-9 * max(1.0, abs(own)):
                print(f"MISMATCH: reader cost {own:.17g} vs solver initial cost {r.initial_cost:.17g}",
                      file=sys.stderr)
                sys.exit(2)
            if r.anomalies != 0:
                print(f"device anomaly counter is {r.anomalies} -- refusing to report", file=sys.stderr)
                sys.exit(3)
            runs[name].append(r)
    out = []
    for name, rs in runs.items():
        r = rs[-1]
        ts = sorted(x.total_time_s for x in rs)
        out.append({
            "file": args.file, "solver": f"gba-amd-lm-{name}", "loss": args.loss,
            "num_cameras": base.num_cameras, "num_points": base.num_points,
            "num_observations": base.num_observations,
            "initial_cost": r.initial_cost, "final_cost": r.final_cost, "iterations": r.iterations,
            "total_time_s": round(r.total_time_s, 6),
            "per_iteration_s": round(r.total_time_s / r.iterations, 6) if r.iterations else 0.0,
            "termination": r.termination, "device": device,
            "pairs": len(rs), "best_s": round(ts[0], 6), "median_s": round(statistics.median(ts), 6),
            "worst_s": round(ts[-1], 6),
        })
    text = "\n".join(json.dumps(o) for o in out)
    print(text)
    if args.json:
        with open(args.json, "w") as fh:
            fh.write(text + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--lambda0", type=float, default=None,
                    help="initial LM damping (default 1e-4; 1e-3 in --mode bal, as bench_bal)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-large", action="store_true", help="skip the preset-4 factor roofline")
    ap.add_argument("--cpu-iterations", type=int, default=2)
    ap.add_argument("--cpu-samples", type=int, default=5, help="interleaved CPU baseline samples")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--file", help="BAL problem for --mode bal")
    ap.add_argument("--iterations", type=int, default=50, help="--mode bal: max LM iterations")
    ap.add_argument("--json", help="--mode bal: also write the JSON objects here")
    ap.add_argument("--loss", choices=["none", "huber"], default="none", help="--mode bal")
    ap.add_argument("--delta", type=float, default=1.0, help="--mode bal: Huber threshold")
    ap.add_argument("--pairs", type=int, default=3, help="--mode bal: interleaved FP64 / mixed pairs")
    ap.add_argument("--mode", choices=["solve", "kernels", "batch", "fp16", "bal", "sparse"], default="solve",
                    help="kernels: only the per-kernel timing loops plus a 512 MiB device copy "
                         "(short run for rocprofv3 counter passes); batch: the batched solver on "
                         "--batch problems of preset 5 (problems/s), mixed and FP64 handles, against "
                         "sequential FP64 solves; "
                         "fp16: preset 4 with FP16 pixel storage, final-cost error vs FP64; "
                         "bal: the reference's bench_bal harness on --file (interleaved FP64 / mixed pairs); "
                         "sparse: preset 7 (or --config 6), envelope vs dense factor layout, plain FP64")
    ap.add_argument("--batch", type=int, default=512, help="problems in --mode batch")
    ap.add_argument("--seq-sample", type=int, default=32,
                    help="problems solved one by one for the sequential rate in --mode batch")
    args = ap.parse_args()
    if args.lambda0 is None:
        args.lambda0 = 1e-3 if args.mode == "bal" else 1e-4
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; a multi-GPU split is not built, "
                                   "run with --gpus 1"}))
        sys.exit(2)

    import torch
    torch.cuda.set_device(0)
    g = load_pkg()
    if args.mode == "batch":
        run_batch(g, args)
        return
    if args.mode == "fp16":
        run_fp16(g, args)
        return
    if args.mode == "sparse":
        run_sparse(g, args)
        return
    if args.mode == "bal":
        if not args.file:
            ap.error("--mode bal needs --file")
        run_bal(g, args)
        return
    prob = g.generate(args.config, args.seed)
    C, P, O = prob.num_cameras, prob.num_points, prob.num_observations
    opts = g.LMOptions(initial_lambda=args.lambda0)

    s64 = g.Solver(prob, g.PRECISION_FP64)
    s32 = g.Solver(prob, g.PRECISION_MIXED)  # the library's choice of refinement (block Jacobi here)
    s32f = g.Solver(prob, g.PRECISION_MIXED, precond=g.PRECOND_FACTOR)
    # the mixed form without its FP64 correction (the FP64 system's residual
    # evaluated and reported only): what the refinement costs
    s32n = g.Solver(prob, g.PRECISION_MIXED, refine_corrections=-1)
    jacobi = s32.precond() == g.PRECOND_JACOBI
    k64 = ("edge_point", "edge_camera", "schur", "factor", "solve", "backsub")
    k32 = (("edge_point", "edge_camera", "schur", "cg_setup", "cg_step", "backsub") if jacobi else k64) + \
        ("refine_residual",) + (("heavy_syrk",) if s32.heavy()[1] > 0 else ())
    if args.mode == "kernels":
        out = {"mode": "kernels", "config": {"cameras": C, "points": P, "observations": O}}
        for name, s, ks in (("fp64", s64, k64), ("mixed", s32, k32), ("mixed_factor", s32f, k64)):
            out[name] = {k: round(s.time_kernel(k, args.lambda0, args.kernel_reps), 4) for k in ks}
        # calibration for FETCH_SIZE / WRITE_SIZE: a plain copy of 512 MiB
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        b.copy_(a)
        ev1.record()
        torch.cuda.synchronize()
        out["copy_512MiB_ms"] = round(ev0.elapsed_time(ev1), 4)
        print(json.dumps(out))
        return
    for _ in range(args.warmup):
        s64.solve(opts)
        s32.solve(opts)
        s32f.solve(opts)
        s32n.solve(opts)
    torch.cuda.synchronize()

    # FP64 reference solves (also the target cost): at least five samples
    r64 = [s64.solve(opts) for _ in range(max(5, args.steps))]
    target = r64[0].final_cost * (1.0 + REL_TOL)
    t64 = [time_to_cost(r, target) for r in r64]

    # timed region: mixed-precision solves
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r32 = [s32.solve(opts) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t32 = [time_to_cost(r, target) for r in r32]
    value = statistics.median(t32)
    # the other refinement of the mixed form (FP32 MFMA factor), at least five samples
    r32f = [s32f.solve(opts) for _ in range(max(5, args.steps))]
    t32f = [time_to_cost(r, target) for r in r32f]
    r32n = [s32n.solve(opts) for _ in range(max(5, args.steps))]
    t32n = [time_to_cost(r, target) for r in r32n]

    # the same measure including allocation, sort and upload (one-shot
    # solve_bal_gpu calls, as the reference's total_time_s counts them)
    once = {}
    for name, prec in (("fp64", g.PRECISION_FP64), ("mixed", g.PRECISION_MIXED)):
        r = g.solve_bal_gpu(prob.copy(), opts, precision=prec)
        once[name] = {"time_to_final_cost_s": round(time_to_cost(r, target), 6),
                      "setup_s": round(r.setup_time_s, 6), "total_s": round(r.total_time_s, 6)}

    # per-kernel durations (HIP events on the solver stream)
    kern = {}
    for name, s, ks in (("fp64", s64, k64), ("mixed", s32, k32), ("mixed_factor", s32f, k64)):
        kern[name] = {k: s.time_kernel(k, args.lambda0, args.kernel_reps) for k in ks}
    st = s32.stats()
    n, npad = st["n"], st["n_padded"]
    trials32 = statistics.median([r.trials for r in r32])
    lins32 = statistics.median([r.iterations for r in r32])
    refine = statistics.median([r.refine_iterations for r in r32])
    outer = statistics.median([r.refine_outer_steps for r in r32])
    km = kern["mixed"]
    kf = kern["mixed_factor"]
    # kernel time per mixed solve (HIP-event averages x launches per solve)
    share = {"edge_point": lins32 * km["edge_point"], "edge_camera": lins32 * km["edge_camera"],
             "schur": trials32 * km["schur"], "backsub": trials32 * km["backsub"]}
    if jacobi:
        # a CG solve runs one step past its last counted one (the launch
        # that ends it); the first solve and the correction each
        share["cg_refinement"] = trials32 * km["cg_setup"] + (refine + (outer or trials32)) * km["cg_step"]
    else:
        share["factor"] = trials32 * km["factor"]
        share["solve"] = (trials32 + refine) * km["solve"]
    share["fp64_residual"] = outer * km["refine_residual"]
    dominant = max(share, key=share.get)
    w_bytes = 108  # FP32 W in the mixed form
    # edge accumulate = point pass + camera pass (one linearisation)
    edge_bytes = O * (16 + 8 + w_bytes) + P * (24 + 72) + 72 * C
    edge_ms = km["edge_point"] + km["edge_camera"]
    edge_gbs = edge_bytes / (edge_ms * 1e-3) / 1e9
    factor_flops = n ** 3 / 3.0
    factor_tflops = factor_flops / (kf["factor"] * 1e-3) / 1e12

    def traffic_of(name):
        """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
        WRITE_SIZE passes, scaled by the copy-kernel calibration."""
        tpath = os.path.join(ROOT, "profiles", name)
        if not (os.path.exists(tpath) and args.config == 3):
            return None
        with open(tpath) as fh:
            t = json.load(fh)
        return round(sum(k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"]
                         for k in t["kernels"].values()) * 1024)

    # The dense factor of S (FP32 MFMA), the other refinement of the mixed
    # form: n^3/3 flops per factorisation over the live HIP-event time of the
    # whole launch family of one factorisation (which also forms the explicit
    # inverse L^-1; its flops are not counted).
    roof_factor = {"kernel": "dense Cholesky of S (k_potrf_tile + k_trsm_tiles + k_chol_update family), "
                             "mixed form with the factor refinement",
                   "bound": "mfma", "achieved": round(factor_tflops, 3), "peak": PEAK["fp32_tflops"],
                   "unit": "TFLOP/s", "frac": round(factor_tflops / PEAK["fp32_tflops"], 4),
                   "traffic": traffic_of("factor_C356_traffic.json"),
                   "basis": f"n^3/3 = {factor_flops:.4e} flop (n = {n}) per factorisation over {kf['factor']:.4f} ms"}
    roof_cg = None
    if jacobi:
        # one block-Jacobi CG step (k_cg_matvec + k_cg_vector): every lower
        # 64x64 FP32 tile of the scaled S read once, plus the vectors (u, p, s,
        # x, r, b read, p, s, x, r, u written: 11 npad doubles) and the tile
        # partials (two 64-double vectors per tile, written and read)
        nblk = npad // 64
        tiles = nblk * (nblk + 1) // 2
        cg_bytes = tiles * 64 * 64 * 4 + 11 * npad * 8 + tiles * 2 * 64 * 8 * 2
        cg_gbs = cg_bytes / (km["cg_step"] * 1e-3) / 1e9
        roof_cg = {"kernel": "block-Jacobi CG step (k_cg_matvec + k_cg_vector), mixed form", "bound": "hbm",
                   "achieved": round(cg_gbs, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                   "frac": round(cg_gbs / PEAK["hbm_gbs"], 4), "traffic": traffic_of("k_cg_C356_traffic.json"),
                   "basis": f"{tiles} tiles x 16 KiB + 11 npad x 8 B + tile partials = {cg_bytes} B per step "
                            f"over {km['cg_step']:.4f} ms"}
    schur_bytes = O * w_bytes + 72 * P + (n * n // 2) * 4
    schur_gbs = schur_bytes / (km["schur"] * 1e-3) / 1e9
    roof_schur = {"kernel": "Schur assembly (k_schur_diag + k_schur + k_heavy_y + k_heavy_syrk), mixed form",
                  "bound": "hbm", "achieved": round(schur_gbs, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                  "frac": round(schur_gbs / PEAK["hbm_gbs"], 4), "traffic": traffic_of("k_schur_C356_traffic.json"),
                  "basis": f"O*108 + 72P + n^2/2*4 = {schur_bytes} B per assembly over {km['schur']:.4f} ms"}
    # the heavy points' dense product S -= Y Y^T inside the assembly (FP32
    # MFMA): 2 x 64 x 64 x width flops per live lower tile, over its own
    # HIP-event time
    roof_syrk = None
    hk, hn, hw = s32.heavy()
    if hn > 0 and "heavy_syrk" in km:
        nblk = npad // 64
        syrk_flops = 2.0 * (nblk * (nblk + 1) // 2) * 64 * 64 * hw
        syrk_tflops = syrk_flops / (km["heavy_syrk"] * 1e-3) / 1e12
        roof_syrk = {"kernel": "heavy points' dense product S -= Y Y^T (k_heavy_syrk), mixed form", "bound": "mfma",
                     "achieved": round(syrk_tflops, 3), "peak": PEAK["fp32_tflops"], "unit": "TFLOP/s",
                     "frac": round(syrk_tflops / PEAK["fp32_tflops"], 4), "traffic": None,
                     "basis": f"{hn} points seen >= {hk} times, width {hw}: 2 x 64^2 x {hw} flop x "
                              f"{nblk * (nblk + 1) // 2} lower tiles = {syrk_flops:.4e} flop per launch over "
                              f"{km['heavy_syrk']:.4f} ms"}
    # one residual of the FP64 reduced camera system (k_refine.hip: rotation
    # terms, point pass, camera chunks, rows): per observation 16 B of pixel
    # and 8 B of indices in each of the two passes, per point X, z (written
    # and read) and V'^-1 / bp, per camera its parameters, U, bc, x and the
    # residual
    res_bytes = O * (16 + 8) * 2 + P * (24 * 2 + 24 * 2 + 48 + 72) + C * (72 + 648 + 72 * 3)
    res_gbs = res_bytes / (km["refine_residual"] * 1e-3) / 1e9
    roof_res = {"kernel": "FP64 residual of the reduced camera system (k_refine.hip), mixed form's refinement",
                "bound": "hbm", "achieved": round(res_gbs, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                "frac": round(res_gbs / PEAK["hbm_gbs"], 4), "traffic": None,
                "basis": f"2 O (16 + 8) + P (48 + 48 + 120) + C (936) = {res_bytes} B per residual over "
                         f"{km['refine_residual']:.4f} ms; the Jacobians are evaluated again in FP64 and the camera "
                         f"pass gathers X and z at random, so the kernels run well below the HBM rate"}
    # the bench line's roofline is the kernel family with the largest share of
    # the measured mixed solve; the others follow as roofline_<name>
    entries = {"schur": roof_schur, "cg_refinement": roof_cg, "factor": roof_factor, "fp64_residual": roof_res,
               "heavy_syrk": roof_syrk}
    dom_key = dominant if dominant in entries and entries[dominant] else ("cg_refinement" if jacobi else "factor")
    roof = dict(entries[dom_key], largest_share=dominant)

    out = {
        "metric": "time-to-final-cost (rel err<=1e-6 vs FP64), C=356; mixed vs FP64 GPU",
        "value": round(value, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 W/S + f64 block-Jacobi CG on the f32-stored S, one f64 iterative-refinement "
                  "correction against the f64 system (matrix-free residual), f64 cost (mixed); "
                  "f32 MFMA factor-preconditioned leg, no-correction leg and f64 reference leg beside it")
                 if jacobi else
                 "f32 W/S/factor + f64 PCG on the f32-stored S, one f64 iterative-refinement correction against "
                 "the f64 system, f64 cost (mixed); f64 reference leg",
        "data": "synthetic (seeded Snavely generator, seed %d)" % args.seed,
        "config": {"workload": f"LM solve C={C} P={P} O={O} (preset {args.config}), "
                               f"lambda0={args.lambda0}, rel_decrease_tol=1e-6",
                   "cameras": C, "points": P, "observations": O},
        "fp64": {"time_to_final_cost_s": round(statistics.median(t64), 6),
                 "samples": len(t64), "best_s": round(min(t64), 6), "worst_s": round(max(t64), 6),
                 "solve_s": round(statistics.median([r.solve_time_s for r in r64]), 6),
                 "final_cost": r64[0].final_cost, "iterations": r64[0].iterations,
                 "trials": r64[0].trials, "termination": r64[0].termination},
        "mixed": {"time_to_final_cost_s": round(value, 6),
                  "samples": len(t32), "best_s": round(min(t32), 6), "worst_s": round(max(t32), 6),
                  "solve_s": round(statistics.median([r.solve_time_s for r in r32]), 6),
                  "final_cost": r32[0].final_cost, "iterations": r32[0].iterations,
                  "trials": r32[0].trials, "refine_iterations": r32[0].refine_iterations,
                  "refine_outer_steps": r32[0].refine_outer_steps,
                  "max_refine_residual_fp64_system": r32[0].max_refine_residual,
                  "refine_floor_trials": r32[0].refine_floor_trials,
                  "inner_solver": "block-Jacobi CG" if jacobi else "FP32 factor PCG",
                  "refine_factor_trials": r32[0].refine_factor_trials,
                  "termination": r32[0].termination,
                  "all_s": [round(t, 6) for t in t32]},
        "mixed_factor": {"time_to_final_cost_s": round(statistics.median(t32f), 6), "samples": len(t32f),
                         "best_s": round(min(t32f), 6), "worst_s": round(max(t32f), 6),
                         "final_cost": r32f[0].final_cost, "trials": r32f[0].trials,
                         "refine_iterations": r32f[0].refine_iterations,
                         "max_refine_residual_fp64_system": r32f[0].max_refine_residual},
        "mixed_no_correction": {"time_to_final_cost_s": round(statistics.median(t32n), 6), "samples": len(t32n),
                                "best_s": round(min(t32n), 6), "worst_s": round(max(t32n), 6),
                                "final_cost": r32n[0].final_cost, "trials": r32n[0].trials,
                                "refine_iterations": r32n[0].refine_iterations,
                                "max_refine_residual_fp64_system": r32n[0].max_refine_residual},
        "including_setup": once,
        "speedup_mixed_vs_fp64": round(statistics.median(t64) / value, 3) if value == value else None,
        "kernels_ms": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in kern.items()},
        "edge_kernel": {"passes": "point + camera", "bytes": edge_bytes, "ms": round(edge_ms, 4),
                        "gbs": round(edge_gbs, 1), "frac": round(edge_gbs / PEAK["hbm_gbs"], 4)},
        "kernel_share_ms": {k: round(v, 3) for k, v in share.items()},
        "factor": {"flops": factor_flops, "tflops_fp32": round(factor_tflops, 3),
                   "tflops_fp64": round(factor_flops / (kern["fp64"]["factor"] * 1e-3) / 1e12, 3)},
        "schur_pairs": st["pair_count"],
        "roofline": roof,
    }
    for k, v in entries.items():
        if v and k != dom_key:
            out["roofline_" + k] = v

    if not args.no_large and args.config == 3:
        # The dense factor where it is throughput-bound rather than
        # latency-bound: preset 4 (C=1,936, n = 17,424), the grouped form
        # (8 tile columns per group), HIP events around the factor launches
        # of one factorisation (S re-assembled before each repetition).
        pl = g.generate(4, args.seed)
        big = {}
        for name, prec in (("fp64", g.PRECISION_FP64), ("mixed", g.PRECISION_MIXED)):
            sl = g.Solver(pl, prec)
            big[name] = sl.time_kernel("factor", args.lambda0, 3)
            st4 = sl.stats()
            n4 = st4["n"]
            if prec == g.PRECISION_MIXED and sl.precond() == g.PRECOND_JACOBI:
                big["cg_step"] = sl.time_kernel("cg_step", args.lambda0, 10)
                nb4 = st4["n_padded"] // 64
                t4 = nb4 * (nb4 + 1) // 2
                big["cg_bytes"] = t4 * 64 * 64 * 4 + 11 * st4["n_padded"] * 8 + t4 * 2 * 64 * 8 * 2
            sl.close()
            torch.cuda.synchronize()
        fl4 = n4 ** 3 / 3.0
        tf64 = fl4 / (big["fp64"] * 1e-3) / 1e12
        tf32 = fl4 / (big["mixed"] * 1e-3) / 1e12
        out["roofline_factor_preset4"] = {
            "kernel": f"dense Cholesky of S at preset 4 (n = {n4}), grouped form", "bound": "mfma",
            "basis": f"n^3/3 = {fl4:.4e} flop per factorisation",
            "fp64": {"ms": round(big["fp64"], 3), "achieved": round(tf64, 2), "peak": PEAK["fp64_tflops"],
                     "unit": "TFLOP/s", "frac": round(tf64 / PEAK["fp64_tflops"], 4)},
            "mixed": {"ms": round(big["mixed"], 3), "achieved": round(tf32, 2), "peak": PEAK["fp32_tflops"],
                      "unit": "TFLOP/s", "frac": round(tf32 / PEAK["fp32_tflops"], 4),
                      "note": "the factor refinement's factor; at this size it preconditions with the factor's "
                              "triangular solves and forms no explicit inverse"},
        }
        if "cg_step" in big:
            gbs4 = big["cg_bytes"] / (big["cg_step"] * 1e-3) / 1e9
            out["roofline_cg_preset4"] = {
                "kernel": f"block-Jacobi CG step at preset 4 (n = {n4}), the default mixed refinement there",
                "bound": "hbm", "achieved": round(gbs4, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                "frac": round(gbs4 / PEAK["hbm_gbs"], 4), "ms": round(big["cg_step"], 4),
                "basis": f"{big['cg_bytes']} B per step (lower tiles of the FP32 S, vectors, tile partials)"}

    if not args.no_cpu:
        # CPU baseline: the oracle (the reference's LM restated, with its host
        # threading: min(C, hardware_concurrency) threads, src/lm.cu:382-383;
        # Phase-2 workers by the 16 MiB rule; single-threaded LLT) on a bounded
        # sample of the same problem -- --cpu-iterations LM iterations --
        # timed --cpu-samples times, interleaved with plain FP64 GPU solves as
        # the reference interleaves its columns (docs/RUNGS.md:136-142).
        sys.path.insert(0, ROOT)
        import oracle
        hw = os.cpu_count() or 1
        threads = min(C, hw)
        try:
            affinity = len(os.sched_getaffinity(0))
        except (AttributeError, OSError):
            affinity = hw
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        cpu_s, gpu_s, its = [], [], 0
        for _ in range(args.cpu_samples):
            p_cpu = g.generate(args.config, args.seed)
            t = time.perf_counter()
            rc = oracle.lm_solve(p_cpu, initial_lambda=args.lambda0, max_iterations=args.cpu_iterations,
                                 nthreads=threads)
            cpu_s.append(time.perf_counter() - t)
            its = rc["iterations"]
            gpu_s.append(time_to_cost(s64.solve(opts), target))
            print(f"[cpu] sample {len(cpu_s)}: {cpu_s[-1]:.2f} s for {its} iterations", file=sys.stderr, flush=True)
        per_it = [t / max(1, its) for t in cpu_s]
        scale = r64[0].iterations
        out["cpu_baseline"] = {
            "value": round(statistics.median(per_it) * scale, 3), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"oracle LM (restated reference: min(C, hardware_concurrency) = {threads} threads, "
                      f"Phase-2 S workers by the 16 MiB rule, single-threaded blocked LLT) for {its} "
                      f"iterations at C={C}, {len(cpu_s)} samples interleaved with FP64 GPU solves; "
                      f"value = median s/iteration x {scale} FP64 GPU iterations",
            "hardware_concurrency": hw, "affinity_cpus": affinity, "cpu_model": model,
            "samples_s": [round(t, 3) for t in cpu_s],
            "best_s": round(min(per_it) * scale, 3), "worst_s": round(max(per_it) * scale, 3),
            "interleaved_gpu_fp64_time_to_final_cost_s": [round(t, 6) for t in gpu_s]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
