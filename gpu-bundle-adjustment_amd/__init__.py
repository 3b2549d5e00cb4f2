"""Python binding of libgba_amd.so (MI355X bundle-adjustment LM solver).

Mirrors the reference's C++ API (gba::BALProblem, gba::LMOptions,
gba::LMResult, gba::solve_bal_gpu, ...) over the C ABI in include/gba_c.h,
with numpy arrays for the problem's structure of arrays.  The library must be
built (gpu-bundle-adjustment_amd/build.py); loading fails loudly otherwise.
"""
import ctypes
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgba_amd.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

CONVERGENCE, MAX_ITERATIONS, FAILURE = 0, 1, 2
TERMINATION = {CONVERGENCE: "CONVERGENCE", MAX_ITERATIONS: "MAX_ITERATIONS", FAILURE: "FAILURE"}
PRECISION_FP64, PRECISION_MIXED = 0, 1
OBS_FP64, OBS_FP16 = 0, 1  # device storage of the observed pixels
NOALLOC_MAX_N = 1024  # GBA_NOALLOC_MAX_N of include/gba_c.h
LAYOUT_AUTO, LAYOUT_DENSE, LAYOUT_ENVELOPE = 0, 1, 2  # enum gba_factor_layout
PRECOND_AUTO, PRECOND_FACTOR, PRECOND_JACOBI = 0, 1, 2  # enum gba_mixed_precond
TRACE_STRIDE = 6
TRIAL_STATUS = {0: "accept", 1: "reject", 2: "nonspd_v", 3: "chol_fail", 4: "denom", 5: "nonfinite"}
KERNELS = {"edge_point": 0, "edge_camera": 1, "schur": 2, "factor": 3, "solve": 4, "backsub": 5,
           "cost": 6, "cg_step": 7, "cg_setup": 8, "refine_residual": 9, "heavy_syrk": 10}


class ProblemView(ctypes.Structure):
    _fields_ = [("num_cameras", ctypes.c_int), ("num_points", ctypes.c_int),
                ("num_observations", ctypes.c_int), ("cameras", _dp), ("points", _dp),
                ("obs_camera", _ip), ("obs_point", _ip), ("obs_pixel", _dp)]


class LMOptionsC(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int), ("initial_lambda", ctypes.c_double),
                ("rel_decrease_tol", ctypes.c_double), ("max_inner_trials", ctypes.c_int),
                ("verbose", ctypes.c_int)]


class LMResultC(ctypes.Structure):
    _fields_ = [("initial_cost", ctypes.c_double), ("final_cost", ctypes.c_double),
                ("iterations", ctypes.c_int), ("termination", ctypes.c_int),
                ("anomalies", ctypes.c_ulonglong), ("total_time_s", ctypes.c_double),
                ("trials", ctypes.c_int), ("solve_time_s", ctypes.c_double),
                ("setup_time_s", ctypes.c_double), ("refine_iterations", ctypes.c_int),
                ("max_refine_residual", ctypes.c_double), ("refine_factor_trials", ctypes.c_int),
                ("refine_outer_steps", ctypes.c_int), ("refine_floor_trials", ctypes.c_int),
                ("max_refine_floor", ctypes.c_double)]


class SolverConfigC(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("refine_tol", ctypes.c_double),
                ("refine_max_iters", ctypes.c_int), ("trace", _dp), ("trace_capacity", ctypes.c_int),
                ("obs_storage", ctypes.c_int), ("huber_delta", ctypes.c_double), ("factor_layout", ctypes.c_int),
                ("mixed_precond", ctypes.c_int), ("jacobi_max_iters", ctypes.c_int),
                ("refine_corrections", ctypes.c_int), ("cg_deflation", ctypes.c_int)]


class SynthSpecC(ctypes.Structure):
    _fields_ = [("num_cameras", ctypes.c_int), ("num_points", ctypes.c_int),
                ("num_observations", ctypes.c_int), ("kmin", ctypes.c_int), ("kmax", ctypes.c_int),
                ("law", ctypes.c_int), ("alpha", ctypes.c_double), ("ring_radius", ctypes.c_double),
                ("pixel_noise", ctypes.c_double), ("perturb", ctypes.c_double), ("view_window", ctypes.c_int)]


# exported symbols and their prototypes (include/gba_c.h)
_PROTOS = {
    "gba_default_options": (None, [ctypes.POINTER(LMOptionsC)]),
    "gba_default_config": (None, [ctypes.POINTER(SolverConfigC)]),
    "gba_solve_bal": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.POINTER(LMOptionsC),
                                     ctypes.POINTER(SolverConfigC), ctypes.POINTER(LMResultC)]),
    "gba_solver_create": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.POINTER(SolverConfigC),
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "gba_solver_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LMOptionsC),
                                        ctypes.POINTER(LMResultC)]),
    "gba_solver_params": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "gba_solver_time_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_int, _dp]),
    "gba_solver_stats": (ctypes.c_int, [ctypes.c_void_p, _dp, _ip, _ip]),
    "gba_solver_layout": (ctypes.c_int, [ctypes.c_void_p, _ip, _dp, _dp]),
    "gba_solver_heavy": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _ip]),
    "gba_solver_mixed_precond": (ctypes.c_int, [ctypes.c_void_p, _ip]),
    "gba_solver_destroy": (None, [ctypes.c_void_p]),
    "gba_batch_create": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "gba_batch_create_precision": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.c_int, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_void_p)]),
    "gba_batch_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LMOptionsC),
                                       ctypes.POINTER(LMResultC), _ip]),
    "gba_batch_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp]),
    "gba_batch_info": (ctypes.c_int, [ctypes.c_void_p, _ip, _dp, _ip]),
    "gba_batch_destroy": (None, [ctypes.c_void_p]),
    "gba_solve_bal_batched": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.c_int,
                                             ctypes.POINTER(LMOptionsC), ctypes.POINTER(LMResultC)]),
    "gba_linearize": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.c_int, _dp, _dp, _dp, _dp,
                                     _dp, _dp, ctypes.POINTER(ctypes.c_ulonglong)]),
    "gba_schur_step": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.c_int, ctypes.c_double, _dp,
                                      _dp, _dp, _ip, _dp, _dp]),
    "gba_encode_obs_fp16": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.c_void_p, ctypes.c_void_p]),
    "gba_linearize_cfg": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.POINTER(SolverConfigC), _dp, _dp,
                                         _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_ulonglong)]),
    "gba_cost_cfg": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.POINTER(SolverConfigC), _dp,
                                    ctypes.POINTER(ctypes.c_ulonglong)]),
    "gba_cost": (ctypes.c_int, [ctypes.POINTER(ProblemView), _dp, ctypes.POINTER(ctypes.c_ulonglong)]),
    "gba_dense_spd_solve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _ip]),
    "gba_dense_spd_solve_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_void_p, ctypes.c_void_p]),
    "gba_batch_chol9_factor_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "gba_batch_inverse3_spd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int, ctypes.c_void_p]),
    "gba_snavely_residual": (None, [_dp, _dp, _dp, _dp]),
    "gba_total_squared_error": (ctypes.c_double, [ctypes.POINTER(ProblemView)]),
    "gba_edge_residual_jacobian_host": (None, [_dp, _dp, _dp, _dp, _dp, _dp]),
    "gba_bal_header": (ctypes.c_int, [ctypes.c_char_p, _ip, _ip, _ip]),
    "gba_load_bal": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ProblemView)]),
    "gba_write_bal": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ProblemView)]),
    "gba_plan_envelope": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.c_double, _ip, _ip, _ip, _dp, _dp]),
    "gba_synth_preset": (ctypes.c_int, [ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(SynthSpecC)]),
    "gba_synth_generate": (ctypes.c_int, [ctypes.POINTER(SynthSpecC), ctypes.c_ulonglong,
                                          ctypes.POINTER(ProblemView)]),
    "gba_synth_launch_bounds": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong,
                                               ctypes.POINTER(ProblemView)]),
    "gba_device_synchronize": (ctypes.c_int, []),
    "gba_last_error": (ctypes.c_char_p, []),
    "gba_version": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """The loaded libgba_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with gpu-bundle-adjustment_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class GBAError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise GBAError(lib().gba_last_error().decode())


def _ptr(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


@dataclasses.dataclass
class BALProblem:
    """Structure of arrays, as gba::BALProblem (reference bal_problem.hpp:15-31)."""
    cameras: np.ndarray     # (C, 9) float64
    points: np.ndarray      # (P, 3) float64
    obs_camera: np.ndarray  # (O,) int32
    obs_point: np.ndarray   # (O,) int32
    obs_pixel: np.ndarray   # (O, 2) float64

    @property
    def num_cameras(self):
        return self.cameras.shape[0]

    @property
    def num_points(self):
        return self.points.shape[0]

    @property
    def num_observations(self):
        return self.obs_camera.shape[0]

    def copy(self):
        return BALProblem(*(np.array(a, copy=True) for a in dataclasses.astuple(self)))

    def normalize(self):
        self.cameras = np.ascontiguousarray(self.cameras, dtype=np.float64).reshape(-1, 9)
        self.points = np.ascontiguousarray(self.points, dtype=np.float64).reshape(-1, 3)
        self.obs_camera = np.ascontiguousarray(self.obs_camera, dtype=np.int32).reshape(-1)
        self.obs_point = np.ascontiguousarray(self.obs_point, dtype=np.int32).reshape(-1)
        self.obs_pixel = np.ascontiguousarray(self.obs_pixel, dtype=np.float64).reshape(-1, 2)
        return self

    def view(self):
        self.normalize()
        return ProblemView(self.num_cameras, self.num_points, self.num_observations,
                           _ptr(self.cameras), _ptr(self.points), _ptr(self.obs_camera, ctypes.c_int),
                           _ptr(self.obs_point, ctypes.c_int), _ptr(self.obs_pixel))

    @staticmethod
    def empty(C, P, O):
        return BALProblem(np.zeros((C, 9)), np.zeros((P, 3)), np.zeros(O, np.int32),
                          np.zeros(O, np.int32), np.zeros((O, 2)))


@dataclasses.dataclass
class LMOptions:
    """gba::LMOptions (reference lm.hpp:13-19)."""
    max_iterations: int = 50
    initial_lambda: float = 1e-3
    rel_decrease_tol: float = 1e-6
    max_inner_trials: int = 16
    verbose: bool = False

    def c(self):
        return LMOptionsC(self.max_iterations, self.initial_lambda, self.rel_decrease_tol,
                          self.max_inner_trials, int(self.verbose))


@dataclasses.dataclass
class LMResult:
    """gba::LMResult (reference lm.hpp:21-29) plus timing / refinement fields."""
    initial_cost: float
    final_cost: float
    iterations: int
    termination: str
    anomalies: int
    total_time_s: float
    trials: int = 0
    solve_time_s: float = 0.0
    setup_time_s: float = 0.0
    refine_iterations: int = 0
    max_refine_residual: float = 0.0
    trace: np.ndarray = None
    refine_factor_trials: int = 0
    refine_outer_steps: int = 0
    refine_floor_trials: int = 0
    max_refine_floor: float = 0.0

    @staticmethod
    def from_c(r, trace=None):
        return LMResult(r.initial_cost, r.final_cost, r.iterations, TERMINATION[r.termination],
                        r.anomalies, r.total_time_s, r.trials, r.solve_time_s, r.setup_time_s,
                        r.refine_iterations, r.max_refine_residual, trace, r.refine_factor_trials,
                        r.refine_outer_steps, r.refine_floor_trials, r.max_refine_floor)


def _config(precision, trace_capacity=0, refine_tol=1e-10, refine_max_iters=12, obs_storage=OBS_FP64,
            huber_delta=0.0, layout=None, precond=None, jacobi_max_iters=None, refine_corrections=None,
            deflation=None):
    c = SolverConfigC()
    lib().gba_default_config(ctypes.byref(c))
    c.precision = precision
    if deflation is not None:
        c.cg_deflation = 0 if deflation else -1
    if refine_corrections is not None:
        c.refine_corrections = refine_corrections
    if precond is not None:
        c.mixed_precond = precond
    if jacobi_max_iters is not None:
        c.jacobi_max_iters = jacobi_max_iters
    if layout is not None:
        c.factor_layout = layout
    c.obs_storage = obs_storage
    c.huber_delta = huber_delta
    c.refine_tol = refine_tol
    c.refine_max_iters = refine_max_iters
    trace = None
    if trace_capacity:
        trace = np.full((trace_capacity, TRACE_STRIDE), np.nan)
        c.trace = _ptr(trace)
        c.trace_capacity = trace_capacity
    return c, trace


def _trim_trace(trace):
    if trace is None:
        return None
    keep = ~np.isnan(trace[:, 4])
    return trace[keep]


def solve_bal_gpu(prob, opts=None, precision=PRECISION_FP64, trace_capacity=1024, obs_storage=OBS_FP64,
                  huber_delta=0.0, layout=None, precond=None):
    """gba::solve_bal_gpu: updates prob.cameras / prob.points in place.
    huber_delta > 0 minimises sum rho(|r|^2) / 2 with the Huber loss; layout
    is a LAYOUT_* value, precond a PRECOND_* value for the mixed refinement
    (default: the library's automatic choices)."""
    opts = opts or LMOptions()
    cfg, trace = _config(precision, trace_capacity, obs_storage=obs_storage, huber_delta=huber_delta,
                         layout=layout, precond=precond)
    res = LMResultC()
    v = prob.view()
    _check(lib().gba_solve_bal(ctypes.byref(v), ctypes.byref(opts.c()), ctypes.byref(cfg),
                               ctypes.byref(res)))
    return LMResult.from_c(res, _trim_trace(trace))


class Solver:
    """Device-resident solver handle: upload once, solve repeatedly."""

    def __init__(self, prob, precision=PRECISION_FP64, trace_capacity=1024, obs_storage=OBS_FP64,
                 huber_delta=0.0, layout=None, precond=None, jacobi_max_iters=None, refine_corrections=None,
                 deflation=None):
        self._cfg, self._trace = _config(precision, trace_capacity, obs_storage=obs_storage,
                                         huber_delta=huber_delta, layout=layout, precond=precond,
                                         jacobi_max_iters=jacobi_max_iters, refine_corrections=refine_corrections,
                                         deflation=deflation)
        self._h = ctypes.c_void_p()
        self._prob = prob
        v = prob.view()
        _check(lib().gba_solver_create(ctypes.byref(v), ctypes.byref(self._cfg), ctypes.byref(self._h)))

    def solve(self, opts=None):
        opts = opts or LMOptions()
        if self._trace is not None:
            self._trace[:] = np.nan
        res = LMResultC()
        _check(lib().gba_solver_solve(self._h, ctypes.byref(opts.c()), ctypes.byref(res)))
        return LMResult.from_c(res, _trim_trace(None if self._trace is None else self._trace.copy()))

    def params(self):
        cams = np.zeros((self._prob.num_cameras, 9))
        pts = np.zeros((self._prob.num_points, 3))
        _check(lib().gba_solver_params(self._h, _ptr(cams), _ptr(pts)))
        return cams, pts

    def time_kernel(self, kernel, lam=1e-4, reps=10):
        ms = ctypes.c_double()
        _check(lib().gba_solver_time_kernel(self._h, KERNELS[kernel], lam, reps, ctypes.byref(ms)))
        return ms.value

    def stats(self):
        pc, n, npad = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        _check(lib().gba_solver_stats(self._h, ctypes.byref(pc), ctypes.byref(n), ctypes.byref(npad)))
        return {"pair_count": pc.value, "n": n.value, "n_padded": npad.value}

    def heavy(self):
        """(heavy_k, heavy points, inner width of the dense product) of the Schur assembly."""
        k, n, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().gba_solver_heavy(self._h, ctypes.byref(k), ctypes.byref(n), ctypes.byref(w)))
        return k.value, n.value, w.value

    def layout(self):
        """(LAYOUT_DENSE or LAYOUT_ENVELOPE, envelope work / dense work, camera-graph density)."""
        lay, ratio, dens = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _check(lib().gba_solver_layout(self._h, ctypes.byref(lay), ctypes.byref(ratio), ctypes.byref(dens)))
        return lay.value, ratio.value, dens.value

    def precond(self):
        """The mixed refinement's preconditioner (PRECOND_FACTOR or PRECOND_JACOBI; 0 for FP64)."""
        m = ctypes.c_int()
        _check(lib().gba_solver_mixed_precond(self._h, ctypes.byref(m)))
        return m.value

    def close(self):
        if self._h:
            lib().gba_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched solver handle (many independent problems, one launch per LM
    stage for the whole batch): upload once, solve repeatedly.  precision
    PRECISION_FP64 (default) or PRECISION_MIXED (FP32 W, diagonally scaled
    reduced system from fixed-point sums, FP64 factor)."""

    def __init__(self, probs, precision=PRECISION_FP64):
        self._probs = list(probs)
        n = len(self._probs)
        self._views = (ProblemView * n)(*[p.view() for p in self._probs])
        self._h = ctypes.c_void_p()
        self.precision = precision
        _check(lib().gba_batch_create_precision(self._views, n, precision, ctypes.byref(self._h)))
        self.rounds = 0

    def info(self):
        """The factor form the autotuner chose ("tiles": tile-parallel across
        the batch, "workgroup": one workgroup per problem), its trial times in
        ms for the two forms, and whether rounds replay captured HIP graphs."""
        tf, gr = ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_double * 2)()
        _check(lib().gba_batch_info(self._h, ctypes.byref(tf), ms, ctypes.byref(gr)))
        return {"factor": "tiles" if tf.value else "workgroup",
                "tune_ms": {"tiles": ms[0], "workgroup": ms[1]}, "graphs": bool(gr.value)}

    def solve(self, opts=None):
        opts = opts or LMOptions()
        n = len(self._probs)
        res = (LMResultC * n)()
        rounds = ctypes.c_int()
        _check(lib().gba_batch_solve(self._h, ctypes.byref(opts.c()), res, ctypes.byref(rounds)))
        self.rounds = rounds.value
        return [LMResult.from_c(r) for r in res]

    def params(self, i):
        p = self._probs[i]
        cams = np.zeros((p.num_cameras, 9))
        pts = np.zeros((p.num_points, 3))
        _check(lib().gba_batch_params(self._h, i, _ptr(cams), _ptr(pts)))
        return cams, pts

    def close(self):
        if self._h:
            lib().gba_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_bal_batched(probs, opts=None):
    """Batched one-shot solve; updates each problem's cameras / points in place."""
    opts = opts or LMOptions()
    probs = list(probs)
    n = len(probs)
    views = (ProblemView * n)(*[p.view() for p in probs])
    res = (LMResultC * n)()
    _check(lib().gba_solve_bal_batched(views, n, ctypes.byref(opts.c()), res))
    return [LMResult.from_c(r) for r in res]


def linearize(prob, precision=PRECISION_FP64, huber_delta=0.0):
    """GPU linearisation in the reference layout: U (C,9,9 lower), V (P,3,3 lower),
    W (O,9,3), bc (C,9), bp (P,3), sse, anomalies.  huber_delta > 0: IRLS-weighted
    by the Huber loss, sse = sum rho(|r|^2)."""
    C, P, O = prob.num_cameras, prob.num_points, prob.num_observations
    U, V, W = np.zeros((C, 9, 9)), np.zeros((P, 3, 3)), np.zeros((O, 9, 3))
    bc, bp = np.zeros((C, 9)), np.zeros((P, 3))
    sse, an = ctypes.c_double(), ctypes.c_ulonglong()
    v = prob.view()
    if huber_delta:
        cfg, _ = _config(precision, huber_delta=huber_delta)
        _check(lib().gba_linearize_cfg(ctypes.byref(v), ctypes.byref(cfg), _ptr(U), _ptr(V), _ptr(W), _ptr(bc),
                                       _ptr(bp), ctypes.byref(sse), ctypes.byref(an)))
    else:
        _check(lib().gba_linearize(ctypes.byref(v), precision, _ptr(U), _ptr(V), _ptr(W), _ptr(bc), _ptr(bp),
                                   ctypes.byref(sse), ctypes.byref(an)))
    return dict(U=U, V=V, W=W, bc=bc, bp=bp, sse=sse.value, anomalies=an.value)


def schur_step(prob, lam, precision=PRECISION_FP64, want_S=False):
    C, P = prob.num_cameras, prob.num_points
    n = 9 * C
    dc, dp = np.zeros(n), np.zeros((P, 3))
    den, st = ctypes.c_double(), ctypes.c_int()
    S = np.zeros((n, n), order="F") if want_S else None
    rhs = np.zeros(n) if want_S else None
    v = prob.view()
    _check(lib().gba_schur_step(ctypes.byref(v), precision, lam, _ptr(dc), _ptr(dp), ctypes.byref(den),
                                ctypes.byref(st), _ptr(S) if want_S else None,
                                _ptr(rhs) if want_S else None))
    return dict(dc=dc, dp=dp, denom=den.value, status=TRIAL_STATUS[st.value], S=S, rhs=rhs)


def encode_obs_fp16(prob):
    """FP16 pixel encoding used by obs_storage=OBS_FP16: (scale float32[C],
    px16 float16[O, 2]); the decoded pixels are px16 * scale[camera]."""
    scale = np.zeros(prob.num_cameras, dtype=np.float32)
    px16 = np.zeros((prob.num_observations, 2), dtype=np.float16)
    v = prob.view()
    _check(lib().gba_encode_obs_fp16(ctypes.byref(v), scale.ctypes.data, px16.ctypes.data))
    return scale, px16


def cost(prob, huber_delta=0.0):
    """(sse, anomalies) at the problem's parameters; with huber_delta > 0 the
    first value is sum rho(|r|^2)."""
    sse, an = ctypes.c_double(), ctypes.c_ulonglong()
    v = prob.view()
    if huber_delta:
        cfg, _ = _config(PRECISION_FP64, huber_delta=huber_delta)
        _check(lib().gba_cost_cfg(ctypes.byref(v), ctypes.byref(cfg), ctypes.byref(sse), ctypes.byref(an)))
    else:
        _check(lib().gba_cost(ctypes.byref(v), ctypes.byref(sse), ctypes.byref(an)))
    return sse.value, an.value


def dense_spd_solve(A, b, precision=PRECISION_FP64):
    A = np.asfortranarray(A, dtype=np.float64)
    x = np.array(b, dtype=np.float64, copy=True)
    ok, it = ctypes.c_int(), ctypes.c_int()
    _check(lib().gba_dense_spd_solve(A.shape[0], precision, _ptr(A), _ptr(x), ctypes.byref(ok),
                                     ctypes.byref(it)))
    return x, bool(ok.value), it.value


def dense_spd_solve_noalloc(A, b, ok):
    """gba_dense_spd_solve_noalloc on torch device tensors: A (n, n) float64
    read as column-major (pass A.T.contiguous() for a row-major matrix; for a
    symmetric A either works), b (n,) float64 overwritten with the solution,
    ok (1,) int32.  Runs on the library's default stream between a torch
    synchronisation and a library synchronisation (see below)."""
    import torch
    torch.cuda.synchronize()
    _check(lib().gba_dense_spd_solve_noalloc(A.shape[0], A.data_ptr(), b.data_ptr(), ok.data_ptr(), None))
    _check(lib().gba_device_synchronize())


# torch ships its own HIP runtime; device pointers are shared between the two
# runtimes of the process, stream handles are not.  The batched launchers
# therefore run on this library's default stream, bracketed by a torch
# synchronisation before and a library synchronisation after.

def batch_chol9_factor_solve(A, b):
    """A (n,9,9), b (n,9) torch device float64 tensors -> (x, ok)."""
    import torch
    n = A.shape[0]
    x = torch.zeros_like(b)
    ok = torch.zeros(n, dtype=torch.int32, device=A.device)
    torch.cuda.synchronize()
    _check(lib().gba_batch_chol9_factor_solve(A.data_ptr(), b.data_ptr(), x.data_ptr(), ok.data_ptr(),
                                              n, None))
    _check(lib().gba_device_synchronize())
    return x, ok


def batch_inverse3_spd(A):
    import torch
    n = A.shape[0]
    Ainv = torch.zeros_like(A)
    ok = torch.zeros(n, dtype=torch.int32, device=A.device)
    torch.cuda.synchronize()
    _check(lib().gba_batch_inverse3_spd(A.data_ptr(), Ainv.data_ptr(), ok.data_ptr(), n, None))
    _check(lib().gba_device_synchronize())
    return Ainv, ok


def snavely_residual(cam, X, obs):
    out = np.zeros(2)
    cam, X, obs = (np.ascontiguousarray(a, dtype=np.float64) for a in (cam, X, obs))
    lib().gba_snavely_residual(_ptr(cam), _ptr(X), _ptr(obs), _ptr(out))
    return out


def total_squared_error(prob):
    v = prob.view()
    return lib().gba_total_squared_error(ctypes.byref(v))


def edge_residual_jacobian_host(cam, X, obs):
    r, Jc, Jp = np.zeros(2), np.zeros((2, 9)), np.zeros((2, 3))
    cam, X, obs = (np.ascontiguousarray(a, dtype=np.float64) for a in (cam, X, obs))
    lib().gba_edge_residual_jacobian_host(_ptr(cam), _ptr(X), _ptr(obs), _ptr(r), _ptr(Jc), _ptr(Jp))
    return r, Jc, Jp


def load_bal(path):
    C, P, O = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().gba_bal_header(str(path).encode(), ctypes.byref(C), ctypes.byref(P), ctypes.byref(O)))
    prob = BALProblem.empty(C.value, P.value, O.value)
    v = prob.view()
    _check(lib().gba_load_bal(str(path).encode(), ctypes.byref(v)))
    return prob


def write_bal(path, prob):
    v = prob.view()
    _check(lib().gba_write_bal(str(path).encode(), ctypes.byref(v)))


def plan_envelope(prob, max_density=1.0):
    """Host-side envelope plan: None when the camera graph is denser than
    max_density, else dict(order, first, work_ratio, density)."""
    C = prob.num_cameras
    nblk = (9 * C + 63) // 64
    order, first = np.zeros(C, np.int32), np.zeros(nblk, np.int32)
    ok, wr, dens = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    v = prob.view()
    _check(lib().gba_plan_envelope(ctypes.byref(v), max_density, ctypes.byref(ok), _ptr(order, ctypes.c_int),
                                   _ptr(first, ctypes.c_int), ctypes.byref(wr), ctypes.byref(dens)))
    if not ok.value:
        return None
    return dict(order=order, first=first, work_ratio=wr.value, density=dens.value)


def synth_spec(config, seed=1):
    s = SynthSpecC()
    _check(lib().gba_synth_preset(config, seed, ctypes.byref(s)))
    return s


def generate(config=1, seed=1, spec=None):
    """Seeded synthetic problem of a preset shape (1..5) or of a given spec."""
    s = spec if spec is not None else synth_spec(config, seed)
    prob = BALProblem.empty(s.num_cameras, s.num_points, s.num_observations)
    v = prob.view()
    _check(lib().gba_synth_generate(ctypes.byref(s), seed, ctypes.byref(v)))
    return prob


def generate_launch_bounds(C=8, P=400000, seed=7):
    prob = BALProblem.empty(C, P, 2 * P)
    v = prob.view()
    _check(lib().gba_synth_launch_bounds(C, P, seed, ctypes.byref(v)))
    return prob


def version():
    return lib().gba_version().decode()
