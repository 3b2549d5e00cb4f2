"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle (oracle/liboracle.so).

The oracle restates the reference LM solve on the CPU (see gba_oracle.c).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, as the checker or the timed CPU baseline; the product never does.
oracle/_ref/libref_bal.so (the reference's own reader/camera-model sources,
compiled by oracle/Makefile when the reference tree is present) is loaded by
ref() for golden-vector checks.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_bal.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_ullp = ctypes.POINTER(ctypes.c_ulonglong)

TRIAL_CODES = {0: "accept", 1: "reject", 2: "nonspd_v", 3: "chol_fail", 4: "denom", 5: "nonfinite"}
STEP_CODES = {0: "ok", 1: "nonspd_v", 2: "llt_fail", 3: "denom"}
TERMINATION = {0: "CONVERGENCE", 1: "MAX_ITERATIONS", 2: "FAILURE"}
# trial status codes of the LM trace (gba_oracle.h ORC_TR_*)
TRIAL_CODES = {0: "accept", 1: "reject", 2: "nonspd_v", 3: "llt_fail", 4: "denom", 5: "nonfinite"}
TRACE_STRIDE = 7


class Options(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int), ("initial_lambda", ctypes.c_double),
                ("rel_decrease_tol", ctypes.c_double), ("max_inner_trials", ctypes.c_int),
                ("verbose", ctypes.c_int), ("huber_delta", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("initial_cost", ctypes.c_double), ("final_cost", ctypes.c_double),
                ("iterations", ctypes.c_int), ("termination", ctypes.c_int),
                ("anomalies", ctypes.c_ulonglong), ("total_time_s", ctypes.c_double),
                ("trials", ctypes.c_int), ("t_linearize", ctypes.c_double),
                ("t_dampinv", ctypes.c_double), ("t_assembly", ctypes.c_double),
                ("t_solve", ctypes.c_double), ("t_trial_cost", ctypes.c_double)]


_lib = None
_ref = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _declare(ctypes.CDLL(LIB))
    return _lib


def variant(path):
    """A separately built copy of the oracle (e.g. one compiled with
    -DORC_MUTATE=k by the schedule tests), declared like lib()."""
    return _declare(ctypes.CDLL(os.path.abspath(path)))


def _declare(L):
    L.orc_snavely_residual.argtypes = [_dp, _dp, _dp, _dp]
    L.orc_edge_residual.argtypes = [_dp, _dp, _dp, _dp]
    L.orc_edge_residual_jacobian.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_total_squared_error.restype = ctypes.c_double
    L.orc_total_squared_error.argtypes = [ctypes.c_int, _dp, _dp, _ip, _ip, _dp]
    L.orc_linearize.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _ip, _ip, _dp] + [_dp] * 6 + [_ullp, ctypes.c_int]
    L.orc_cost.restype = ctypes.c_double
    L.orc_cost.argtypes = [ctypes.c_int, _dp, _dp, _ip, _ip, _dp, _ullp, ctypes.c_int]
    L.orc_linearize_loss.argtypes = ([ctypes.c_int] * 3 + [_dp, _dp, _ip, _ip, _dp, ctypes.c_double] +
                                     [_dp] * 6 + [_ullp, ctypes.c_int])
    L.orc_cost_loss.restype = ctypes.c_double
    L.orc_cost_loss.argtypes = [ctypes.c_int, _dp, _dp, _ip, _ip, _dp, ctypes.c_double, _ullp, ctypes.c_int]
    L.orc_chol9_factor.restype = ctypes.c_int
    L.orc_chol9_factor.argtypes = [_dp, _dp]
    L.orc_chol9_solve.argtypes = [_dp, _dp, _dp]
    L.orc_inverse3_spd.restype = ctypes.c_int
    L.orc_inverse3_spd.argtypes = [_dp, _dp]
    L.orc_batch_chol9_factor_solve.argtypes = [_dp, _dp, _dp, _ip, ctypes.c_int]
    L.orc_batch_inverse3_spd.argtypes = [_dp, _dp, _ip, ctypes.c_int]
    L.orc_damp_invert_v.restype = ctypes.c_long
    L.orc_damp_invert_v.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp]
    L.orc_dense_cholesky.restype = ctypes.c_int
    L.orc_dense_cholesky.argtypes = [ctypes.c_int, _dp]
    L.orc_dense_cholesky_solve.argtypes = [ctypes.c_int, _dp, _dp]
    L.orc_schur_step.restype = ctypes.c_int
    L.orc_schur_step.argtypes = [ctypes.c_int] * 3 + [_ip, _ip] + [_dp] * 5 + [ctypes.c_double] + [_dp] * 5 + [ctypes.c_int]
    L.orc_lm_solve.restype = ctypes.c_int
    L.orc_lm_solve.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _ip, _ip, _dp, ctypes.POINTER(Options),
                                                     ctypes.POINTER(Result), _dp, ctypes.c_int, _ip,
                                                     ctypes.c_int]
    return L


def ref():
    """oracle/_ref/libref_bal.so, or None when it was not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        R = ctypes.CDLL(REF_LIB)
        R.ref_snavely_residual.argtypes = [_dp, _dp, _dp, _dp]
        R.ref_total_squared_error.restype = ctypes.c_double
        R.ref_total_squared_error.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _ip, _ip, _dp]
        R.ref_load_bal_header.restype = ctypes.c_int
        R.ref_load_bal_header.argtypes = [ctypes.c_char_p, _ip, _ip, _ip, ctypes.c_char_p, ctypes.c_int]
        R.ref_load_bal.restype = ctypes.c_int
        R.ref_load_bal.argtypes = [ctypes.c_char_p, _dp, _dp, _ip, _ip, _dp]
        _ref = R
    return _ref


def _p(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _arrs(prob):
    cams = np.ascontiguousarray(prob.cameras, dtype=np.float64)
    pts = np.ascontiguousarray(prob.points, dtype=np.float64)
    oc = np.ascontiguousarray(prob.obs_camera, dtype=np.int32)
    op = np.ascontiguousarray(prob.obs_point, dtype=np.int32)
    px = np.ascontiguousarray(prob.obs_pixel, dtype=np.float64)
    return cams, pts, oc, op, px


def snavely_residual(cam, X, obs):
    r = np.zeros(2)
    cam, X, obs = (np.ascontiguousarray(a, dtype=np.float64) for a in (cam, X, obs))
    lib().orc_snavely_residual(_p(cam), _p(X), _p(obs), _p(r))
    return r


def edge_residual(cam, X, obs):
    r = np.zeros(2)
    cam, X, obs = (np.ascontiguousarray(a, dtype=np.float64) for a in (cam, X, obs))
    lib().orc_edge_residual(_p(cam), _p(X), _p(obs), _p(r))
    return r


def edge_residual_jacobian(cam, X, obs):
    r, Jc, Jp = np.zeros(2), np.zeros((2, 9)), np.zeros((2, 3))
    cam, X, obs = (np.ascontiguousarray(a, dtype=np.float64) for a in (cam, X, obs))
    lib().orc_edge_residual_jacobian(_p(cam), _p(X), _p(obs), _p(r), _p(Jc), _p(Jp))
    return r, Jc, Jp


def total_squared_error(prob):
    cams, pts, oc, op, px = _arrs(prob)
    return lib().orc_total_squared_error(len(oc), _p(cams), _p(pts), _p(oc, ctypes.c_int),
                                         _p(op, ctypes.c_int), _p(px))


def linearize(prob, nthreads=8, huber_delta=0.0):
    cams, pts, oc, op, px = _arrs(prob)
    C, P, O = cams.shape[0], pts.shape[0], len(oc)
    U, V, W = np.zeros((C, 9, 9)), np.zeros((P, 3, 3)), np.zeros((O, 9, 3))
    bc, bp = np.zeros((C, 9)), np.zeros((P, 3))
    sse, an = ctypes.c_double(), ctypes.c_ulonglong()
    lib().orc_linearize_loss(C, P, O, _p(cams), _p(pts), _p(oc, ctypes.c_int), _p(op, ctypes.c_int), _p(px),
                             huber_delta, _p(U), _p(V), _p(W), _p(bc), _p(bp), ctypes.byref(sse),
                             ctypes.byref(an), nthreads)
    return dict(U=U, V=V, W=W, bc=bc, bp=bp, sse=sse.value, anomalies=an.value)


def cost(prob, nthreads=8, huber_delta=0.0):
    cams, pts, oc, op, px = _arrs(prob)
    an = ctypes.c_ulonglong()
    s = lib().orc_cost_loss(len(oc), _p(cams), _p(pts), _p(oc, ctypes.c_int), _p(op, ctypes.c_int), _p(px),
                            huber_delta, ctypes.byref(an), nthreads)
    return s, an.value


def chol9_factor(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    L = np.zeros((9, 9))
    ok = lib().orc_chol9_factor(_p(A), _p(L))
    return L, bool(ok)


def batch_chol9_factor_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = A.shape[0]
    x, ok = np.zeros((n, 9)), np.zeros(n, np.int32)
    lib().orc_batch_chol9_factor_solve(_p(A), _p(b), _p(x), _p(ok, ctypes.c_int), n)
    return x, ok


def batch_inverse3_spd(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    Ai, ok = np.zeros((n, 3, 3)), np.zeros(n, np.int32)
    lib().orc_batch_inverse3_spd(_p(A), _p(Ai), _p(ok, ctypes.c_int), n)
    return Ai, ok


def damp_invert_v(V, lam):
    V = np.ascontiguousarray(V, dtype=np.float64)
    Vi = np.zeros_like(V)
    bad = lib().orc_damp_invert_v(V.shape[0], _p(V), lam, _p(Vi))
    return Vi, bad


def dense_cholesky_solve(A, b):
    """Column-major lower Cholesky + solve (the oracle's LLT stand-in)."""
    n = A.shape[0]
    L = np.array(A, dtype=np.float64, order="F", copy=True)
    x = np.array(b, dtype=np.float64, copy=True)
    ok = lib().orc_dense_cholesky(n, _p(L))
    if ok:
        lib().orc_dense_cholesky_solve(n, _p(L), _p(x))
    return x, bool(ok), np.tril(L)


def schur_step(prob, lin, lam, nthreads=8, want_S=False):
    _, _, oc, op, _ = _arrs(prob)
    C, P, O = prob.cameras.shape[0], prob.points.shape[0], len(oc)
    n = 9 * C
    dc, dp, den = np.zeros(n), np.zeros((P, 3)), ctypes.c_double()
    S = np.zeros((n, n), order="F") if want_S else None
    rhs = np.zeros(n) if want_S else None
    U, V, W, bc, bp = (np.ascontiguousarray(lin[k]) for k in ("U", "V", "W", "bc", "bp"))
    st = lib().orc_schur_step(C, P, O, _p(oc, ctypes.c_int), _p(op, ctypes.c_int), _p(U), _p(V), _p(W),
                              _p(bc), _p(bp), lam, _p(dc), _p(dp), ctypes.byref(den),
                              _p(S) if want_S else None, _p(rhs) if want_S else None, nthreads)
    return dict(dc=dc, dp=dp, denom=den.value, status=STEP_CODES[st], S=S, rhs=rhs)


def lm_solve(prob, max_iterations=50, initial_lambda=1e-3, rel_decrease_tol=1e-6,
             max_inner_trials=16, verbose=False, nthreads=8, max_trace=4096, huber_delta=0.0, library=None):
    """Runs the restated reference LM; updates prob.cameras / prob.points in
    place.  out["trace"] has one row per trial: outer iteration, lambda used,
    trial cost, rho, status code (TRIAL_CODES), nu, gain denominator.
    library: a variant() to run instead of the oracle proper."""
    cams, pts, oc, op, px = _arrs(prob)
    opts = Options(max_iterations, initial_lambda, rel_decrease_tol, max_inner_trials, int(verbose),
                   huber_delta)
    res = Result()
    trace = np.zeros((max_trace, TRACE_STRIDE))
    ntr = ctypes.c_int()
    (library or lib()).orc_lm_solve(cams.shape[0], pts.shape[0], len(oc), _p(cams), _p(pts), _p(oc, ctypes.c_int),
                       _p(op, ctypes.c_int), _p(px), ctypes.byref(opts), ctypes.byref(res), _p(trace),
                       max_trace, ctypes.byref(ntr), nthreads)
    prob.cameras[...] = cams.reshape(prob.cameras.shape)
    prob.points[...] = pts.reshape(prob.points.shape)
    out = {f: getattr(res, f) for f, _ in Result._fields_}
    out["termination"] = TERMINATION[res.termination]
    out["trace"] = trace[: ntr.value]
    return out
