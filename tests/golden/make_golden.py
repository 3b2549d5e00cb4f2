"""Writes golden vectors from the reference's own host sources.

Runs in the build container only: oracle/_ref/libref_bal.so is compiled by
oracle/Makefile from the reference's src/bal_reader.cpp (snavely_residual,
total_squared_error, load_bal).  Inputs are seeded numpy draws; the outputs are
stored as data (npz, no pickles) next to this script:
  residuals.npz  : cams (N,9), pts (N,3), obs (N,2) -> res (N,2); rows 0..15
                   have |rodrigues| <= 1e-12 (the reader's zero-order branch)
  problem_sse.npz: a small problem (C,P,O arrays) -> reference SSE
  bal_roundtrip.txt + bal_roundtrip.npz: a BAL text file and the arrays the
                   reference loader returns for it
Usage: python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def p(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref_bal.so missing: run make -C oracle ref")
    rng = np.random.default_rng(20261203)
    N = 256
    cams = np.empty((N, 9))
    cams[:, 0:3] = rng.normal(0, 0.8, (N, 3))
    cams[:16, 0:3] = rng.normal(0, 1e-14, (16, 3))  # zero-order branch
    cams[:, 3:6] = rng.normal(0, 1, (N, 3))
    cams[:, 5] -= 30.0
    cams[:, 6] = rng.uniform(300, 700, N)
    cams[:, 7] = rng.normal(0, 0.05, N)
    cams[:, 8] = rng.normal(0, 0.005, N)
    pts = rng.uniform(-5, 5, (N, 3))
    obs = rng.normal(0, 200, (N, 2))
    res = np.zeros((N, 2))
    for i in range(N):
        R.ref_snavely_residual(p(cams[i]), p(pts[i]), p(obs[i]), p(res[i]))
    np.savez(os.path.join(HERE, "residuals.npz"), cams=cams, pts=pts, obs=obs, res=res)

    C, P, O = 5, 40, 150
    pc = cams[16:16 + C].copy()
    pp = pts[:P].copy()
    oc = rng.integers(0, C, O).astype(np.int32)
    op = rng.integers(0, P, O).astype(np.int32)
    px = rng.normal(0, 150, (O, 2))
    sse = R.ref_total_squared_error(C, P, O, p(pc), p(pp), p(oc, ctypes.c_int), p(op, ctypes.c_int), p(px))
    np.savez(os.path.join(HERE, "problem_sse.npz"), cams=pc, pts=pp, obs_camera=oc, obs_point=op,
             obs_pixel=px, sse=np.array(sse))

    # BAL text with mixed formatting (exponents, tabs, CRLF) read by the reference loader
    path = os.path.join(HERE, "bal_roundtrip.txt")
    lines = [f"{C} {P} {O}"]
    for i in range(O):
        lines.append(f"{oc[i]}\t{op[i]}   {px[i,0]:.10e} {px[i,1]:.17g}")
    for v in pc.reshape(-1):
        lines.append(f"{v:.17g}")
    for v in pp.reshape(-1):
        lines.append(f"{v:.12e}")
    with open(path, "w", newline="") as f:
        f.write("\r\n".join(lines) + "\n")
    c_, p_, o_ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    err = ctypes.create_string_buffer(256)
    assert R.ref_load_bal_header(path.encode(), ctypes.byref(c_), ctypes.byref(p_), ctypes.byref(o_), err, 256) == 0
    lc, lp = np.zeros(9 * c_.value), np.zeros(3 * p_.value)
    loc, lop = np.zeros(o_.value, np.int32), np.zeros(o_.value, np.int32)
    lpx = np.zeros(2 * o_.value)
    assert R.ref_load_bal(path.encode(), p(lc), p(lp), p(loc, ctypes.c_int), p(lop, ctypes.c_int), p(lpx)) == 0
    np.savez(os.path.join(HERE, "bal_roundtrip.npz"), cameras=lc, points=lp, obs_camera=loc,
             obs_point=lop, obs_pixel=lpx)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
