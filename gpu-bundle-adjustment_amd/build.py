"""Builds libgba_amd.so (HIP kernels for gfx950 + host driver + C ABI) in-tree.

Every source is compiled with hipcc for gfx950 only; objects go to
gpu-bundle-adjustment_amd/_build/, the shared library next to this file.
Packed FP32 instructions are disabled in device code, FP64 atomics are native.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgba_amd.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = [
    "k_linearize.hip",
    "k_blocks.hip",
    "k_schur.hip",
    "k_dense.hip",
    "k_precond.hip",
    "k_cg.hip",
    "k_heavy.hip",
    "k_refine.hip",
    "k_backsub.hip",
    "k_batch.hip",
    "k_pairs.hip",
    "k_layout.hip",
    "noalloc.hip",
    "solver.cpp",
    "batch.cpp",
    "obs_fp16.cpp",
    "capi.cpp",
    "cpp_api.cpp",
    "host_model.cpp",
    "synth.cpp",
    "camera_graph.cpp",
]

ARCH = os.environ.get("GBA_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags():
    return [
        "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++20",
        "-munsafe-fp-atomics",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
        f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
    ]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + common_flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for (o, err), s in zip(results, SOURCES):
            warn = "\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l)
            if warn.strip():
                print(f"[{s}]\n{warn}", file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
