"""Host check of the Jacobian product forms the FP64 refinement residual
uses (snavely_model.hpp: edge_jc_x, edge_jpT_m, edge_jp_z, edge_jcT_t):
tests/cpp/edge_products.cpp compares each with the product of the explicit
Jacobians of snavely_eval<true> (the linearisation's) over 30,000 random
observations in the three rotation regimes (closed form, series below
theta = 1e-2, first-order model at theta = 0), relative to the sum of the
magnitudes of the product's terms, and checks that the residual is the same.
A wrong cross-product order, a dropped rotation term or a missing transpose
gives O(1) differences."""
import os
import re
import subprocess

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "edge_products.cpp")
INC = os.path.join(ROOT, "gpu-bundle-adjustment_amd", "csrc")


def test_edge_product_forms_match_explicit_jacobians(tmp_path):
    exe = str(tmp_path / "edge_products")
    r = subprocess.run(["g++", "-std=c++17", "-O2", f"-I{INC}", SRC, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"cases (\d+) jc_x (\S+) jpT_m (\S+) jp_z (\S+) jcT_t (\S+)", r.stdout)
    assert m, r.stdout
    assert int(m.group(1)) > 29000
    for v in m.groups()[1:]:
        assert float(v) < 1e-12, r.stdout
