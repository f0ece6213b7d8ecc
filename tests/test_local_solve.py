"""The product's local update (cufim-nvlink_amd/csrc/local_solve.hpp),
compiled for the host with g++, against the oracle's restatement on random
element geometries and node times.  FP64 must agree to rounding; the
reference update (compat mode) must agree bit for bit; the FP32 exact path
must agree to FP32 accuracy.  CPU only."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cufim-nvlink_amd", "csrc")

SHIM = r"""
#include "local_solve.hpp"
extern "C" {
double ls_exact64(const double *g, const double *ph, int t) {
  const double inf = 1e300;
  switch (t) {
    case 0: return fim::cand_exact<0, double>(g, ph, inf);
    case 1: return fim::cand_exact<1, double>(g, ph, inf);
    case 2: return fim::cand_exact<2, double>(g, ph, inf);
    default: return fim::cand_exact<3, double>(g, ph, inf);
  }
}
float ls_exact32(const float *g, const float *ph, int t) {
  const float inf = 3.0e38f;
  switch (t) {
    case 0: return fim::cand_exact<0, float>(g, ph, inf);
    case 1: return fim::cand_exact<1, float>(g, ph, inf);
    case 2: return fim::cand_exact<2, float>(g, ph, inf);
    default: return fim::cand_exact<3, float>(g, ph, inf);
  }
}
double ls_ref(const double *g, const double *ph, int t) { return fim::cand_ref(g, ph, t); }
}
"""


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("ls")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    f"-I{CSRC}", str(src), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    pd, pf = C.POINTER(C.c_double), C.POINTER(C.c_float)
    lib.ls_exact64.restype = C.c_double
    lib.ls_exact64.argtypes = [pd, pd, C.c_int]
    lib.ls_exact32.restype = C.c_float
    lib.ls_exact32.argtypes = [pf, pf, C.c_int]
    lib.ls_ref.restype = C.c_double
    lib.ls_ref.argtypes = [pd, pd, C.c_int]
    return lib


def _cases(n, seed):
    """Geometry of random non-degenerate tets under a random anisotropic
    metric, node times spread so that every branch (edge, triangle, face,
    unusable node) is taken."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        x = rng.normal(size=(4, 3)) * 500.0
        a = rng.normal(size=(3, 3))
        M = a @ a.T + 0.2 * np.eye(3)
        M /= 0.36
        if abs(np.linalg.det(x[1:] - x[0])) < 1e5:
            continue
        g = O.geo_elem(x, np.arange(4), M)
        ph = rng.uniform(0, 2000.0, 4)
        if rng.random() < 0.15:
            ph[rng.integers(4)] = 1e300
        out.append((np.ascontiguousarray(g), ph))
    return out


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def test_exact_fp64_matches_oracle(shim):
    worst = 0.0
    for g, ph in _cases(3000, 1):
        for t in range(4):
            ours = shim.ls_exact64(_ptr(g, C.c_double), _ptr(ph, C.c_double), t)
            ref = O.local(O.LOCAL_EXACT, g, ph, t)
            if ref >= 1e299:
                assert ours >= 1e299
                continue
            worst = max(worst, abs(ours - ref) / max(1.0, abs(ref)))
    assert worst < 1e-9, worst


def test_reference_update_bit_identical(shim):
    for g, ph in _cases(3000, 2):
        ph = np.where(ph >= 1e299, np.finfo(np.float64).max, ph)
        for t in range(4):
            ours = shim.ls_ref(_ptr(g, C.c_double), _ptr(ph, C.c_double), t)
            ref = O.local(O.LOCAL_REF, g, ph, t)
            assert ours == ref or (np.isnan(ours) and np.isnan(ref))


def test_exact_fp32_close_to_oracle(shim):
    worst = 0.0
    for g, ph in _cases(3000, 3):
        gf = g.astype(np.float32)
        pf = np.where(ph >= 1e299, np.float32(3.0e38), ph).astype(np.float32)
        for t in range(4):
            ours = shim.ls_exact32(_ptr(gf, C.c_float), _ptr(pf, C.c_float), t)
            ref = O.local(O.LOCAL_EXACT, g, ph, t)
            if ref >= 1e299:
                continue
            assert ours < 3.0e38  # an edge candidate always exists here
            worst = max(worst, abs(ours - ref) / max(1.0, abs(ref)))
    assert worst < 2e-4, worst
