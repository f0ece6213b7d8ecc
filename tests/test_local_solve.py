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
    """Geometry of random well-shaped tets (a regular tet of side ~600 um,
    vertices moved by up to 30 % of a side, random orientation) under a
    random transversely anisotropic metric (speeds 0.2-0.6 um/us, as in
    cardiac tissue), node times spread so that every branch (edge,
    triangle, face, unusable node) is taken."""
    rng = np.random.default_rng(seed)
    base = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, np.sqrt(3) / 2, 0.0],
                     [0.5, np.sqrt(3) / 6, np.sqrt(2.0 / 3.0)]]) * 600.0
    out = []
    while len(out) < n:
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        x = (base + rng.uniform(-0.3, 0.3, (4, 3)) * 600.0) @ q
        if abs(np.linalg.det(x[1:] - x[0])) < 1e6:
            continue
        r, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        v = rng.uniform(0.2, 0.6, 3)
        M = r @ np.diag(1.0 / v ** 2) @ r.T
        packed = np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])
        g = O.geo_elem(x, np.arange(4), packed)
        assert np.all(g > 0)  # SPD metric: positive squared lengths
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


SHIM_SH = r"""
#include "local_solve.hpp"
extern "C" int ls_shared_equal(const float *g, const float *ph) {
  float sq[6], rc[6];
  for (int k = 0; k < 6; ++k) { sq[k] = fim::csqrt(g[k]); rc[k] = 1.0f / g[k]; }
  const float inf = 3.0e38f;
  const float a[4] = {fim::cand_exact<0, float>(g, ph, inf), fim::cand_exact<1, float>(g, ph, inf),
                      fim::cand_exact<2, float>(g, ph, inf), fim::cand_exact<3, float>(g, ph, inf)};
  const float b[4] = {fim::cand_exact_sh<0, float>(g, sq, rc, ph, inf),
                      fim::cand_exact_sh<1, float>(g, sq, rc, ph, inf),
                      fim::cand_exact_sh<2, float>(g, sq, rc, ph, inf),
                      fim::cand_exact_sh<3, float>(g, sq, rc, ph, inf)};
  int bad = 0;
  for (int k = 0; k < 4; ++k) {
    const float tol = 1e-4f * (a[k] < inf ? a[k] : 1.0f);
    if (!((a[k] >= inf && b[k] >= inf) || (a[k] - b[k] <= tol && b[k] - a[k] <= tol))) ++bad;
  }
  return bad;
}
"""


def test_shared_root_candidate_equals_plain_candidate(tmp_path):
    """the variant with per-element square roots and reciprocals (device
    path) gives the same candidates.  On the host a/b and a*(1/b) round
    differently, so a validity test at a face or edge boundary can flip and
    select the neighbouring candidate, which is continuous there: 1e-4
    relative; a wrong edge or reciprocal would be off by far more."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "sh.cpp"
    src.write_text(SHIM_SH)
    so = tmp_path / "sh.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{CSRC}",
                    str(src), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    pf = C.POINTER(C.c_float)
    lib.ls_shared_equal.argtypes = [pf, pf]
    bad = 0
    for g, ph in _cases(2000, 7):
        gf = g.astype(np.float32)
        pf_ = np.where(ph >= 1e299, np.float32(3.0e38), ph).astype(np.float32)
        bad += lib.ls_shared_equal(_ptr(gf, C.c_float), _ptr(pf_, C.c_float))
    assert bad == 0
