"""Builds libcufim.so (HIP kernels for gfx950 + C++ host code + C ABI) and
the reference-format CLI `cufim` in-tree:

    python cufim-nvlink_amd/build.py          # -> cufim-nvlink_amd/lib/libcufim.so
                                              #    cufim-nvlink_amd/bin/cufim
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libcufim.so")
BIN = os.path.join(HERE, "bin", "cufim")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
# device code: gfx950 only; packed-FP32 instructions disabled
HIPFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]

SOURCES = [
    ("solver.hip", True),
    ("surrogate.hip", True),
    ("mesh.cpp", False),
    ("partition.cpp", False),
    ("host_capi.cpp", False),
    ("config.cpp", False),
    ("solver_fim.cpp", False),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def _compile(src, is_hip):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    if is_hip:
        cmd = [HIPCC, *COMMON, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    else:
        # host-only C++; same compiler driver so the objects link cleanly
        cmd = [HIPCC, *COMMON, "-x", "c++", "-c", os.path.join(CSRC, src), "-o", out]
    _run(cmd)
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    srcs = [(s, h) for s, h in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda a: _compile(*a), srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-soname,libcufim.so"])
    os.replace(tmp, LIB)
    cli = os.path.join(CSRC, "cli_main.cpp")
    if os.path.exists(cli):
        tmpb = BIN + ".tmp"
        _run([HIPCC, *COMMON, "-x", "c++", cli, "-o", tmpb, f"-L{os.path.dirname(LIB)}", "-lcufim",
              "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(tmpb, BIN)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
