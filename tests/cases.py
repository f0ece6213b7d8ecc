"""Shared workloads for tests and the benchmark (sizes in micrometres,
speeds in micrometres per microsecond, times in microseconds)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UT = os.path.join(GOLDEN, "ut_mesh")


def splitmix64(seed):
    """Deterministic stream of 64-bit integers (SplitMix64)."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def pick(n_choices, k, seed):
    """k distinct indices in [0, n_choices) drawn from SplitMix64(seed)."""
    g = splitmix64(seed)
    out = []
    while len(out) < k:
        v = next(g) % n_choices
        if v not in out:
            out.append(v)
    return out


# configuration 1: unit cube 1 cm, 20^3 hexes x 6, jitter 10 %, 0.6 m/s,
# pacing at the origin
C1 = dict(lx=10000.0, ly=10000.0, lz=10000.0, nx=20, ny=20, nz=20, jitter=0.1,
          angle0=0.0, angle1=0.0, vf=0.6, vc=0.6)
# configuration 2: slab 10 x 10 x 2 cm, 160 x 160 x 13 hexes x 6, fibres
# rotating -60..+60 degrees through the thickness, vf 0.6, vc 0.24,
# five pacing vertices on z = 0
C2 = dict(lx=100000.0, ly=100000.0, lz=20000.0, nx=160, ny=160, nz=13, jitter=0.1,
          angle0=-60.0, angle1=60.0, vf=0.6, vc=0.24)


def scaled(cfg, f):
    """Same physical box with the hex counts divided by f (test sizes)."""
    c = dict(cfg)
    for k in ("nx", "ny", "nz"):
        c[k] = max(2, int(round(cfg[k] / f)))
    return c


def box(cfg, seed):
    """Product mesh of a box config and the oracle's view of the same input
    (vertex coordinates, sorted element nodes, per-element fibre/speeds)."""
    import fimpy as F
    import pyoracle as O
    m = F.Mesh.gen_box(seed=seed, **cfg)
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()
    return m, O.Mesh(xyz, e2n, fib=fib, vel=vel)


def face_vertices(cfg, z_index=0):
    nx, ny = cfg["nx"], cfg["ny"]
    NX, NY = nx + 1, ny + 1
    return np.array([i + NX * (j + NY * z_index) for j in range(NY) for i in range(NX)], dtype=np.int64)


def pacing_sites(cfg, k, seed):
    face = face_vertices(cfg, 0)
    return face[pick(face.shape[0], k, seed)]


def ut_plan_stimuli():
    plan = json.load(open(os.path.join(UT, "ut_mesh.plan.json")))
    st = plan["simulation"]["stimuli"]["S0"]
    return [(st["active"], st["protocol"]["start"] * 1000.0, st["electrode"]["vtxdata"])]
