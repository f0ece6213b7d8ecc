"""Shared workloads for tests and the benchmark (sizes in micrometres,
speeds in micrometres per microsecond, times in microseconds)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UT = os.path.join(GOLDEN, "ut_mesh")


from fimpy.workloads import CUBE, SLAB, pacing_sites, pick, scaled, splitmix64  # noqa: F401,E402


def box(cfg, seed):
    """Product mesh of a box config and the oracle's view of the same input
    (vertex coordinates, sorted element nodes, per-element fibre/speeds)."""
    import fimpy as F
    import pyoracle as O
    m = F.Mesh.gen_box(seed=seed, **cfg)
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()
    return m, O.Mesh(xyz, e2n, fib=fib, vel=vel)


def ut_plan_stimuli():
    plan = json.load(open(os.path.join(UT, "ut_mesh.plan.json")))
    st = plan["simulation"]["stimuli"]["S0"]
    return [(st["active"], st["protocol"]["start"] * 1000.0, st["electrode"]["vtxdata"])]
