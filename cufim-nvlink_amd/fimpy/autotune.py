"""Autotuner of the converge-mode solver: picks the submesh size, the
workgroup shape, the local sweeps per global iteration, the time window, the
compute kernel's register budget, the work-unit size and the number of
contexts the batch is split over for a mesh and batch size by measuring
short batched solves on the GPU.

Search: coordinate descent from the defaults over
    max_nn      {256, 512, 1024, 1536, 2048, 2816}  vertex copies per submesh
    block       {64, 128, 256, 512}                 workgroup size
    local_iters {0, 1, 2, 4, 8}               sweeps per activation (0 = to local convergence)
    window      {500, 1000, 2000, 4000}       microseconds per global iteration
    waves       {0, 5}                        waves per SIMD of the compute kernel (0 = LDS bound;
                                              workgroups of 256 only)
    unit        {1, 2, 3, 4}                  queued instances of a submesh per work unit
    contexts    {1, 2}                        contexts the batch is split over, run together
                                              (fim_run_multi)
Each trial times one batch solve after one warm-up solve.  Results are cached
in a JSON file keyed by a mesh signature (sizes, bounding box, coordinate
checksum) and the batch size.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import PART_RCB, Mesh, Pack, Solver, SolverGroup

SPACE = {
    "max_nn": [256, 512, 1024, 1536, 2048, 2816],
    "block": [64, 128, 256, 512],
    "local_iters": [0, 1, 2, 4, 8],
    "window": [500.0, 1000.0, 2000.0, 4000.0],
    "waves": [0, 5],
    "unit": [1, 2, 3, 4],
    "contexts": [1, 2],
}
DEFAULT = {"max_nn": 2816, "block": 512, "local_iters": 4, "window": 2000.0, "waves": 0, "unit": 2,
           "contexts": 1}


def signature(mesh: Mesh, n_instances: int) -> str:
    xyz, e2n, _ = mesh.arrays()
    h = hashlib.sha256()
    h.update(np.array([mesh.nn, mesh.ne, n_instances], dtype=np.int64).tobytes())
    h.update(np.round(xyz.min(axis=0), 3).tobytes())
    h.update(np.round(xyz.max(axis=0), 3).tobytes())
    h.update(np.float64(np.round(xyz.sum(), 1)).tobytes())
    h.update(np.int64(e2n.sum()).tobytes())
    return h.hexdigest()[:24]


def _measure(mesh, packs, cfg, seeds, scales, seed):
    key = cfg["max_nn"]
    if key not in packs:
        packs[key] = Pack(mesh, PART_RCB, max_nn=key, seed=seed)
    n = len(seeds)
    nctx = max(1, min(int(cfg.get("contexts", 1)), n))
    cut = np.linspace(0, n, nctx + 1).astype(int)
    ss = [Solver(packs[key], n_instances=int(cut[k + 1] - cut[k]), block=cfg["block"],
                 local_iters=cfg["local_iters"], window=cfg["window"], waves=cfg.get("waves", 0),
                 unit=cfg.get("unit", 0)) for k in range(nctx)]
    s = ss[0] if nctx == 1 else SolverGroup(ss)
    parts = [(ss[0], 0, n)] if nctx == 1 else s.parts(n)
    try:
        def solve():
            for ctx, a, b in parts:
                ctx.reset()
                if scales is not None:
                    for i in range(a, b):
                        ctx.set_scaling(scales[i], inst=i - a)
                ctx.seed_batch(seeds[a:b], 0.0)
            s.run(max_iters=1_000_000)
        solve()
        t0 = time.perf_counter()
        solve()
        dt = time.perf_counter() - t0
    finally:
        s.close()
    return len(seeds) / dt


def tune(mesh: Mesh, seeds, scales=None, space=None, cache=None, seed=0, rounds=2, log=None):
    """Returns (best config, list of (config, maps/s) trials).  `seeds` is a
    list of per-instance pacing vertex arrays; `scales` optional (n, 4)."""
    space = space or SPACE
    sig = signature(mesh, len(seeds))
    if cache and os.path.exists(cache):
        db = json.load(open(cache))
        if sig in db:
            return db[sig]["best"], db[sig]["trials"]
    packs = {}
    trials = []
    seen = {}

    def canonical(cfg):
        # the register budget only selects a kernel variant for workgroups of
        # 256; other workgroup sizes always run the 8-wave variant
        if "waves" in cfg and cfg.get("block", DEFAULT["block"]) != 256:
            return dict(cfg, waves=0)
        return cfg

    def run(cfg):
        cfg = canonical(cfg)
        k = json.dumps(cfg, sort_keys=True)
        if k not in seen:
            seen[k] = _measure(mesh, packs, cfg, seeds, scales, seed)
            trials.append((dict(cfg), seen[k]))
            if log:
                log(f"[autotune] {cfg} -> {seen[k]:.1f} maps/s")
        return seen[k]

    best = canonical({k: (DEFAULT[k] if DEFAULT[k] in v else v[0]) for k, v in space.items()})
    best_v = run(best)
    for _ in range(rounds):
        improved = False
        for k in [k for k in ("window", "max_nn", "block", "local_iters", "waves", "unit", "contexts")
                  if k in space]:
            for v in space[k]:
                cand = canonical(dict(best, **{k: v}))
                r = run(cand)
                if r > best_v:
                    best, best_v, improved = cand, r, True
        if not improved:
            break
    if cache:
        db = json.load(open(cache)) if os.path.exists(cache) else {}
        db[sig] = {"best": best, "maps_per_s": best_v, "trials": trials}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        json.dump(db, open(cache, "w"), indent=1)
    return best, trials
