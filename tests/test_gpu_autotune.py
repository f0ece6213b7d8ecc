"""Autotuner (fimpy/autotune.py) on a batched multi-mesh workload: it picks
a configuration among the candidates by measurement, caches it by mesh
signature, and the tuned configuration still reaches the oracle's maps."""
import json

import numpy as np
import pytest

import cases
import fimpy as F
import pyoracle as O
from fimpy import autotune as A
from fimpy import workloads as W

pytestmark = pytest.mark.gpu


def test_autotune_picks_measured_best_and_caches(tmp_path):
    parts = [(n, cases.scaled(c, 4)) for n, c in W.MULTI]
    m, offs, cfgs = W.multi_mesh(parts, 3)
    solves = W.multi_solves(cfgs, offs, 16, 5)
    seeds = [s["sites"] for s in solves]
    scales = np.stack([s["scale"] for s in solves])
    space = {"max_nn": [256, 512], "block": [128, 256], "local_iters": [0, 2], "window": [1000.0],
             "waves": [0, 5], "unit": [2, 3], "contexts": [1, 2]}
    cache = str(tmp_path / "tune.json")
    best, trials = A.tune(m, seeds, scales, space=space, cache=cache, rounds=1)
    assert best["max_nn"] in space["max_nn"] and best["block"] in space["block"]
    rates = {json.dumps(c, sort_keys=True): r for c, r in trials}
    assert rates[json.dumps(best, sort_keys=True)] == max(rates.values())
    again, _ = A.tune(m, seeds, scales, space=space, cache=cache)
    assert again == best  # served from the cache
    # the tuned configuration solves correctly (one instance vs the oracle)
    p = F.Pack(m, F.PART_RCB, max_nn=best["max_nn"], seed=0)
    s = F.Solver(p, n_instances=1, block=best["block"], local_iters=best["local_iters"],
                 window=best["window"], waves=best["waves"])
    s.seed(seeds[1], 0.0)
    s.set_scaling(scales[1])
    s.run()
    phi, _ = s.download()
    s.close()
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()
    reg = np.concatenate([W.quadrant_regions(F.Mesh.gen_box(seed=3 + k, **c)) for k, (_, c) in enumerate(parts)])
    om = O.Mesh(xyz, e2n, fib=fib, vel=vel * scales[1].astype(np.float64)[reg][:, None])
    ref, _ = O.fim_solve(om, seeds[1], variant=O.LOCAL_EXACT)
    fin = ref < 1e50
    assert np.array_equal(fin, phi < 1e50)
    assert np.max(np.abs(phi[fin] - ref[fin])) <= 1e-4 * ref[fin].max()
