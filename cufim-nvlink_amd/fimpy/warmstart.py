"""Surrogate warm start: training maps from the batched solver, training of
the coordinate MLP, and two ways to use it on a batch of new solves.

  scheduled (the default, `scheduled_batch`): the surrogate predicts each
    solve's latest activation time, and every instance of the batch gets its
    own time window (fim_set_window) so that all of them finish in about the
    same number of global iterations.  Windows change only the processing
    order, never the fixed point, so no check of the result is needed.
  upper bound (`warm_solve_all`): prediction + delta as the initial field of
    every vertex, the support check as judge and a cold re-solve as
    fallback.  Kept for comparison: under the window order the front lowers
    every vertex below any certified upper bound before it acts, so this
    prunes nothing.

A solve is (sites, scales): 1-3 pacing vertices (time 0) and four regional
speed factors.  All work runs through libcufim.so (fim_nn_*, fim_*).
"""
from __future__ import annotations

import time

import numpy as np

from . import PART_RCB, Pack, Solver, SolverGroup, Surrogate  # noqa: F401
from .workloads import pacing_sites, splitmix64


def draw_solves(cfg, n, seed, max_sites=3):
    """n solves on a box config: 1..max_sites pacing vertices of the z = 0
    face and four factors uniform in [0.5, 1.5], from SplitMix64(seed)."""
    g = splitmix64(seed)
    out = []
    for _ in range(n):
        k = 1 + next(g) % max_sites
        sites = pacing_sites(cfg, k, next(g))
        sc = np.array([0.5 + (next(g) >> 11) * 2.0 ** -53 for _ in range(4)], dtype=np.float32)
        out.append((sites, sc))
    return out


def _chunks(seq, n):
    for i in range(0, len(seq), n):
        yield i, seq[i:i + n]


def window_sample(mesh, k=1000, seed=0):
    """Vertices at which the latest activation of a solve is estimated: the
    vertex nearest each corner of the mesh's bounding box and k seeded
    random others (the latest activation lies on the boundary, mostly at a
    corner)."""
    xyz, _, _ = mesh.arrays()
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    corners = [int(np.argmin(np.sum(((xyz - np.array([a, b, c])) / span) ** 2, axis=1)))
               for a in (lo[0], hi[0]) for b in (lo[1], hi[1]) for c in (lo[2], hi[2])]
    rng = np.random.default_rng(seed)
    rand = rng.choice(xyz.shape[0], min(k, xyz.shape[0]), replace=False)
    return np.unique(np.concatenate([corners, rand])).astype(np.int64)


def scheduled_windows(tmax, window, quantile=0.75, floor=250.0):
    """Per-solve time windows (us per global iteration) under which the
    solves of one batch need about the same number of iterations: the batch
    takes K = quantile(tmax) / window iterations and solve j gets
    tmax_j / K, at least `floor`.  A solve's window sets how much of its
    front one iteration may process; the fixed point does not depend on it."""
    tmax = np.asarray(tmax, dtype=np.float64)
    k = max(1.0, float(np.quantile(tmax, quantile)) / window)
    return np.maximum(floor, tmax / k)


def _contexts(solver, n):
    """(context, first, end) parts of instances 0..n-1: a SolverGroup's
    contexts, or the one context with every instance"""
    return solver.parts(n) if isinstance(solver, SolverGroup) else [(solver, 0, n)]


def _batch(solver, solves, wins):
    """Set-up and solve of len(solves) instances, instance i at window
    wins[i]: every context of a SolverGroup takes its own instances and the
    contexts run together."""
    for ctx, a, b in _contexts(solver, len(solves)):
        ctx.reset()
        for i in range(a, b):
            ctx.set_scaling(solves[i][1], inst=i - a)
            ctx.set_window(float(wins[i]), inst=i - a)
        ctx.seed_batch([s for s, _ in solves[a:b]], 0.0)
    solver.run(max_iters=1_000_000)
    return solver.stats()


def windowed_batch(solver, solves, wins):
    """Batched solve of len(solves) instances, instance i at window wins[i],
    on a Solver or a SolverGroup; returns the solver stats."""
    return _batch(solver, solves, wins)


def scheduled_batch(nn: Surrogate, solver, solves, sample, window=2000.0, quantile=0.75,
                    wins_out=None):
    """Surrogate-scheduled batched solve of len(solves) instances: predicted
    latest activation per solve (one device call over the vertex sample),
    per-instance windows, regional factors, seeds, then FIM to the fixed
    point.  Returns (stats, seconds of prediction, seconds of the rest); the
    windows are appended to wins_out when given."""
    t0 = time.perf_counter()
    tmax = nn.predict_max(solves, sample)
    t1 = time.perf_counter()
    wins = scheduled_windows(tmax, window, quantile)
    if wins_out is not None:
        wins_out.extend(float(w) for w in wins)
    st = windowed_batch(solver, solves, wins)
    t2 = time.perf_counter()
    return st, t1 - t0, t2 - t1


def cold_batch(solver, solves, window=None):
    """Cold batched solve of len(solves) instances, every instance at one
    window (the context's, or `window` when given); returns the solver
    stats."""
    w = solver.window if window is None else window
    return _batch(solver, solves, np.full(len(solves), w))


def _note(on_stats, st):
    if on_stats is not None:
        on_stats(st)


def build_dataset(nn: Surrogate, pack: Pack, solves, batch=256, window=2000.0, unit=0,
                  profile=False, on_stats=None, local_iters=0, block=0):
    """Solves every solve cold in batches and appends the maps to the
    surrogate's dataset.  Returns seconds spent.  on_stats(stats) is called
    after every batch (profile=True times every compute launch)."""
    t0 = time.perf_counter()
    s = Solver(pack, n_instances=batch, window=window, unit=unit, profile=profile,
               local_iters=local_iters, block=block)
    try:
        for i, chunk in _chunks(solves, batch):
            _note(on_stats, cold_batch(s, chunk + [chunk[0]] * (batch - len(chunk))))
            nn.add_maps(s, 0, [c[0] for c in chunk], np.stack([c[1] for c in chunk]))
    finally:
        s.close()
    return time.perf_counter() - t0


def surrogate_errors(nn: Surrogate, pack: Pack, solves, batch=256, window=2000.0, profile=False,
                     on_stats=None, local_iters=0, block=0):
    """Per-solve (max underestimate, max |error|, mean |error|, max T) of the
    bare surrogate against the converged FIM maps, in us."""
    out = []
    s = Solver(pack, n_instances=batch, window=window, profile=profile, local_iters=local_iters,
               block=block)
    try:
        for i, chunk in _chunks(solves, batch):
            _note(on_stats, cold_batch(s, chunk + [chunk[0]] * (batch - len(chunk))))
            for j, (sites, sc) in enumerate(chunk):
                ref, _ = s.download(inst=j)
                pred = nn.predict_map(sites, sc)
                d = ref - pred
                out.append((float(d.max()), float(np.abs(d).max()), float(np.abs(d).mean()),
                            float(ref.max())))
    finally:
        s.close()
    return out


def warm_batch(nn: Surrogate, solver: Solver, solves, delta, all_active=True):
    """Warm-started batch: prediction + delta as the initial field of every
    instance, then FIM to the fixed point.  Returns (stats, seconds of
    inference + initialisation, seconds of FIM)."""
    solver.reset()
    t0 = time.perf_counter()
    for i, (sites, sc) in enumerate(solves):
        solver.set_scaling(sc, inst=i)
        nn.warm_start(solver, i, sites, sc, delta, all_active=all_active)
        if not all_active:
            solver.seed(sites, 0.0, inst=i)
    t1 = time.perf_counter()
    solver.run(max_iters=1_000_000)
    t2 = time.perf_counter()
    return solver.stats(), t1 - t0, t2 - t1


def judge(solver: Solver, n, tol=1e-4):
    """Instances whose field lies below its fixed point (support > tol)."""
    return [i for i in range(n) if solver.support(inst=i) > tol]


def warm_solve_all(nn: Surrogate, pack: Pack, solves, delta, batch=256, window=2000.0, unit=0,
                   all_active=True, tol=1e-4, judge_maps=True, profile=False, on_stats=None,
                   local_iters=0, block=0):
    """Warm-started batched solves with the support check as judge and a cold
    re-solve of the rejected ones.  Returns a dict of times and counts."""
    s = Solver(pack, n_instances=batch, window=window, unit=unit, profile=profile,
               local_iters=local_iters, block=block)
    fb = Solver(pack, n_instances=batch, window=window, unit=unit, profile=profile,
                local_iters=local_iters, block=block)
    rep = dict(maps=0, infer_s=0.0, fim_s=0.0, judge_s=0.0, fallback_s=0.0, fallbacks=0,
               global_iters=0, activations=0, elem_updates=0)
    try:
        for i, chunk in _chunks(solves, batch):
            full = chunk + [chunk[0]] * (batch - len(chunk))
            st, ti, tf = warm_batch(nn, s, full, delta, all_active)
            _note(on_stats, st)
            rep["infer_s"] += ti
            rep["fim_s"] += tf
            for k in ("global_iters", "activations", "elem_updates"):
                rep[k] += st[k]
            if judge_maps:
                t0 = time.perf_counter()
                bad = judge(s, len(chunk), tol)
                rep["judge_s"] += time.perf_counter() - t0
                if bad:
                    t0 = time.perf_counter()
                    redo = [chunk[j] for j in bad]
                    _note(on_stats, cold_batch(fb, redo + [redo[0]] * (batch - len(redo))))
                    rep["fallback_s"] += time.perf_counter() - t0
                    rep["fallbacks"] += len(bad)
            rep["maps"] += len(chunk)
    finally:
        s.close()
        fb.close()
    return rep
