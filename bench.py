#!/usr/bin/env python3
"""bench.py -- activation maps/s of the MI355X FIM solver.

Workload (one GPU): the anisotropic slab -- 160 x 160 x 13 hexes x 6 tets =
1,996,800 tets, 362,894 vertices over 10 x 10 x 2 cm, fibres rotating
-60..+60 degrees through the thickness, v_fibre 0.6 and v_cross 0.24 m/s --
with five pacing vertices per map on the z = 0 face.  One step = one batch of
B maps (different pacing sets) solved to the eps = 1e-6 fixed point by the
batched HIP solver, mesh resident in HBM; reset and seeding are inside the
timed region.  The map of the first instance is checked against the FP64
fixed point of the CPU oracle (max |dT| <= 0.1 ms is required for a valid
number), and that oracle run, timed on one host core, is the CPU baseline.
Plain single-solve GPU FIM (B = 1) is timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cufim-nvlink_amd"))

import fimpy as F  # noqa: E402
from fimpy import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
METRIC = "maps/s @max|ΔT|≤0.1ms, 2M-tet slab: batched+NN warm-start vs plain FIM"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pacing_sets(cfg, batch, seed):
    return [W.pacing_sites(cfg, 5, seed * 100003 + i) for i in range(batch)]


def solve_batch(s, sets):
    s.reset()
    for i, src in enumerate(sets):
        s.seed(src, 0.0, inst=i)
    return s.run(max_iters=1_000_000)


def traffic_from_profiles(args):
    """HBM bytes per compute launch from the committed PMC summary, when this
    run's configuration is the one that was profiled (else None)."""
    p = os.path.join(ROOT, "profiles", "conv_band_slab_pmc.md")
    if not os.path.exists(p):
        return None
    text = open(p).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    c = re.search(r"traffic_config:\s*(.*)", text)
    mine = (f"batch={args.batch} max_nn={args.max_nn} block={args.block} unit={args.unit} "
            f"window={float(args.window)} staged={args.staged}")
    src = re.search(r"traffic_source:\s*([0-9a-f]+)", text)
    if not m or not c or c.group(1).strip() != mine or not src or src.group(1) != kernel_source_digest():
        return None
    return float(m.group(1))


def kernel_source_digest():
    """sha256 (first 16 hex digits) of the kernel sources; a PMC summary
    applies only to the sources it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in ("solver.hip", "local_solve.hpp"):
        h.update(open(os.path.join(ROOT, "cufim-nvlink_amd", "csrc", f), "rb").read())
    return h.hexdigest()[:16]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--max-nn", type=int, default=1024)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--unit", type=int, default=2,
                    help="queued instances of one submesh run back to back by one workgroup")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--graph-steps", type=int, default=0,
                    help="also time this many steps with graph-captured iterations")
    ap.add_argument("--staged", action="store_true",
                    help="stage whole submeshes in LDS (default: element data from L2)")
    ap.add_argument("--window", type=float, default=2000.0,
                    help="time window per global iteration in us (0 = plain FIM order)")
    # plain single-solve GPU FIM: its own best setting (fewer, smaller work
    # items want smaller submeshes and a finer window)
    ap.add_argument("--plain-max-nn", type=int, default=256)
    ap.add_argument("--plain-window", type=float, default=1000.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X only: the batched solve would shard by map across "
              "GPUs with no exchange; only the single-GPU measurement is implemented.", file=sys.stderr)
        sys.exit(1)

    cfg = W.SLAB
    t0 = time.perf_counter()
    mesh = W.box(cfg, args.seed)
    pack = F.Pack(mesh, F.PART_RCB, max_nn=args.max_nn, seed=args.seed)
    info = pack.info
    log(f"[bench] mesh {mesh.nn} vertices {mesh.ne} tets, {info['nsub']} submeshes, "
        f"setup {time.perf_counter() - t0:.2f}s")
    sets = pacing_sets(cfg, args.batch, args.seed)

    # Every solve below launches its iterations individually with a HIP
    # event pair around each compute-kernel launch on the solver's stream
    # (profile mode: the same kernel sequence the graph replays).  The
    # roofline uses the average over all those launches, which is also what
    # `rocprofv3 --kernel-trace --stats` reports for this command.
    kern = dict(alg_bytes=0, compute_ms=0.0, compute_launches=0)

    def account(st):
        for k in kern:
            kern[k] += st[k]

    # ---- batched solver, timed -----------------------------------------
    s = F.Solver(pack, n_instances=args.batch, block=args.block, profile=True, unit=args.unit,
                 window=args.window, staged=args.staged)
    for i in range(args.warmup):
        solve_batch(s, sets)
        account(s.stats())
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    tot = dict(elem_updates=0, global_iters=0, activations=0, local_iters=0, vertex_updates=0)
    t0 = time.perf_counter()
    for i in range(args.steps):
        solve_batch(s, sets)
        st = s.stats()
        account(st)
        for k in tot:
            tot[k] += st[k]
    elapsed = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} steps of {args.batch} maps in {elapsed:.3f}s")
    phi0, _ = s.download(inst=0)
    res0 = s.residual(inst=0)
    s.close()
    ms_per_step = 1e3 * elapsed / args.steps
    maps_per_s = args.batch * args.steps / elapsed

    # ---- optionally: the same batch with graph-captured iterations (the
    # library default; its launches cannot carry events, so it is off here
    # to keep every compute launch of this command timed)
    graph_maps_per_s = None
    if args.graph_steps > 0:
        g = F.Solver(pack, n_instances=args.batch, block=args.block, window=args.window, unit=args.unit,
                     staged=args.staged)
        solve_batch(g, sets)
        t0 = time.perf_counter()
        for i in range(args.graph_steps):
            solve_batch(g, sets)
        graph_maps_per_s = args.batch * args.graph_steps / (time.perf_counter() - t0)
        phi0g, _ = g.download(inst=0)
        g.close()
        assert np.array_equal(phi0, phi0g), "graph replay differs from direct launches"

    # ---- plain single-solve GPU FIM (the denominator of the target) -----
    ppack = pack if args.plain_max_nn == args.max_nn else \
        F.Pack(mesh, F.PART_RCB, max_nn=args.plain_max_nn, seed=args.seed)
    p1 = F.Solver(ppack, n_instances=1, block=args.block, profile=True, window=args.plain_window,
                  staged=args.staged)
    solve_batch(p1, sets[:1])
    account(p1.stats())
    t0 = time.perf_counter()
    for i in range(args.plain_steps):
        solve_batch(p1, sets[i % len(sets):i % len(sets) + 1])
        account(p1.stats())
    plain_s = (time.perf_counter() - t0) / args.plain_steps
    p1.close()
    log(f"[bench] plain single solve {1e3 * plain_s:.2f} ms")

    # ---- accuracy and CPU baseline (FP64 oracle, one host core) ---------
    max_dt = None
    cpu = None
    if not (args.no_check and args.no_cpu_baseline):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O  # accuracy check and CPU baseline only
        xyz, e2n, _ = mesh.arrays()
        fib, vel = mesh.material()
        om = O.Mesh(xyz, e2n, fib=fib, vel=vel)
        t0 = time.perf_counter()
        ref, ost = O.fim_solve(om, sets[0], variant=O.LOCAL_EXACT)
        cpu_s = time.perf_counter() - t0
        fin = ref < 1e50
        if not np.array_equal(fin, phi0 < 1e50):
            max_dt = float("inf")
        else:
            max_dt = float(np.max(np.abs(phi0[fin] - ref[fin])))
        log(f"[bench] oracle {cpu_s:.2f}s, max|dT| = {max_dt:.3f} us (max T {ref[fin].max():.0f} us)")
        cpu = {"value": 1.0 / cpu_s, "unit": "maps/s", "cores": 1, "kind": "port",
               "sample": "one map of the same slab and pacing set (instance 0), FP64 "
                         "label-correcting fixed point of the same local update, one host core"}
    valid = max_dt is not None and max_dt <= 100.0

    bytes_per_launch = kern["alg_bytes"] / max(kern["compute_launches"], 1)
    avg_launch_s = kern["compute_ms"] / 1e3 / max(kern["compute_launches"], 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": maps_per_s if (valid or args.no_check) else 0.0,
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "anisotropic slab 160x160x13 hexes x 6 = 1,996,800 tets / 362,894 vertices, "
                        "fibres -60..+60 deg, v_f 0.6 / v_c 0.24 m/s, 5 pacing sites per map; "
                        f"batch of {args.batch} maps per step solved to eps 1e-6 "
                        "(batched FIM; NN warm start not included)",
            "maps_per_step": args.batch,
            "partition": f"recursive bisection, <= {args.max_nn} vertex copies per submesh, "
                         f"{info['nsub']} submeshes",
            "workgroup": args.block,
            "unit": args.unit,
            "window_us": args.window,
            "staged": args.staged,
        },
        "max_abs_dT_us": max_dt,
        "tolerance_us": 100.0,
        "valid": valid,
        "residual_inst0": res0,
        "graph_maps_per_s": graph_maps_per_s,
        "plain_maps_per_s": 1.0 / plain_s,
        "plain_config": {"max_nn": args.plain_max_nn, "window_us": args.plain_window},
        "batched_over_plain": maps_per_s * plain_s,
        "solves_per_s": maps_per_s,
        "elem_evals_per_s": tot["elem_updates"] / elapsed,
        "vertex_updates_per_s": tot["vertex_updates"] / elapsed,
        "global_iters_per_step": tot["global_iters"] / args.steps,
        "activations_per_step": tot["activations"] / args.steps,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_conv_compute" if args.staged else "k_conv_band",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(args),
            "alg_bytes_per_launch": bytes_per_launch,
            "avg_launch_us": 1e6 * avg_launch_s,
            "launches": kern["compute_launches"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
