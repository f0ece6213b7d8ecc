#!/usr/bin/env python3
"""bench.py -- activation maps/s of the MI355X FIM solver, batched and
surrogate-scheduled, against plain single-solve FIM.

Workload (one GPU): the anisotropic slab -- 160 x 160 x 13 hexes x 6 tets =
1,996,800 tets, 362,894 vertices over 10 x 10 x 2 cm, fibres rotating
-60..+60 degrees through the thickness, v_fibre 0.6 and v_cross 0.24 m/s --
split into four quadrant regions.  A solve is 1-3 pacing vertices on the
z = 0 face plus four regional speed factors U[0.5, 1.5] (seeded).

The surrogate (coordinate MLP, BF16 MFMA) is trained in the run on 4,096
maps of the batched solver; that is set-up, outside the timed region.  One
step = the 512 held-out solves as one batch: the surrogate predicts each
solve's latest activation time (GPU inference over a vertex sample), every
instance gets its own time window from it (fim_set_window), and the batched
FIM runs to the eps = 1e-6 fixed point as two solver contexts of 256 run
together (fim_run_multi: one context's launch tails filled by the other's
work), iterations replayed from captured hipGraphs, mesh resident in HBM.  `value` is that rate; inference, seeding
and reset are inside the timed region.

Legs of the same run, in this order:
  surrogate    dataset (batched solves, every compute launch event-timed)
               and training;
  accounting   one scheduled batch of the held-out solves, in one context, with every compute
               launch timed by HIP events on the solver's stream (profile
               mode, the kernel sequence the graph replays): the
               k_conv_band dispatches right after the dataset's, from which
               the roofline comes;
  headline     W warm-up and K timed scheduled batches (graph replay);
  cold         the same solves at one uniform window (the method without
               the surrogate), same contexts, K timed batches;
  plain        plain single-solve GPU FIM on held-out solves (graph replay,
               its own best settings) -- the target's denominator;
  slab batch   the five-site slab maps of the earlier headline: a batch of
               2048 maps of the batched solver alone;
  compat       the reference's time-window algorithm (compat mode, FP64) on
               one five-site map: N fim_step windows, timed like the
               reference (step loop only);
  CPU          the same compat map on all host cores (the oracle's OpenMP
               restatement of the reference algorithm, step loop timed): the
               CPU baseline, compared with the GPU map bit for bit; then the
               FP64 oracle's fixed points of the first held-out solves, one
               per host thread, which validate the number (max |dT| <= 0.1 ms).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
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
from fimpy import warmstart as WS  # noqa: E402
from fimpy import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every
# 4 cycles at 2.4 GHz
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
METRIC = "maps/s @max|ΔT|≤0.1ms, 2M-tet slab: batched+NN warm-start vs plain FIM"
NN_PARAMS = 333569


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_source_digest():
    """sha256 (first 16 hex digits) of the kernel sources without // comments
    and blank lines; a counter summary applies only to the code it was
    measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in ("solver.hip", "local_solve.hpp"):
        for line in open(os.path.join(ROOT, "cufim-nvlink_amd", "csrc", f)).read().splitlines():
            code = re.sub(r"//.*$", "", line).rstrip()
            if code:
                h.update(code.encode() + b"\n")
    return h.hexdigest()[:16]


def config_key(args):
    return (f"held_out={args.test} max_nn={args.max_nn} block={args.block} unit={args.unit} "
            f"local_iters={args.local_iters} window={float(args.window)} quantile={args.quantile} "
            f"sample={args.sample}")


def profile_figures(args, name):
    """key: value figures of a committed counter summary, when it was
    measured on this configuration and this kernel source (else None)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    text = open(p).read()
    src = re.search(r"^source:\s*([0-9a-f]+)", text, re.M)
    cfg = re.search(r"^config:\s*(.*)$", text, re.M)
    if not src or src.group(1) != kernel_source_digest() or not cfg or cfg.group(1).strip() != config_key(args):
        return None
    return {k: float(v) for k, v in re.findall(r"^([a-z_]+):\s*([-0-9.eE+]+)\s*$", text, re.M)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def compat_leg(args, mesh, sites):
    """The reference's time-window algorithm (compat mode, FP64) on the slab:
    the fim_step loop timed like the reference's solve (step loop only,
    cuda_fim.cu:2424-2447), one pacing set, own partition.  Returns the
    report and what the CPU side needs to replay the same solve."""
    pack = F.Pack(mesh, F.PART_RCB, max_nn=args.compat_max_nn, seed=args.seed)
    s = F.Solver(pack, mode=F.MODE_COMPAT, local=F.LOCAL_REF)
    fw = mesh.front_width()
    stims = [(1, 0.0, [int(v) for v in sites])]
    best = None
    for rep in range(2):  # the first run warms the device up
        s.reset()
        s.seed(sites, 0.0)
        cur = 0.0
        t0 = time.perf_counter()
        for _ in range(args.compat_steps):
            ps = cur
            cur += args.compat_dt
            s.step(ps, cur, args.compat_dt, fw)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    phi, st = s.download()
    s.close()
    e2s, _ = pack.partition()
    rep = {"workload": "one five-site map of the slab (no regional factors)",
           "max_nn": args.compat_max_nn, "submeshes": pack.info["nsub"], "dt_us": args.compat_dt,
           "steps": args.compat_steps, "gpu_ms_per_step": 1e3 * best / args.compat_steps,
           "gpu_maps_per_s": 1.0 / best, "reached": int(np.sum(phi < 1e50))}
    log(f"[bench] compat: {args.compat_steps} windows of {args.compat_dt:.0f} us, "
        f"{rep['gpu_ms_per_step']:.3f} ms per window on the GPU")
    return rep, (phi, st, e2s, stims, fw)


def slab_batch_leg(args, mesh_plain):
    """The batched solver alone on five-site maps of the slab (no regional
    factors): args.slab_batch maps per batch."""
    cfg = W.SLAB
    pack = F.Pack(mesh_plain, F.PART_RCB, max_nn=args.max_nn, seed=args.seed)
    sets = [W.pacing_sites(cfg, 5, args.seed * 100003 + i) for i in range(args.slab_batch)]
    s = F.Solver(pack, n_instances=args.slab_batch, block=args.block, unit=args.unit,
                 local_iters=args.local_iters, window=args.window)

    def one():
        s.reset()
        s.seed_batch(sets, 0.0)
        s.run(max_iters=1_000_000)

    one()
    t0 = time.perf_counter()
    for _ in range(args.slab_steps):
        one()
    dt = (time.perf_counter() - t0) / args.slab_steps
    st = s.stats()
    s.close()
    rep = {"workload": f"{args.slab_batch} five-site maps of the slab per batch (no regional factors), "
                       "batched FIM at one window, graph replay",
           "maps_per_s": args.slab_batch / dt, "ms_per_batch": 1e3 * dt, "global_iters": st["global_iters"],
           "elem_evals_per_elem_per_map": st["elem_updates"] / args.slab_batch / mesh_plain.ne}
    log(f"[bench] five-site slab batch: {rep['maps_per_s']:.1f} maps/s")
    return rep


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--test", type=int, default=512, help="held-out solves = maps per step (one batch)")
    ap.add_argument("--max-nn", type=int, default=2816)
    ap.add_argument("--block", type=int, default=512)
    ap.add_argument("--local-iters", type=int, default=4,
                    help="sweeps per activation before requeueing (0 = to local convergence)")
    ap.add_argument("--unit", type=int, default=3,
                    help="queued instances of one submesh run back to back by one workgroup "
                         "(3 with two contexts; 2 is the best for one)")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--window", type=float, default=2000.0,
                    help="time window per global iteration in us: the cold batch's, and the base of "
                         "the scheduled windows")
    ap.add_argument("--quantile", type=float, default=0.75,
                    help="scheduled windows: the batch runs quantile(predicted latest time) / window "
                         "iterations")
    ap.add_argument("--sample", type=int, default=1000,
                    help="random vertices (besides the 8 bounding-box corners) the latest activation "
                         "is predicted at")
    ap.add_argument("--nn-maps", type=int, default=4096)
    ap.add_argument("--nn-steps", type=int, default=4000)
    ap.add_argument("--nn-batch", type=int, default=65536)
    ap.add_argument("--plain-steps", type=int, default=5)
    # plain single-solve GPU FIM: its own best setting, found by a sweep of
    # submesh size (256..2816 copies), workgroup (128/256/512), sweeps (2..8)
    # and window (1..4 ms): one map wants more, smaller submeshes
    ap.add_argument("--plain-max-nn", type=int, default=1024)
    ap.add_argument("--plain-window", type=float, default=2000.0)
    ap.add_argument("--plain-local-iters", type=int, default=3)
    ap.add_argument("--plain-block", type=int, default=512)
    ap.add_argument("--slab-batch", type=int, default=2048)
    ap.add_argument("--slab-steps", type=int, default=2)
    ap.add_argument("--compat-steps", type=int, default=150,
                    help="windows of the reference time-window algorithm (compat mode, FP64) on one "
                         "map, GPU and all host cores (0 = skip); 150 x 2 ms covers the slab's maps")
    ap.add_argument("--compat-dt", type=float, default=2000.0)
    ap.add_argument("--compat-max-nn", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU legs (0 = all cores this process may use)")
    ap.add_argument("--check-maps", type=int, default=16,
                    help="held-out maps checked against the FP64 oracle (one per host thread)")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--windows-out", default="",
                    help="save the accounting batch's per-instance windows (.npy)")
    ap.add_argument("--windows-in", default="",
                    help="counter runs: replay saved per-instance windows for the accounting batch "
                         "without the surrogate, print the roofline fields and stop")
    ap.add_argument("--contexts", type=int, default=2,
                    help="solver contexts the held-out batch is split over (fim_run_multi: their "
                         "streams overlap, one context's launch tails filled by another's work)")
    ap.add_argument("--nn-only", action="store_true",
                    help="counter runs: build the surrogate's dataset (--nn-maps) and train "
                         "(--nn-steps), print the training rate and stop")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X only: the batched solve would shard by map across "
              "GPUs with no exchange; only the single-GPU measurement is implemented.", file=sys.stderr)
        sys.exit(1)

    cfg = W.SLAB
    t0 = time.perf_counter()
    mesh = W.box(cfg, args.seed)
    reg = W.quadrant_regions(mesh)
    mesh.set_regions(reg)
    pack = F.Pack(mesh, F.PART_RCB, max_nn=args.max_nn, seed=args.seed)
    info = pack.info
    log(f"[bench] mesh {mesh.nn} vertices {mesh.ne} tets, {info['nsub']} submeshes, "
        f"setup {time.perf_counter() - t0:.2f}s")
    host_cores = len(os.sched_getaffinity(0))
    n_threads = args.cpu_threads if args.cpu_threads > 0 else host_cores
    train = WS.draw_solves(cfg, args.nn_maps, 1001 + args.seed)
    test = WS.draw_solves(cfg, args.test, 3003 + args.seed)
    kern = dict(alg_bytes=0, compute_ms=0.0, compute_launches=0, elem_updates=0, activations=0,
                vertex_updates=0, global_iters=0, local_iters=0)

    pre = dict(compute_launches=0)
    if args.windows_in:
        # the accounting batch alone, at the windows a full run saved: the
        # same k_conv_band dispatches with no surrogate kernel in the process
        acc = F.Solver(pack, n_instances=args.test, block=args.block, profile=True, unit=args.unit,
                       local_iters=args.local_iters, window=args.window)
        st = WS.windowed_batch(acc, test, np.load(args.windows_in))
        acc.close()
        # and the reference algorithm's windows (compat mode), for its counters
        compat = None
        if args.compat_steps > 0:
            compat, _ = compat_leg(args, W.box(cfg, args.seed), W.pacing_sites(cfg, 5, args.seed * 100003))
        print(json.dumps({"accounting_only": True, "roofline": {
            "launches": st["compute_launches"], "launches_skipped": 0,
            "avg_launch_us": 1e3 * st["compute_ms"] / max(st["compute_launches"], 1),
            "alg_bytes_per_launch": st["alg_bytes"] / max(st["compute_launches"], 1),
            "vertex_updates_per_launch": st["vertex_updates"] / max(st["compute_launches"], 1)},
            "compat": compat}), flush=True)
        return

    # ---- surrogate: training maps and training (set-up) ----------------------
    nn = F.Surrogate(mesh, seed=args.seed, max_maps=args.nn_maps, max_batch=args.nn_batch)

    def count(st):
        pre["compute_launches"] += st["compute_launches"]

    t_data = WS.build_dataset(nn, pack, train, batch=args.test, window=args.window, unit=args.unit,
                              profile=True, on_stats=count, local_iters=args.local_iters, block=args.block)
    log(f"[bench] surrogate dataset: {args.nn_maps} maps in {t_data:.2f}s")
    t0 = time.perf_counter()
    done, loss = 0, float("nan")
    while done < args.nn_steps:
        k = min(1000, args.nn_steps - done)
        loss = nn.train(k, batch=args.nn_batch, seed=args.seed * 7919 + done)
        done += k
        log(f"[bench] surrogate training {done}/{args.nn_steps} steps, loss {loss:.3e}")
    t_train = time.perf_counter() - t0
    if args.nn_only:
        nn.close()
        print(json.dumps({"surrogate_only": True, "steps": args.nn_steps, "batch": args.nn_batch,
                          "ms_per_step": 1e3 * t_train / max(args.nn_steps, 1), "loss": loss}), flush=True)
        return
    sample = WS.window_sample(mesh, k=args.sample, seed=args.seed)

    # ---- accounting: event-timed compute launches of one scheduled batch ----
    acc = F.Solver(pack, n_instances=args.test, block=args.block, profile=True, unit=args.unit,
                   local_iters=args.local_iters, window=args.window)
    acc_wins = []
    st, _, _ = WS.scheduled_batch(nn, acc, test, sample, window=args.window, quantile=args.quantile,
                                  wins_out=acc_wins)
    if args.windows_out:
        np.save(args.windows_out, np.array(acc_wins, dtype=np.float64))
    for k in kern:
        kern[k] += st[k]
    acc_map0 = acc.download(inst=0)[0]
    acc.close()
    log(f"[bench] accounting batch: {st['compute_launches']} launches after the dataset's "
        f"{pre['compute_launches']}, {st['compute_ms']:.1f} ms of compute kernel")

    # ---- headline: surrogate-scheduled batches, graph replay ----------------
    # the batch split over `contexts` solver contexts that run together; every
    # map is the one a single context gives (checked below against the
    # accounting batch's first map)
    nctx = max(1, min(args.contexts, args.test))
    cut = np.linspace(0, args.test, nctx + 1).astype(int)
    s = F.SolverGroup([F.Solver(pack, n_instances=int(cut[k + 1] - cut[k]), block=args.block, unit=args.unit,
                                local_iters=args.local_iters, window=args.window) for k in range(nctx)])
    for i in range(args.warmup):
        WS.scheduled_batch(nn, s, test, sample, window=args.window, quantile=args.quantile)
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    tot = dict(elem_updates=0, global_iters=0, activations=0, local_iters=0, vertex_updates=0)
    t_inf = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        st, ti, _ = WS.scheduled_batch(nn, s, test, sample, window=args.window, quantile=args.quantile)
        t_inf += ti
        for k in tot:
            tot[k] += st[k]
        log(f"[bench] step {i + 1}/{args.steps}")
    elapsed = time.perf_counter() - t0
    maps_per_s = args.test * args.steps / elapsed
    log(f"[bench] timed {args.steps} scheduled batches of {args.test} held-out solves in {elapsed:.3f}s "
        f"({maps_per_s:.1f} maps/s)")
    n_check = max(1, min(args.check_maps, args.test))
    sched_maps = [s.download(inst=i)[0] for i in range(n_check)]
    res0 = s.residual(inst=0)
    assert np.array_equal(sched_maps[0], acc_map0), "graph replay differs from direct launches"

    # ---- cold: the same solves at one window (no surrogate) ----------------
    WS.cold_batch(s, test)
    cold_tot = dict(elem_updates=0, global_iters=0, activations=0)
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = WS.cold_batch(s, test)
        for k in cold_tot:
            cold_tot[k] += st[k]
    cold_s = (time.perf_counter() - t0) / args.steps
    cold_maps = [s.download(inst=i)[0] for i in range(n_check)]
    cold_tmax = np.array([s.download(inst=i)[0].max() for i in range(args.test)])
    s.close()
    log(f"[bench] cold batches: {args.test / cold_s:.1f} maps/s")
    pred_tmax = nn.predict_max(test, sample)

    # ---- plain single-solve GPU FIM (the denominator of the target) --------
    ppack = F.Pack(mesh, F.PART_RCB, max_nn=args.plain_max_nn, seed=args.seed)
    p1 = F.Solver(ppack, n_instances=1, block=args.plain_block, window=args.plain_window,
                  local_iters=args.plain_local_iters)
    WS.cold_batch(p1, test[:1])
    t0 = time.perf_counter()
    for i in range(args.plain_steps):
        WS.cold_batch(p1, test[i % len(test):i % len(test) + 1])
    plain_s = (time.perf_counter() - t0) / args.plain_steps
    p1.close()
    log(f"[bench] plain single solve {1e3 * plain_s:.2f} ms (graph replay)")

    # ---- the five-site slab batch and the reference algorithm ---------------
    mesh_plain = W.box(cfg, args.seed)
    slab = slab_batch_leg(args, mesh_plain) if args.slab_batch > 0 else None
    five = W.pacing_sites(cfg, 5, args.seed * 100003)
    compat, compat_replay = (None, None)
    if args.compat_steps > 0:
        compat, compat_replay = compat_leg(args, mesh_plain, five)

    # ---- CPU legs: the oracle is the checker and the CPU baseline only ------
    max_dt, cpu, sur_vs_oracle = None, None, None
    if not (args.no_check and args.no_cpu_baseline):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O  # accuracy check and CPU baselines only
        from concurrent.futures import ThreadPoolExecutor
        xyz, e2n, _ = mesh.arrays()
        fib, vel = mesh.material()
        if compat_replay is not None and not args.no_cpu_baseline:
            cphi, cst, e2s, stims, fw = compat_replay
            omc = O.Mesh(xyz, e2n, fib=fib, vel=vel, atags=e2s.astype(np.int32))
            t0 = time.perf_counter()
            ophi, ost, sec = O.compat_solve_par(omc, stims, args.compat_dt, args.compat_steps, n_threads,
                                                front_width=fw)
            wall = time.perf_counter() - t0
            del omc
            compat["bit_identical"] = bool(np.array_equal(cphi, ophi) and np.array_equal(cst, ost))
            compat["cpu_ms_per_step"] = 1e3 * sec[1] / args.compat_steps
            compat["gpu_over_cpu"] = compat["cpu_ms_per_step"] / compat["gpu_ms_per_step"]
            log(f"[bench] compat on {n_threads} host threads: step loop {sec[1]:.2f}s, setup {sec[0]:.2f}s, "
                f"bit-identical to the GPU {compat['bit_identical']}")
            cpu = {"value": 1.0 / sec[1], "unit": "maps/s", "cores": n_threads, "kind": "port",
                   "sample": f"one five-site map of the slab, the reference's time-window algorithm "
                             f"(cuda_fim.cu:1752-2334) as the oracle's FP64 restatement on {n_threads} "
                             f"OpenMP threads, one submesh per thread at a time ({compat['submeshes']} "
                             f"submeshes of <= {args.compat_max_nn} copies, own partition): "
                             f"{args.compat_steps} windows of {args.compat_dt:.0f} us; value = 1 / step-loop "
                             f"seconds ({sec[1]:.2f} s), timed as the reference times its solve",
                   "end_to_end_maps_per_s": 1.0 / wall, "setup_s": float(sec[0]),
                   "nproc": host_cores, "cpu_model": cpu_model()}
        if not args.no_check:
            def oracle_map(i):  # the C oracle releases the GIL
                sites, sc = test[i]
                om = O.Mesh(xyz, e2n, fib=fib, vel=vel * sc.astype(np.float64)[reg][:, None])
                t_start = time.perf_counter()
                ref_i, _ = O.fim_solve(om, sites, variant=O.LOCAL_EXACT)
                return ref_i, time.perf_counter() - t_start

            nth = min(n_check, n_threads)
            log(f"[bench] FP64 oracle: {n_check} held-out maps on {nth} host threads")
            t0 = time.perf_counter()
            with ThreadPoolExecutor(nth) as ex:
                res = list(ex.map(oracle_map, range(n_check)))
            wall = time.perf_counter() - t0
            max_dt, cold_dt = 0.0, 0.0
            for (ref, _), gmap, cmap in zip(res, sched_maps, cold_maps):
                fin = ref < 1e50
                if not np.array_equal(fin, gmap < 1e50):
                    max_dt = float("inf")
                else:
                    max_dt = max(max_dt, float(np.max(np.abs(gmap[fin] - ref[fin]))))
                    cold_dt = max(cold_dt, float(np.max(np.abs(cmap[fin] - ref[fin]))))
            times = [dt for _, dt in res]
            log(f"[bench] oracle {n_check} maps in {wall:.2f}s (per map {min(times):.1f}-{max(times):.1f}s), "
                f"max|dT| scheduled {max_dt:.3f} us, cold {cold_dt:.3f} us")
            ref0 = res[0][0]
            pred0 = nn.predict_map(*test[0])
            sur_vs_oracle = {"map": 0, "mean_us": float(np.mean(np.abs(pred0 - ref0))),
                             "max_us": float(np.max(np.abs(pred0 - ref0))),
                             "cold_fim_max_us": cold_dt,
                             "oracle_maps_per_s_on_threads": n_check / wall, "threads": nth}
    nn.close()
    valid = max_dt is not None and max_dt <= 100.0

    # ---- roofline of k_conv_band from the accounting launches ---------------
    nl = max(kern["compute_launches"], 1)
    bytes_per_launch = kern["alg_bytes"] / nl
    avg_launch_s = kern["compute_ms"] / 1e3 / nl
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    pmc = profile_figures(args, "conv_band_c5_counters.md")
    valu, traffic = None, None
    if pmc:
        traffic = pmc.get("traffic_bytes_per_launch")
        ipl = pmc.get("valu_insts_per_launch")
        cand = kern["vertex_updates"] / nl
        if ipl:
            valu = {"insts_per_launch": ipl,
                    "issue_frac": ipl / avg_launch_s / VALU_ISSUE_PEAK if avg_launch_s > 0 else None,
                    "insts_per_candidate": ipl / cand if cand else None,
                    "lane_slots_per_candidate": 64 * ipl / cand if cand else None,
                    "lds_conflict_cycles_per_lds_inst": pmc.get("lds_conflicts_per_lds_inst"),
                    "wait_frac": pmc.get("wait_frac"),
                    "peak_insts_per_s": VALU_ISSUE_PEAK}
    one_pass = 32.0 * mesh.ne + 10.0 * mesh.nn
    nmaps_acc = args.test
    cold_rate = args.test / cold_s
    out = {
        "metric": METRIC,
        "value": maps_per_s if (valid or args.no_check) else 0.0,
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "anisotropic slab 160x160x13 hexes x 6 = 1,996,800 tets / 362,894 vertices, fibres "
                        "-60..+60 deg, v_f 0.6 / v_c 0.24 m/s, four quadrant regions; "
                        f"{args.test} held-out solves per step as one batch split over {nctx} solver "
                        "contexts run together (1-3 pacing sites on z = 0, "
                        "regional speed factors U[0.5, 1.5]), each instance's time window from the "
                        "surrogate's predicted latest activation (inference inside the timed region), "
                        "batched FIM to eps 1e-6, graph replay; surrogate trained in the run on "
                        f"{args.nn_maps} batched-FIM maps (set-up, not timed)",
            "maps_per_step": args.test,
            "contexts": nctx,
            "partition": f"recursive bisection, <= {args.max_nn} vertex copies per submesh, "
                         f"{info['nsub']} submeshes",
            "max_nn": args.max_nn,
            "workgroup": args.block,
            "unit": args.unit,
            "local_iters": args.local_iters,
            "window_us": args.window,
            "quantile": args.quantile,
            "sample": args.sample,
            "sample_vertices": len(sample),
        },
        "max_abs_dT_us": max_dt,
        "tolerance_us": 100.0,
        "valid": valid,
        "residual_inst0": res0,
        "plain_maps_per_s": 1.0 / plain_s,
        "plain_config": {"max_nn": args.plain_max_nn, "window_us": args.plain_window,
                         "local_iters": args.plain_local_iters, "workgroup": args.plain_block,
                         "timing": f"graph replay, one held-out solve per solve, solves 0..{args.plain_steps - 1}"},
        "batched_over_plain": maps_per_s * plain_s,
        "solves_per_s": maps_per_s,
        "elem_evals_per_s": tot["elem_updates"] / elapsed,
        "vertex_updates_per_s": tot["vertex_updates"] / elapsed,
        "global_iters_per_step": tot["global_iters"] / args.steps,
        "activations_per_step": tot["activations"] / args.steps,
        "nn_warm_start": {
            "method": "surrogate-scheduled windows: predicted latest activation per solve (BF16 MLP on a "
                      f"{len(sample)}-vertex sample), instance window = predicted / K with K = "
                      f"quantile_{args.quantile}(predicted) / {args.window:.0f} us; the fixed point does not "
                      "depend on the windows, so no judge or fallback is needed",
            "maps_per_s": maps_per_s,
            "cold_maps_per_s": cold_rate,
            "gain_over_cold": maps_per_s / cold_rate,
            "warm_over_plain": maps_per_s * plain_s,
            "inference_ms_per_batch": 1e3 * t_inf / args.steps,
            "global_iters": {"cold": cold_tot["global_iters"] / args.steps,
                             "scheduled": tot["global_iters"] / args.steps},
            "activations_per_map": {"cold": cold_tot["activations"] / args.steps / args.test,
                                    "scheduled": tot["activations"] / args.steps / args.test},
            "elem_evals_per_map": {"cold": cold_tot["elem_updates"] / args.steps / args.test,
                                   "scheduled": tot["elem_updates"] / args.steps / args.test},
            "fallbacks": 0,
            "predicted_latest_time_rel_err": {
                "mean": float(np.mean(np.abs(pred_tmax - cold_tmax) / cold_tmax)),
                "max": float(np.max(np.abs(pred_tmax - cold_tmax) / cold_tmax))},
            "latest_time_ms": {"min": float(cold_tmax.min() / 1e3), "median": float(np.median(cold_tmax) / 1e3),
                               "max": float(cold_tmax.max() / 1e3)},
            "surrogate_vs_oracle_us": sur_vs_oracle,
            "training": {"maps": args.nn_maps, "dataset_s": t_data, "steps": args.nn_steps,
                         "batch": args.nn_batch, "train_s": t_train, "final_loss": loss,
                         "tflops": 6.0 * NN_PARAMS * args.nn_batch * args.nn_steps / t_train / 1e12,
                         "mfma_peak_tflops": MFMA_BF16_PEAK_TFLOPS},
        },
        "sweep_overhead": {
            "activations_per_submesh_per_map": kern["activations"] / nmaps_acc / info["nsub"],
            "elem_evals_per_elem_per_map": kern["elem_updates"] / nmaps_acc / mesh.ne,
            "candidates_per_vertex_per_map": kern["vertex_updates"] / nmaps_acc / mesh.nn,
            "sweeps_per_activation": kern["local_iters"] / max(kern["activations"], 1),
            "alg_bytes_per_map_over_one_pass": kern["alg_bytes"] / nmaps_acc / one_pass,
            "one_pass_bytes_per_map": one_pass,
        },
        "roofline": {
            "bound": "valu" if valu else "hbm",
            "kernel": "k_conv_band",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "alg_bytes_per_launch": bytes_per_launch,
            "bytes_per_unit": "32 per distinct element evaluated + 10 per vertex copy + 12 per interface "
                              "copy of each activation",
            "avg_launch_us": 1e6 * avg_launch_s,
            "launches": kern["compute_launches"],
            "launches_skipped": pre["compute_launches"],
            "launches_are": "k_conv_band dispatches launches_skipped+1 .. launches_skipped+launches of the "
                            "process (the accounting batch, after the dataset's)",
            "valu": valu,
        },
        "slab_batch": slab,
        "cpu_baseline": cpu,
        "compat": compat,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
