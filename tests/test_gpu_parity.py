"""GPU parity: the HIP solver (through the C ABI) against the CPU oracle.

Tolerances
  compat mode (FP64, reference local update): bit-identical to the oracle's
    restatement on the same partition; against the reference golden
    (printed with 6 decimals) within 1e-6 us.
  converge mode (FP32, exact local update): max |dT| <= 1e-4 * max(T)
    against the FP64 fixed point of the oracle, the per-config bound the
    solver is specified to.
"""
import numpy as np
import pytest

import cases
import fimpy as F
import pyoracle as O
import vtkio

pytestmark = pytest.mark.gpu


def _compat_run(mesh, pack, stims, dt, n_steps, local=F.LOCAL_REF):
    s = F.Solver(pack, mode=F.MODE_COMPAT, local=local)
    fw = mesh.front_width()
    s.reset()
    for active, tstart, vtx in stims:  # update_events_gpu at cur_time 0
        if not (tstart >= 0.0 and active):
            break
        s.seed(vtx, 0.0)
    cur = 0.0
    for _ in range(n_steps):
        ps = cur
        cur += dt
        s.step(ps, cur, dt, fw)
    phi, st = s.download()
    s.close()
    return phi, st


def test_compat_ut_mesh_matches_golden_and_oracle():
    m = F.Mesh.load_vtk(f"{cases.UT}/ut_mesh.vtk")
    m.read_tags(f"{cases.UT}/ut_mesh.tags")
    p = F.Pack(m, F.PART_TAGS)
    stims = cases.ut_plan_stimuli()
    phi, st = _compat_run(m, p, stims, 150e3, 3)
    g = vtkio.read_vtk(f"{cases.UT}/cuda_fim_res.vtk")
    gp, gs = g["point_data"]["phi"], g["point_data"]["states"]
    assert np.array_equal(st, gs.astype(np.int32))
    ours = np.where(phi < 1e50, phi, -1.0)
    assert np.max(np.abs(ours - gp)) <= 1e-6
    xyz, e2n, _ = m.arrays()
    om = O.Mesh(xyz, e2n, atags=np.loadtxt(f"{cases.UT}/ut_mesh.tags", dtype=np.int32))
    ophi, ost, _ = O.compat_solve(om, stims, 150e3, 3)
    assert np.array_equal(st, ost)
    assert np.array_equal(phi, ophi)  # bit-identical FP64


@pytest.mark.parametrize("steps", [1, 2, 4])
def test_compat_ut_mesh_windows(steps):
    m = F.Mesh.load_vtk(f"{cases.UT}/ut_mesh.vtk")
    m.read_tags(f"{cases.UT}/ut_mesh_p.tags")
    p = F.Pack(m, F.PART_TAGS)
    stims = cases.ut_plan_stimuli()
    phi, st = _compat_run(m, p, stims, 150e3, steps)
    xyz, e2n, _ = m.arrays()
    om = O.Mesh(xyz, e2n, atags=np.loadtxt(f"{cases.UT}/ut_mesh_p.tags", dtype=np.int32))
    ophi, ost, _ = O.compat_solve(om, stims, 150e3, steps)
    assert np.array_equal(st, ost)
    assert np.array_equal(phi, ophi)


def test_compat_box_rcb_partition():
    """reference semantics on a generated mesh with the own partitioner"""
    cfg = cases.scaled(cases.CUBE, 2)
    m, om = cases.box(cfg, seed=3)
    p = F.Pack(m, F.PART_RCB, max_nn=128, seed=3)
    e2s, _ = p.partition()
    om = O.Mesh(om.xyz, om.e2n, fib=om.fib, vel=om.vel, atags=e2s.astype(np.int32))
    stims = [(1, 0.0, [0])]
    fw = m.front_width()
    n_steps = int(np.ceil(1.8 * 10000 / 0.6 / 1000.0)) + 2
    phi, st = _compat_run(m, p, stims, 1000.0, n_steps)
    ophi, ost, _ = O.compat_solve(om, stims, 1000.0, n_steps, front_width=fw)
    assert np.array_equal(st, ost)
    assert np.array_equal(phi, ophi)


def _converge(m, pack, sources, **kw):
    s = F.Solver(pack, **kw)
    s.seed(sources, 0.0)
    it = s.run()
    phi, st = s.download()
    res = s.residual()
    stats = s.stats()
    s.close()
    return phi, st, it, res, stats


def _check_fp32(phi, ref):
    fin = ref < 1e50
    assert np.array_equal(fin, phi < 1e50)
    err = np.max(np.abs(phi[fin] - ref[fin]))
    tol = 1e-4 * ref[fin].max()
    assert err <= tol, (err, tol)
    return err


def test_converge_cube_48k_vs_fp64_oracle():
    m, om = cases.box(cases.CUBE, seed=1)
    p = F.Pack(m, F.PART_RCB, max_nn=512, seed=1)
    phi, st, it, res, stats = _converge(m, p, [0])
    ref, _ = O.fim_solve(om, [0], variant=O.LOCAL_EXACT)
    _check_fp32(phi, ref)
    assert res < 1e-5
    assert np.all(st == F.FROZEN)


def test_converge_anisotropic_slab_vs_fp64_oracle():
    cfg = cases.scaled(cases.SLAB, 4)
    m, om = cases.box(cfg, seed=2)
    p = F.Pack(m, F.PART_RCB, max_nn=512, seed=2)
    src = cases.pacing_sites(cfg, 5, seed=2)
    phi, st, it, res, stats = _converge(m, p, src)
    ref, _ = O.fim_solve(om, src, variant=O.LOCAL_EXACT)
    _check_fp32(phi, ref)


@pytest.mark.parametrize("staged", [False, True])
@pytest.mark.parametrize("block", [64, 128, 256, 512])
def test_converge_workgroup_shapes_agree(block, staged):
    cfg = cases.scaled(cases.SLAB, 8)
    m, om = cases.box(cfg, seed=5)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=5)
    src = cases.pacing_sites(cfg, 3, seed=5)
    phi, *_ = _converge(m, p, src, block=block, staged=staged)
    ref, _ = O.fim_solve(om, src, variant=O.LOCAL_EXACT)
    _check_fp32(phi, ref)


def test_converge_long_element_lists_and_pair_path_agree():
    # 2048-copy submeshes seeded on their whole bottom faces: sweeps mark more
    # elements than one element-list pass holds, so they run in several
    # passes.  Workgroups of 256 and 512 evaluate (element, target) pairs, of
    # 128 one element per lane: the same sweeps, so bit-identical maps, all
    # within tolerance of the FP64 oracle.
    cfg = cases.scaled(cases.SLAB, 4)
    m, om = cases.box(cfg, seed=17)
    p = F.Pack(m, F.PART_RCB, max_nn=2048, seed=17)
    assert p.info["ne_max"] < (1 << 14)  # the pair path is taken at 256
    xyz, _, _ = m.arrays()
    src = np.nonzero(xyz[:, 2] < xyz[:, 2].min() + 0.2 * cfg["lz"] / cfg["nz"])[0]
    assert len(src) > 1000  # a bottom layer of ~2,400 tets per submesh is marked at once
    maps = {}
    for block in (128, 256, 512):
        s = F.Solver(p, block=block, window=0.0, local_iters=4)
        s.seed(src, 0.0)
        s.run()
        maps[block] = s.download()[0]
        s.close()
    assert np.array_equal(maps[128], maps[256]) and np.array_equal(maps[128], maps[512])
    ref, _ = O.fim_solve(om, src, variant=O.LOCAL_EXACT)
    _check_fp32(maps[256], ref)


@pytest.mark.parametrize("max_nn", [256, 1280])
def test_converge_register_variants_agree(max_nn):
    # the compute kernel built for 5..8 waves per SIMD: same arithmetic, so the
    # same fixed point bit for bit, single and batched, and the FP64 oracle's
    cfg = cases.scaled(cases.SLAB, 8)
    m, om = cases.box(cfg, seed=11)
    p = F.Pack(m, F.PART_RCB, max_nn=max_nn, seed=11)
    srcs = [cases.pacing_sites(cfg, k + 1, seed=40 + k) for k in range(3)]
    ref, _ = O.fim_solve(om, srcs[0], variant=O.LOCAL_EXACT)
    base = None
    for waves in (0, 5, 6, 7, 8):
        s = F.Solver(p, n_instances=3, window=50.0, local_iters=4, waves=waves, block=256)
        for i, src in enumerate(srcs):
            s.seed(src, 0.0, inst=i)
        s.run()
        maps = [s.download(inst=i)[0] for i in range(3)]
        s.close()
        _check_fp32(maps[0], ref)
        if base is None:
            base = maps
        for a_, b_ in zip(base, maps):
            assert np.array_equal(a_, b_), waves


def test_seed_batch_equals_per_instance_seeds():
    cfg = cases.scaled(cases.SLAB, 8)
    m, _ = cases.box(cfg, seed=13)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=13)
    sets = [cases.pacing_sites(cfg, k % 3 + 1, seed=60 + k) for k in range(4)]
    a = F.Solver(p, n_instances=5, window=50.0)
    b = F.Solver(p, n_instances=5, window=50.0)
    for i, src in enumerate(sets):
        a.seed(src, 0.0, inst=i + 1)
    b.seed_batch(sets + [np.zeros(0, np.int64)], 0.0, insts=[1, 2, 3, 4, 0])
    a.run()
    b.run()
    for i in range(5):
        pa, sa = a.download(inst=i)
        pb, sb = b.download(inst=i)
        assert np.array_equal(pa, pb) and np.array_equal(sa, sb), i
    assert np.all(a.download(inst=0)[0] > 1e50)  # instance 0 was never seeded
    with pytest.raises(F.FimError):
        b.seed_batch([sets[0]], 0.0, insts=[5])  # instance out of range
    with pytest.raises(F.FimError):
        b.seed_batch([np.array([m.nn], np.int64)], 0.0)  # vertex out of range
    a.close()
    b.close()


def test_stream_maps_equal_lockstep_maps():
    # continuous batching: 10 solves through 3 instances, recycled as they
    # converge; every map equals the same solve run on its own, bit for bit
    from fimpy.stream import stream_solve
    cfg = cases.scaled(cases.SLAB, 8)
    m, om = cases.box(cfg, seed=19)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=19)
    sets = [cases.pacing_sites(cfg, k % 3 + 1, seed=70 + k) for k in range(10)]
    got = {}
    s = F.Solver(p, n_instances=3, window=200.0, local_iters=4)

    def keep(j, inst):
        got[j] = s.download(inst=inst)[0]

    stream_solve(s, sets, on_map=keep)
    s.close()
    assert sorted(got) == list(range(10))
    for j in (0, 4, 9):
        one = F.Solver(p, n_instances=1, window=200.0, local_iters=4)
        one.seed(sets[j], 0.0)
        one.run()
        ref = one.download()[0]
        one.close()
        assert np.array_equal(got[j], ref), j
    ref, _ = O.fim_solve(om, sets[9], variant=O.LOCAL_EXACT)
    _check_fp32(got[9], ref)
    # an instance with queued work cannot be recycled
    s = F.Solver(p, n_instances=2, window=200.0)
    s.seed(sets[0], 0.0, inst=0)
    with pytest.raises(F.FimError):
        s.recycle([0])
    s.close()


def test_converge_deterministic_and_batched():
    cfg = cases.scaled(cases.SLAB, 8)
    m, om = cases.box(cfg, seed=7)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=7)
    srcs = [cases.pacing_sites(cfg, k + 1, seed=100 + k) for k in range(4)]
    singles = [_converge(m, p, s)[0] for s in srcs]
    again = _converge(m, p, srcs[0])[0]
    assert np.array_equal(singles[0], again)  # run-to-run bitwise
    b = F.Solver(p, n_instances=4)
    for i, s in enumerate(srcs):
        b.seed(s, 0.0, inst=i)
    b.run()
    for i in range(4):
        phi, _ = b.download(inst=i)
        assert np.array_equal(phi, singles[i])  # batching changes nothing
    b.close()


@pytest.mark.parametrize("unit", [1, 3, 16])
@pytest.mark.parametrize("grid", [0, 8, 13])
def test_converge_work_units_and_grid_change_nothing(unit, grid):
    """work units of 1..16 instances, dealt to grids of 8, 13 or all
    workgroups, give bitwise the same maps (12 instances, uneven buckets)"""
    cfg = cases.scaled(cases.SLAB, 8)
    m, _ = cases.box(cfg, seed=9)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=9)
    srcs = [cases.pacing_sites(cfg, 1 + k % 3, seed=200 + k) for k in range(12)]

    def solve(**kw):
        s = F.Solver(p, n_instances=12, window=500.0, **kw)
        for i, src in enumerate(srcs):
            s.seed(src, 0.0, inst=i)
        s.run()
        out = [s.download(inst=i)[0] for i in range(12)]
        s.close()
        return out

    base = solve(unit=4)
    for a, b in zip(base, solve(unit=unit, grid=grid)):
        assert np.array_equal(a, b)


def test_run_multi_contexts_change_nothing():
    """12 instances at their own windows, solved by one context and by three
    contexts of 5 + 4 + 3 instances run together (fim_run_multi, one host
    thread feeding three streams): bitwise the same maps, so the oracle pins
    of the single context carry over"""
    cfg = cases.scaled(cases.SLAB, 8)
    m, _ = cases.box(cfg, seed=9)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=9)
    srcs = [cases.pacing_sites(cfg, 1 + k % 3, seed=300 + k) for k in range(12)]
    wins = [250.0 + 75.0 * k for k in range(12)]
    one = F.Solver(p, n_instances=12, window=500.0)
    for i, src in enumerate(srcs):
        one.set_window(wins[i], inst=i)
        one.seed(src, 0.0, inst=i)
    one.run()
    base = [one.download(inst=i)[0] for i in range(12)]
    iters_one = one.stats()["global_iters"]
    one.close()
    g = F.SolverGroup([F.Solver(p, n_instances=n, window=500.0) for n in (5, 4, 3)])
    for ctx, a, b in g.parts(12):
        for i in range(a, b):
            ctx.set_window(wins[i], inst=i - a)
            ctx.seed(srcs[i], 0.0, inst=i - a)
    its = F.run_many(g.solvers)
    assert len(its) == 3 and max(its) <= iters_one and all(k > 0 for k in its)
    for i in range(12):
        assert np.array_equal(g.download(inst=i)[0], base[i])
    # a second run of the same group after a reset gives the same maps again
    for ctx, a, b in g.parts(12):
        ctx.reset()
        for i in range(a, b):
            ctx.seed(srcs[i], 0.0, inst=i - a)
    g.run()
    for i in range(12):
        assert np.array_equal(g.download(inst=i)[0], base[i])
    g.close()
    # profile mode (direct launches, every compute launch event-timed)
    g = F.SolverGroup([F.Solver(p, n_instances=n, window=500.0, profile=True) for n in (6, 6)])
    for ctx, a, b in g.parts(12):
        for i in range(a, b):
            ctx.set_window(wins[i], inst=i - a)
            ctx.seed(srcs[i], 0.0, inst=i - a)
    g.run()
    st = g.stats()
    assert st["compute_launches"] > 0 and st["compute_ms"] > 0.0
    for i in range(12):
        assert np.array_equal(g.download(inst=i)[0], base[i])
    g.close()


def test_run_multi_rejects_bad_context_lists():
    cfg = cases.scaled(cases.SLAB, 8)
    m, _ = cases.box(cfg, seed=9)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=9)
    a = F.Solver(p, n_instances=1)
    c = F.Solver(p, mode=F.MODE_COMPAT)
    with pytest.raises(F.FimError, match="listed twice"):
        F.run_many([a, a])
    with pytest.raises(F.FimError, match="converge mode"):
        F.run_many([a, c])
    with pytest.raises(F.FimError, match="max_iters"):
        F.run_many([a], max_iters=0)
    # an iteration cap below what the solve needs ends every context's run
    # and reports FIM_ENOCONV; a later run finishes the maps
    b = F.Solver(p, n_instances=1, window=500.0)
    b.seed(cases.pacing_sites(cfg, 1, seed=301), 0.0)
    with pytest.raises(F.FimError, match="iteration cap"):
        F.run_many([b], max_iters=1)
    assert b.stats()["global_iters"] >= 1
    F.run_many([b])
    b.close()
    a.close()
    c.close()


@pytest.mark.parametrize("staged", [False, True])
@pytest.mark.parametrize("window", [300.0, 1000.0, 4000.0])
def test_converge_time_window_same_fixed_point(window, staged):
    """the per-iteration time window changes the processing order only"""
    cfg = cases.scaled(cases.SLAB, 4)
    m, om = cases.box(cfg, seed=2)
    p = F.Pack(m, F.PART_RCB, max_nn=512, seed=2)
    src = cases.pacing_sites(cfg, 5, seed=2)
    phi, st, it, res, stats = _converge(m, p, src, window=window, staged=staged)
    ref, _ = O.fim_solve(om, src, variant=O.LOCAL_EXACT)
    _check_fp32(phi, ref)
    assert res < 1e-5


def test_per_instance_windows_same_fixed_point():
    """fim_set_window: four instances of one batch with windows 0 (plain FIM
    order), 0.5, 1 and 4 ms reach the oracle's fixed point, and each equals
    the same solve run alone with that window as the context's, bit for bit"""
    cfg = cases.scaled(cases.SLAB, 4)
    m, om = cases.box(cfg, seed=2)
    p = F.Pack(m, F.PART_RCB, max_nn=512, seed=2)
    src = cases.pacing_sites(cfg, 5, seed=2)
    wins = (0.0, 500.0, 1000.0, 4000.0)
    s = F.Solver(p, n_instances=len(wins), window=1000.0)
    for i, w in enumerate(wins):
        s.set_window(w, inst=i)
        s.seed(src, 0.0, inst=i)
    s.run()
    maps = [s.download(inst=i)[0] for i in range(len(wins))]
    with pytest.raises(F.FimError):
        s.set_window(-1.0, inst=0)
    s.close()
    ref, _ = O.fim_solve(om, src, variant=O.LOCAL_EXACT)
    for i, w in enumerate(wins):
        _check_fp32(maps[i], ref)
        one, *_ = _converge(m, p, src, window=w)
        assert np.array_equal(one, maps[i]), w


def test_warm_start_from_upper_bound_reaches_same_fixed_point():
    """init_field with every vertex active: an upper bound (oracle + 5 %)
    converges to the cold-start fixed point"""
    cfg = cases.scaled(cases.SLAB, 8)
    m, om = cases.box(cfg, seed=9)
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=9)
    src = cases.pacing_sites(cfg, 2, seed=9)
    cold, *_ = _converge(m, p, src)
    ref, _ = O.fim_solve(om, src, variant=O.LOCAL_EXACT)
    s = F.Solver(p)
    s.seed(src, 0.0)
    s.init_field((ref * 1.05 + 100.0).astype(np.float32))
    s.seed(src, 0.0)
    s.run()
    warm, _ = s.download()
    s.close()
    _check_fp32(warm, ref)
    assert np.max(np.abs(warm - cold)) <= 1e-4 * ref.max()


def test_regional_scaling_equals_scaled_mesh():
    """one instance with all four regional factors s equals the mesh with
    both speeds multiplied by s (times divide by s)"""
    cfg = cases.scaled(cases.CUBE, 4)
    m, om = cases.box(cfg, seed=4)
    p = F.Pack(m, F.PART_RCB, max_nn=128, seed=4)
    s = F.Solver(p, n_instances=2)
    s.seed([0], 0.0, inst=0)
    s.seed([0], 0.0, inst=1)
    s.set_scaling(np.full(4, 1.25, np.float32), inst=1)
    s.run()
    a, _ = s.download(inst=0)
    b, _ = s.download(inst=1)
    s.close()
    fin = a < 1e50
    assert np.allclose(b[fin] * 1.25, a[fin], rtol=2e-6, atol=1e-3)


def test_multi_mesh_batch_with_regional_scalings():
    """one batch over a union of four meshes (cube + three slabs, reduced),
    each instance paced on one component with four quadrant speed factors;
    each map equals the oracle's fixed point on the union mesh with every
    element's speeds multiplied by its region's factor"""
    from fimpy import workloads as W
    parts = [(n, cases.scaled(c, 3)) for n, c in W.MULTI]
    m, offs, cfgs = W.multi_mesh(parts, 7)
    solves = W.multi_solves(cfgs, offs, 8, 11)
    p = F.Pack(m, F.PART_RCB, max_nn=512, seed=7)
    s = F.Solver(p, n_instances=len(solves), window=1000.0)
    for i, sv in enumerate(solves):
        s.seed(sv["sites"], 0.0, inst=i)
        s.set_scaling(sv["scale"], inst=i)
    s.run()
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()
    reg = np.concatenate([W.quadrant_regions(F.Mesh.gen_box(seed=7 + k, **c))
                          for k, (_, c) in enumerate(parts)])
    bounds = offs + [m.nn]
    for i, sv in enumerate(solves):
        phi, _ = s.download(inst=i)
        om = O.Mesh(xyz, e2n, fib=fib, vel=vel * sv["scale"].astype(np.float64)[reg][:, None])
        ref, _ = O.fim_solve(om, sv["sites"], variant=O.LOCAL_EXACT)
        _check_fp32(phi, ref)
        k = sv["part"]
        reached = np.flatnonzero(ref < 1e50)
        assert reached.min() >= bounds[k] and reached.max() < bounds[k + 1]
    s.close()


def test_scarred_shell_vs_fp64_oracle():
    """prolate shell with twelve slow scars at the 1M-tet resolution the
    large workload is checked at, three pacing sites on the inner surface"""
    from fimpy import workloads as W
    m = W.shell(1, h=850.0)
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()
    om = O.Mesh(xyz, e2n, fib=fib, vel=vel)
    p = F.Pack(m, F.PART_RCB, max_nn=1024, seed=1)
    sites = W.inner_surface_sites(m, 3, 5)
    phi, st, it, res, stats = _converge(m, p, sites, window=2000.0)
    ref, _ = O.fim_solve(om, sites, variant=O.LOCAL_EXACT)
    _check_fp32(phi, ref)


def test_compat_timed_events_two_beats_bit_identical():
    """multi-beat pacing with events fired in the window holding their start
    time (re-entry through tail -> far): GPU compat mode equals the oracle's
    timed restatement bit for bit"""
    cfg = cases.scaled(cases.CUBE, 2)
    m, om = cases.box(cfg, seed=3)
    p = F.Pack(m, F.PART_RCB, max_nn=128, seed=3)
    e2s, _ = p.partition()
    om = O.Mesh(om.xyz, om.e2n, fib=om.fib, vel=om.vel, atags=e2s.astype(np.int32))
    dt, apd, steps = 10e3, 50e3, 18
    stims = [(1, 0.0, [0]), (1, 120e3, [0]), (1, 125e3, [1330])]
    s = F.Solver(p, mode=F.MODE_COMPAT, local=F.LOCAL_REF, apd=apd)
    fw = m.front_width()
    s.reset()
    cur = 0.0
    for _ in range(steps):
        ps = cur
        cur += dt
        for active, tstart, vtx in stims:
            if active and ps <= tstart < cur:
                s.seed(vtx, tstart)
        s.step(ps, cur, dt, fw)
    phi, st = s.download()
    s.close()
    ophi, ost, _ = O.compat_solve(om, stims, dt, steps, apd=apd, front_width=fw, timed=True)
    assert np.array_equal(st, ost)
    assert np.array_equal(phi, ophi)


def test_converge_edge_cases():
    """a seed time shifts the map; an unreached component stays unreached;
    an iteration cap reports FIM_ENOCONV and a further run finishes the
    same map"""
    from fimpy import workloads as W
    cfg = cases.scaled(cases.SLAB, 8)
    a = F.Mesh.gen_box(seed=1, **cfg)
    b = F.Mesh.gen_box(seed=2, **cases.scaled(cases.CUBE, 4))
    u = F.Mesh.concat([a, b])
    p = F.Pack(u, F.PART_RCB, max_nn=256, seed=1)
    src = cases.pacing_sites(cfg, 2, seed=3)
    s = F.Solver(p, n_instances=3, window=1000.0)
    s.seed(src, 0.0, inst=0)
    s.seed(src, 5000.0, inst=1)
    s.run()
    p0, _ = s.download(inst=0)
    p1, _ = s.download(inst=1)
    p2, st2 = s.download(inst=2)  # never seeded
    fin = p0 < 1e50
    assert np.array_equal(np.flatnonzero(fin), np.arange(a.nn))  # only the first component
    assert np.allclose(p1[fin], p0[fin] + 5000.0, rtol=0, atol=1e-4 * p0[fin].max())
    assert np.all(p2 >= 1e50) and np.all(st2 == F.FAR)
    # iteration cap, then completion
    s.reset()
    s.seed(src, 0.0, inst=0)
    with pytest.raises(F.FimError, match="FIM_ENOCONV"):
        s.run(max_iters=3)
    s.run()
    again, _ = s.download(inst=0)
    assert np.array_equal(again, p0)
    s.close()
