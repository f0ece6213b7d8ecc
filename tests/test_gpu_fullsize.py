"""Parity at the full sizes of the measured workloads, through the C ABI.

  anisotropic slab, 160 x 160 x 13 hexes x 6 = 1,996,800 tets: a batch of
    64 maps at the benchmark's settings, two of them against the FP64 fixed
    point of the oracle (run in host threads);
  256 solves over the union of the 48k / 404k / 996k / 2M-tet meshes with
    1-3 pacing sites and four quadrant speed factors U[0.5, 1.5] each: one
    instance per mesh against the oracle on that mesh with every element's
    speeds multiplied by its region's factor, and every instance confined to
    its own component;
  the surrogate on the slab: trained on 4,096 batched-FIM maps, then the 512
    held-out solves as one surrogate-scheduled batch -- every map within
    tolerance of the cold batch's, two against the FP64 oracle, and fewer
    global iterations than the cold batch (what the network gains).

Tolerance: max |dT| <= 1e-4 * max(T) (FP32 against FP64), which on these
meshes is well inside the 0.1 ms the throughput figure is quoted at; the
slab maps are also held to 100 us directly.  The oracle needs 10-30 s per
2M-tet map on one host thread."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import fimpy as F
import pyoracle as O
from fimpy import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def _check(phi, ref, abs_tol=None):
    fin = ref < 1e50
    assert np.array_equal(fin, phi < 1e50)
    err = float(np.max(np.abs(phi[fin] - ref[fin])))
    tol = 1e-4 * float(ref[fin].max())
    assert err <= tol, (err, tol)
    if abs_tol is not None:
        assert err <= abs_tol, err
    return err


def test_full_slab_batch_vs_fp64_oracle():
    cfg = W.SLAB
    m = W.box(cfg, 2)
    assert (m.nn, m.ne) == (362894, 1996800)
    p = F.Pack(m, F.PART_RCB, max_nn=2816, seed=2)
    sets = [W.pacing_sites(cfg, 5, 2 * 100003 + i) for i in range(64)]
    s = F.Solver(p, n_instances=64, block=512, unit=2, local_iters=4, window=2000.0)
    s.seed_batch(sets, 0.0)
    s.run(max_iters=1_000_000)
    got = {i: s.download(inst=i)[0] for i in (0, 63)}
    res = [s.residual(inst=i) for i in (0, 63)]
    s.close()
    assert max(res) < 1e-5
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()
    om = O.Mesh(xyz, e2n, fib=fib, vel=vel)
    with ThreadPoolExecutor(2) as ex:
        refs = dict(zip((0, 63), ex.map(lambda i: O.fim_solve(om, sets[i], variant=O.LOCAL_EXACT)[0],
                                        (0, 63))))
    for i in (0, 63):
        _check(got[i], refs[i], abs_tol=100.0)


def test_four_mesh_union_256_solves_vs_fp64_oracle():
    m, offs, cfgs = W.multi_mesh(W.MULTI, 7)
    solves = W.multi_solves(cfgs, offs, 256, 11)
    assert m.ne == sum(6 * c["nx"] * c["ny"] * c["nz"] for c in cfgs)
    p = F.Pack(m, F.PART_RCB, max_nn=2816, seed=7)
    s = F.Solver(p, n_instances=len(solves), window=2000.0, local_iters=4)
    for i, sv in enumerate(solves):
        s.set_scaling(sv["scale"], inst=i)
    s.seed_batch([sv["sites"] for sv in solves], 0.0)
    s.run(max_iters=1_000_000)
    bounds = offs + [m.nn]
    # every instance stays in its own component
    for i, sv in enumerate(solves):
        phi, _ = s.download(inst=i)
        reached = np.flatnonzero(phi < 1e50)
        k = sv["part"]
        assert reached.min() >= bounds[k] and reached.max() < bounds[k + 1], i
    # one instance per mesh against the oracle on that mesh alone
    picks = [next(i for i, sv in enumerate(solves) if sv["part"] == k) for k in range(len(cfgs))]
    got = {i: s.download(inst=i)[0] for i in picks}
    s.close()

    def oracle(i):
        sv = solves[i]
        k = sv["part"]
        part = W.box(cfgs[k], 7 + k)
        xyz, e2n, _ = part.arrays()
        fib, vel = part.material()
        reg = W.quadrant_regions(part)
        om = O.Mesh(xyz, e2n, fib=fib, vel=vel * sv["scale"].astype(np.float64)[reg][:, None])
        return O.fim_solve(om, np.asarray(sv["sites"]) - offs[k], variant=O.LOCAL_EXACT)[0]

    with ThreadPoolExecutor(len(picks)) as ex:
        refs = dict(zip(picks, ex.map(oracle, picks)))
    for i in picks:
        k = solves[i]["part"]
        _check(got[i][bounds[k]:bounds[k + 1]], refs[i])


def test_surrogate_scheduled_full_slab_512_held_out():
    from fimpy import warmstart as WS
    cfg = W.SLAB
    m = W.box(cfg, 2)
    reg = W.quadrant_regions(m)
    m.set_regions(reg)
    p = F.Pack(m, F.PART_RCB, max_nn=2816, seed=2)
    train = WS.draw_solves(cfg, 4096, 1003)
    test = WS.draw_solves(cfg, 512, 3005)
    nn = F.Surrogate(m, seed=2, max_maps=4096, max_batch=65536)
    WS.build_dataset(nn, p, train, batch=512, window=2000.0, local_iters=4, block=512)
    assert nn.dataset_size() == 4096
    l0 = nn.train(1, batch=65536, seed=1)
    l1 = nn.train(4000, batch=65536, seed=2)
    assert l1 < 0.05 * l0, (l0, l1)
    s = F.Solver(p, n_instances=512, window=2000.0, unit=2, local_iters=4, block=512)
    cold = WS.cold_batch(s, test)
    cold_maps = [s.download(inst=i)[0] for i in range(512)]
    sample = WS.window_sample(m)
    st, _, _ = WS.scheduled_batch(nn, s, test, sample, window=2000.0, quantile=0.75)
    # the network's gain: the batch needs fewer global iterations
    assert st["global_iters"] < cold["global_iters"], (st["global_iters"], cold["global_iters"])
    for i in range(512):  # windows change the order only: the cold maps within tolerance
        _check(s.download(inst=i)[0], cold_maps[i])
    got = {i: s.download(inst=i)[0] for i in (0, 511)}
    # the surrogate alone is a usable but coarse map
    pred = nn.predict_map(*test[0])
    assert np.mean(np.abs(pred - cold_maps[0])) < 0.05 * cold_maps[0].max()
    s.close()
    nn.close()
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()

    def oracle(i):
        sites, sc = test[i]
        om = O.Mesh(xyz, e2n, fib=fib, vel=vel * sc.astype(np.float64)[reg][:, None])
        return O.fim_solve(om, sites, variant=O.LOCAL_EXACT)[0]

    with ThreadPoolExecutor(2) as ex:
        refs = dict(zip((0, 511), ex.map(oracle, (0, 511))))
    for i in (0, 511):
        _check(got[i], refs[i], abs_tol=100.0)

