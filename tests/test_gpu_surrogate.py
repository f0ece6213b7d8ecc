"""GPU surrogate (BF16 MFMA MLP, include/cufim.h fim_nn_*) against its CPU
oracle (oracle/mlp_oracle.py), and the warm start it feeds.

Tolerances: the oracle rounds to BF16 where the GPU does, so forward values
agree to FP32-accumulation level (relative 1e-2 of the output spread) and
gradients per parameter block to 2e-2 relative L2 with cosine > 0.999.
Warm-started maps must equal cold maps within 1e-4 x max(T)."""
import numpy as np
import pytest

import cases
import fimpy as F
import mlp_oracle as M

pytestmark = pytest.mark.gpu

TS = 1e5  # us per output unit (library default)


def _blocks():
    """(name, slice) of each parameter block in logical order"""
    out, q = [], 0
    for layer in range(M.NL):
        fan = M.NIN if layer == 0 else M.HID
        out.append((f"W{layer}", slice(q, q + M.HID * fan)))
        q += M.HID * fan
        out.append((f"b{layer}", slice(q, q + M.HID)))
        q += M.HID
    out.append(("w_out", slice(q, q + M.HID)))
    out.append(("b_out", slice(q + M.HID, q + M.HID + 1)))
    return out


@pytest.fixture(scope="module")
def small_slab():
    cfg = cases.scaled(cases.SLAB, 4)
    m = F.Mesh.gen_box(seed=3, **cfg)
    from fimpy import workloads as W
    m.set_regions(W.quadrant_regions(m))
    return cfg, m


def test_forward_matches_oracle(small_slab):
    _, m = small_slab
    nn = F.Surrogate(m, seed=1)
    p = M.init_params(5)
    nn.set_params(p)
    assert np.array_equal(nn.params(), p)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (200, M.NIN)).astype(np.float32)  # 200: padded rows are ignored
    y = nn.predict(x)
    ref, _ = M.forward(p, x)
    ref = ref * TS
    assert np.max(np.abs(y - ref)) <= 1e-2 * (np.ptp(ref) + 1e-3 * TS), np.max(np.abs(y - ref))
    nn.close()


@pytest.mark.parametrize("b", [256, 2112])
def test_gradient_matches_oracle(small_slab, b):
    """b = 256: four row tiles, eight weight-gradient spans of one 32-row
    step; b = 2112: 33 tiles, 66 steps over 40 spans of one or two steps."""
    _, m = small_slab
    nn = F.Surrogate(m, seed=1)
    p = M.init_params(7)
    nn.set_params(p)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (b, M.NIN)).astype(np.float32)
    t = rng.uniform(0.0, 2.0 * TS, b).astype(np.float32)
    loss, g = nn.gradient(x, t)
    rl, rg = M.loss_and_grad(p, x, t.astype(np.float64) / TS)
    assert abs(loss - rl) <= 1e-2 * rl
    for name, sl in _blocks():
        a, b = g[sl].astype(np.float64), rg[sl]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        assert err < 2e-2 and cos > 0.999, (name, err, cos)
    # the parameters are unchanged by a gradient evaluation
    assert np.array_equal(nn.params(), p)
    nn.close()


def test_adam_steps_match_oracle(small_slab):
    """k_adam against the oracle's Adam: three steps with known gradients
    through the C ABI; FP32 masters and both moments agree with the FP64
    update (tolerance: 4 FP32 ulps of the parameter plus 1e-5 of the step
    size, far below the lr-sized differences a wrong bias correction or a
    swapped beta gives).  The forward pass afterwards uses the updated BF16
    copies of the hidden-layer weights: it matches the oracle forward with
    the new parameters, and not the one with the old."""
    _, m = small_slab
    lr = 0.05
    nn = F.Surrogate(m, seed=1, lr=lr)
    p0 = M.init_params(9)
    nn.set_params(p0)
    rng = np.random.default_rng(4)
    b1, b2, eps = float(np.float32(0.9)), float(np.float32(0.999)), float(np.float32(1e-8))
    p, mo, vo = p0.astype(np.float64), np.zeros(M.NPARAMS), np.zeros(M.NPARAMS)
    gsum, g2sum = np.zeros(M.NPARAMS), np.zeros(M.NPARAMS)  # scales of the moments' FP32 rounding
    for t in (1, 2, 3):
        g = (rng.normal(0, 1e-2, M.NPARAMS) * rng.choice([1e-3, 1.0, 1e2], M.NPARAMS)).astype(np.float32)
        g[::97] = 0.0
        if t == 2:
            g[:5000] = -g[:5000]
        nn.adam_step(g)
        p, mo, vo = M.adam_step(p, mo, vo, g.astype(np.float64), t, lr, b1, b2, eps)
        gsum += np.abs(g)
        g2sum += g.astype(np.float64) ** 2
    got = nn.params().astype(np.float64)
    gm, gv, steps = nn.moments()
    assert steps == 3
    tol = 4 * np.spacing(np.abs(p).astype(np.float32)).astype(np.float64) + 1e-5 * lr
    assert np.all(np.abs(got - p) <= tol), float(np.max(np.abs(got - p) - tol))
    # moments to FP32 accumulation (a few roundings of the terms summed;
    # m can cancel to near zero after a sign change, so the bound is on the
    # magnitudes that were summed, not on m itself)
    assert np.all(np.abs(gm - mo) <= 1e-6 * gsum + 1e-30), float(np.max(np.abs(gm - mo) - 1e-6 * gsum))
    assert np.all(np.abs(gv - vo) <= 1e-6 * g2sum + 1e-30), float(np.max(np.abs(gv - vo) - 1e-6 * g2sum))
    assert np.max(np.abs(got - p0)) > 0.1 * lr  # the steps moved the parameters
    x = rng.uniform(-1, 1, (256, M.NIN)).astype(np.float32)
    y = nn.predict(x)
    new, _ = M.forward(got.astype(np.float32), x)
    old, _ = M.forward(p0, x)
    err_new = np.max(np.abs(y - new * TS))
    err_old = np.max(np.abs(y - old * TS))
    assert err_new <= 1e-2 * (np.ptp(new * TS) + 1e-3 * TS), err_new
    assert err_old > 5 * err_new, (err_old, err_new)
    nn.close()


def test_features_layout(small_slab):
    cfg, m = small_slab
    nn = F.Surrogate(m, seed=1)
    xyz, _, _ = m.arrays()
    sites = [5, 17]
    x = nn.features([0, 5, 100], sites, [1.0, 0.5, 1.5, 1.2])
    ctr = 0.5 * (xyz.min(0) + xyz.max(0))
    half = 0.5 * np.ptp(xyz, axis=0).max()
    assert np.allclose(x[:, :3], (xyz[[0, 5, 100]] - ctr) / half, atol=1e-6)
    assert np.allclose(x[1, 3:6], x[1, :3])  # vertex 5 is the first site
    assert np.allclose(x[:, 9:12], x[:, 3:6])  # third site repeats the first
    assert np.allclose(x[0, 12:], [0.0, -0.5, 0.5, 0.2], atol=1e-7)
    nn.close()


def _solve_batch(p, solves, **kw):
    s = F.Solver(p, n_instances=len(solves), window=500.0, **kw)
    for i, (sites, sc) in enumerate(solves):
        s.seed(sites, 0.0, inst=i)
        s.set_scaling(sc, inst=i)
    s.run()
    return s


def _draw(cfg, n, seed):
    from fimpy import workloads as W
    g = W.splitmix64(seed)
    out = []
    for _ in range(n):
        k = 1 + next(g) % 3
        sites = W.pacing_sites(cfg, k, next(g))
        sc = np.array([0.5 + (next(g) >> 11) * 2.0 ** -53 for _ in range(4)], dtype=np.float32)
        out.append((sites, sc))
    return out


def test_training_and_warm_start_reach_the_cold_fixed_point(small_slab):
    cfg, m = small_slab
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=3)
    train = _draw(cfg, 1024, 11)
    s = _solve_batch(p, train)
    nn = F.Surrogate(m, seed=2, max_maps=1024, max_batch=4096)
    nn.add_maps(s, 0, [x[0] for x in train], np.stack([x[1] for x in train]))
    assert nn.dataset_size() == 1024
    s.close()
    l0 = nn.train(1, batch=4096, seed=1)
    l1 = nn.train(3000, batch=4096, seed=2)
    assert l1 < 0.02 * l0, (l0, l1)
    # held-out solves: surrogate error, then warm start vs cold
    test = _draw(cfg, 4, 99)
    cold = _solve_batch(p, test)
    warm = F.Solver(p, n_instances=len(test), window=500.0)
    for i, (sites, sc) in enumerate(test):
        ref, _ = cold.download(inst=i)
        pred = nn.predict_map(sites, sc)
        err = np.max(np.abs(pred - ref))
        # a usable surrogate on held-out solves: mean error a few % of max T
        assert np.mean(np.abs(pred - ref)) < 0.1 * ref.max(), (np.mean(np.abs(pred - ref)), err)
        delta = 1.5 * err  # an upper bound by construction for this test
        warm.set_scaling(sc, inst=i)
        nn.warm_start(warm, i, sites, sc, delta, all_active=True)
    warm.run()
    for i, (sites, sc) in enumerate(test):
        ref, _ = cold.download(inst=i)
        got, _ = warm.download(inst=i)
        assert np.max(np.abs(got - ref)) <= 1e-4 * ref.max()
        assert warm.support(inst=i) <= 1e-4
    cold.close()
    warm.close()
    nn.close()


@pytest.mark.parametrize("max_batch", [0, 64])
def test_predicted_maximum_equals_the_map_maximum(small_slab, max_batch):
    """fim_nn_predict_max over a vertex sample equals the largest value of
    fim_nn_predict_map at those vertices (the same rows through the same
    kernels, so bit for bit), for several solves in one call; with the
    smallest row capacity (the mesh's vertices) the solves take several
    forward passes"""
    from fimpy import warmstart as WS
    cfg, m = small_slab
    nn = F.Surrogate(m, seed=3, max_batch=max_batch)
    nn.set_params(M.init_params(11))
    solves = _draw(cfg, 37, 5)
    sample = WS.window_sample(m, k=500, seed=1)
    got = nn.predict_max(solves, sample)
    for j in (0, 17, 36):
        t = nn.predict_map(*solves[j])
        assert got[j] == t[sample].max(), j
    with pytest.raises(F.FimError):
        nn.predict_max(solves, np.array([m.nn], dtype=np.int64))
    nn.close()


def test_support_check_flags_a_field_below_the_fixed_point(small_slab):
    cfg, m = small_slab
    p = F.Pack(m, F.PART_RCB, max_nn=256, seed=3)
    sites = cases.pacing_sites(cfg, 2, seed=4)
    s = F.Solver(p, n_instances=2, window=500.0)
    s.seed(sites, 0.0, inst=0)
    s.run()
    ref, _ = s.download(inst=0)
    assert s.support(inst=0) <= 1e-5
    low = (0.7 * ref).astype(np.float32)  # not an upper bound
    s.init_field(low, inst=1)
    s.seed(sites, 0.0, inst=1)
    s.run()
    assert s.support(inst=1) > 1e-2
    s.close()
