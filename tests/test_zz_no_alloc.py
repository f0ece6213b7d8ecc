"""fim_support_noalloc -- the support check with its per-vertex scratch in a
static device array of the library instead of a hipMalloc per call --
against fim_support through the C ABI: equal results on two calls in a row
(the second on a smaller mesh, so the shared array holds the first call's
larger contents), on a field reached from an upper bound (support within the
convergence threshold) and on one started below the fixed point (flagged);
a mesh above the array's 2^20 vertices is rejected with FIM_ELIMIT."""
import numpy as np
import pytest

import cases
import fimpy as F

pytestmark = pytest.mark.gpu


def _solved(cfg, seed, n_inst=2):
    m = F.Mesh.gen_box(seed=seed, **cfg)
    p = F.Pack(m, F.PART_RCB, max_nn=512, seed=seed)
    s = F.Solver(p, n_instances=n_inst, window=500.0)
    sites = cases.pacing_sites(cfg, 2, seed=seed)
    s.seed(sites, 0.0, inst=0)
    s.run()
    ref, _ = s.download(inst=0)
    s.init_field((0.7 * ref).astype(np.float32), inst=1)  # below the fixed point
    s.seed(sites, 0.0, inst=1)
    s.run()
    return m, p, s


def test_support_noalloc_equals_support():
    big, pb, sb = _solved(cases.scaled(cases.SLAB, 2), 21)     # 81 x 81 x 8 hexes
    small, ps, ss = _solved(cases.scaled(cases.SLAB, 8), 22)   # 20 x 20 x 2 hexes
    assert big.nn > 4 * small.nn
    for s in (sb, ss):  # the second call runs on a smaller mesh
        for inst in (0, 1):
            a = s.support(inst=inst)
            b = s.support_noalloc(inst=inst)
            assert a == b, (inst, a, b)
        assert s.support_noalloc(inst=0) <= 1e-5
        assert s.support_noalloc(inst=1) > 1e-2
    # and back to the larger mesh: nothing of the small call leaks into it
    assert sb.support_noalloc(inst=1) == sb.support(inst=1)
    assert sb.support_noalloc(inst=0) == sb.support(inst=0)
    sb.close()
    ss.close()


def test_support_noalloc_rejects_meshes_above_its_scratch():
    cfg = dict(cases.CUBE, nx=101, ny=101, nz=101)
    m = F.Mesh.gen_box(seed=1, **cfg)
    assert m.nn > (1 << 20)
    p = F.Pack(m, F.PART_RCB, max_nn=2816, seed=1)
    s = F.Solver(p, n_instances=1, window=2000.0)
    with pytest.raises(F.FimError, match="FIM_ELIMIT"):
        s.support_noalloc(inst=0)
    s.close()
