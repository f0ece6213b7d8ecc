"""The surrogate's CPU oracle (oracle/mlp_oracle.py), pinned without a GPU:
its gradient against central finite differences of its own FP64 loss, the
BF16 rounding against known values, and the parameter count of the layout
the C ABI declares."""
import numpy as np

import mlp_oracle as M


def test_parameter_count_matches_header():
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    h = open(os.path.join(root, "include", "cufim.h")).read()
    n = int(re.search(r"#define FIM_NN_PARAMS (\d+)", h).group(1))
    assert n == M.NPARAMS == 333569
    assert 2 * (16 * 256 + 5 * 256 * 256 + 256) == 664064  # forward FLOP per point


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 3.14159265, -2.5e-3], dtype=np.float32)
    r = M.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie -> even (1.0)
    assert r[2] == 1.0 + 2.0 ** -7  # above the tie -> up
    assert abs(r[3] - 3.140625) < 1e-12
    assert abs(r[4] + 0.0025024414) < 1e-9


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1).astype(np.float64)
    x = rng.uniform(-1, 1, (6, M.NIN))
    t = rng.uniform(0, 2, 6)
    loss, g = M.loss_and_grad(p, x, t, bf16_emulation=False)
    idx = np.concatenate([rng.choice(M.NPARAMS - 257, 120, replace=False),
                          np.arange(M.NPARAMS - 257, M.NPARAMS)[::16]])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = M.loss_and_grad(q, x, t, bf16_emulation=False)
        q[i] -= 2 * h
        lm, _ = M.loss_and_grad(q, x, t, bf16_emulation=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_emulation_stays_close_to_exact():
    rng = np.random.default_rng(1)
    p = M.init_params(2)
    x = rng.uniform(-1, 1, (64, M.NIN))
    t = rng.uniform(0, 2, 64)
    l1, g1 = M.loss_and_grad(p, x, t, bf16_emulation=True)
    l0, g0 = M.loss_and_grad(p, x, t, bf16_emulation=False)
    assert abs(l1 - l0) <= 0.05 * l0
    cos = g1 @ g0 / np.linalg.norm(g1) / np.linalg.norm(g0)
    assert cos > 0.99
