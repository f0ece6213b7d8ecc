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


def test_oracle_adam_closed_forms_and_planted_mistakes():
    """Adam pinned by what its definition fixes: the first step moves every
    parameter by lr * g / (|g| + eps) (the bias corrections cancel the
    (1 - beta) factors); a constant gradient keeps m_hat = g, v_hat = g^2 at
    every step; after g, -g the corrected moments are -(1 - b1) / (1 + b1) g
    and g^2.  A dropped bias correction, swapped betas or eps inside the
    square root each fail one of these."""
    rng = np.random.default_rng(3)
    n = 1000
    p0 = rng.normal(0, 0.1, n)
    g = rng.normal(0, 1e-2, n) * rng.choice([1e-3, 1.0, 1e3], n)
    z = np.zeros(n)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8

    def close(a, b):
        return np.allclose(a, b, rtol=1e-12, atol=1e-15)

    def checks(step):
        p1, m1, v1 = step(p0, z, z, g, 1)
        ok = close(p1, p0 - lr * g / (np.abs(g) + eps))
        p, m, v = p0, z, z
        for t in range(1, 6):
            q, m, v = step(p, m, v, g, t)
            ok &= close(q, p - lr * g / (np.abs(g) + eps))
            p = q
        p2, m2, v2 = step(p1, m1, v1, -g, 2)
        mh = -(1 - b1) / (1 + b1) * g
        ok &= close(p2, p1 - lr * mh / (np.abs(g) + eps))
        return bool(ok)

    assert checks(lambda p, m, v, gg, t: M.adam_step(p, m, v, gg, t, lr, b1, b2, eps))

    def no_correction(p, m, v, gg, t):
        m = b1 * m + (1 - b1) * gg
        v = b2 * v + (1 - b2) * gg * gg
        return p - lr * m / (np.sqrt(v) + eps), m, v

    def swapped(p, m, v, gg, t):
        return M.adam_step(p, m, v, gg, t, lr, b2, b1, eps)

    def eps_inside(p, m, v, gg, t):
        m = b1 * m + (1 - b1) * gg
        v = b2 * v + (1 - b2) * gg * gg
        return p - lr * (m / (1 - b1 ** t)) / np.sqrt(v / (1 - b2 ** t) + eps), m, v

    for bad in (no_correction, swapped, eps_inside):
        assert not checks(bad), bad.__name__
