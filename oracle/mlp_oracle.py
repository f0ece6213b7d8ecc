"""CPU restatement of the surrogate MLP (cufim-nvlink_amd/csrc/surrogate.hip):
forward pass, squared-error loss and its gradient, and the Adam update, in
FP64 numpy.

TEST INFRASTRUCTURE ONLY: the parity oracle for the GPU network; only
tests/ may import it.  The network has no counterpart in the reference.

With ``bf16=True`` every value the GPU rounds to BF16 is rounded here at the
same place (inputs, hidden-layer weights, pre-activations Z kept for the
backward pass, activations H, backward dZ; the weight gradients take
H = SiLU(Z) of the stored Z, rounded), so the two differ only by
accumulation order (FP32 on the GPU) and the hardware exponential and
reciprocal.  With ``bf16=False`` the pass is exact FP64, and its gradient is
pinned by central finite differences.

Adam (``adam_step``) is Kingma & Ba's Algorithm 1 with bias-corrected
moments; it is pinned by closed forms (first step, constant and alternating
gradients) in tests/test_surrogate_oracle.py.

Parameter order (333,569 values): W0[256 x 16], b0[256], then W_l[256 x 256],
b_l[256] for l = 1..5, then w_out[256], b_out.
"""
from __future__ import annotations

import numpy as np

NIN, HID, NL = 16, 256, 6
NPARAMS = HID * NIN + HID + (NL - 1) * (HID * HID + HID) + HID + 1


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.shape == (NPARAMS,)
    out, q = [], 0
    for layer in range(NL):
        fan = NIN if layer == 0 else HID
        W = p[q:q + HID * fan].reshape(HID, fan)
        q += HID * fan
        b = p[q:q + HID]
        q += HID
        out.append((W, b))
    wo = p[q:q + HID]
    bo = p[q + HID]
    return out, wo, bo


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf16_emulation=True):
    rnd = bf16 if bf16_emulation else (lambda a: np.asarray(a, dtype=np.float64))
    layers, wo, bo = unpack(p)
    h = rnd(x)
    cache = [h]
    for W, b in layers:
        z = h @ rnd(W).T + b
        Z, h = rnd(z), rnd(silu(z))
        cache.append((Z, h))
    y = h @ wo + bo
    return y, cache


def loss_and_grad(p, x, t, bf16_emulation=True):
    """Mean squared error of y against t (already divided by the time scale)
    and its gradient in parameter order."""
    rnd = bf16 if bf16_emulation else (lambda a: np.asarray(a, dtype=np.float64))
    layers, wo, bo = unpack(p)
    y, cache = forward(p, x, bf16_emulation)
    B = x.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / B
    Z5, H5 = cache[NL]
    g_wo = H5.T @ dy
    g_bo = dy.sum()
    dz = rnd(dy[:, None] * wo[None, :] * dsilu(Z5))
    grads = [None] * NL
    for layer in range(NL - 1, -1, -1):
        # the weight gradient recomputes H_{l-1} from the stored Z_{l-1}
        hin = cache[layer] if layer == 0 else rnd(silu(cache[layer][0]))
        gW = dz.T @ hin
        gb = dz.sum(axis=0)
        grads[layer] = (gW, gb)
        if layer > 0:
            W = rnd(layers[layer][0])
            Zp = cache[layer][0]
            dz = rnd((dz @ W) * dsilu(Zp))
    flat = []
    for gW, gb in grads:
        flat += [gW.reshape(-1), gb]
    flat += [g_wo, np.array([g_bo])]
    return loss, np.concatenate(flat)


def adam_step(p, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1): t is the step number
    this update makes (1 for the first).  Returns (p, m, v) after it."""
    p, m, v, g = (np.asarray(a, dtype=np.float64) for a in (p, m, v, g))
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed, scale=1.0):
    """He-uniform hidden layers, small output layer, small random biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for layer in range(NL):
        fan = NIN if layer == 0 else HID
        a = np.sqrt(6.0 / fan) * scale
        parts += [rng.uniform(-a, a, HID * fan), rng.uniform(-0.1, 0.1, HID)]
    parts += [rng.uniform(-0.1, 0.1, HID) * np.sqrt(6.0 / HID), rng.uniform(-0.1, 0.1, 1)]
    return np.concatenate(parts).astype(np.float32)
