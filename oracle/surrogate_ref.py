"""FP64 reference of the surrogate MLP (test infrastructure only).

The surrogate has no counterpart in the reference project; this module is
the CPU oracle its gfx950 kernels are checked against: the same network
(inputs standardised and zero-padded to 32, hidden layers with SiLU, a
linear output layer padded to 32), the same flat parameter layout per
ensemble member (gs_surrogate_layout), and the same loss
L = sum w (y - t)^2 / sum w with its exact gradient.  Its backward pass is
pinned by central finite differences in tests/test_oracle_surrogate.py.

Only tests/ may import this module.
"""
import numpy as np

PAD = 32


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def bf16_round(a):
    """Round float values to bfloat16 (nearest, ties to even), as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, layout):
    """[(W [out][in], b [out]) ...] of one member's flat parameter block."""
    out = []
    for woff, boff, o, i in layout:
        W = params[woff:woff + o * i].reshape(o, i)
        b = params[boff:boff + o]
        out.append((W, b))
    return out


def pad_inputs(x, mean, sd):
    n, k = x.shape
    xp = np.zeros((n, PAD))
    xp[:, :k] = (x - mean) / sd
    return xp


def forward(layers, xp, emulate_bf16=False):
    """Outputs [n][PAD] and the cache the backward pass needs."""
    r = bf16_round if emulate_bf16 else (lambda v: v)
    acts = [r(xp)]
    zs = []
    for li, (W, b) in enumerate(layers):
        z = acts[-1] @ r(W).T + b
        if li == len(layers) - 1:
            return z, (acts, zs)
        zs.append(z)
        acts.append(r(silu(z)))


def loss_and_grads(layers, xp, y, w, emulate_bf16=False):
    """Loss and [(dW, db) ...] for targets y, weights w ([n][PAD], zero padded)."""
    out, (acts, zs) = forward(layers, xp, emulate_bf16)
    r = bf16_round if emulate_bf16 else (lambda v: v)
    ws = w.sum()
    d = out - y
    loss = float((w * d * d).sum() / ws)
    dz = 2.0 * w * d / ws
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        dzr = r(dz)
        grads[li] = (dzr.T @ acts[li], dz.sum(axis=0))
        if li == 0:
            break
        dz = (dzr @ r(W)) * dsilu(zs[li - 1])
    return loss, grads


def flatten(grads, layout, size):
    g = np.zeros(size)
    for (woff, boff, o, i), (dW, db) in zip(layout, grads):
        g[woff:woff + o * i] = dW.reshape(-1)
        g[boff:boff + o] = db
    return g
