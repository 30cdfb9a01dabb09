"""CPU oracle of the surrogate network (test infrastructure only): the forward pass, the
mean-square-error loss and its gradient for an MLP with SiLU hidden layers and a linear output
layer, in numpy.

bf16=True rounds every operand of a product to bfloat16 (round to nearest even) at the points
where the GPU kernels round it (weights, layer inputs, the loss gradient and the hidden-layer
gradients before each product), with float32 accumulation; bf16=False computes in float64 and is
the version the finite-difference check pins.
"""
import numpy as np


def to_bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(ws, bs, x, bf16=True):
    """Pre-activations of every layer; the last is the output."""
    dt = np.float32 if bf16 else np.float64
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    h = r(np.asarray(x, dtype=dt))
    zs = []
    for l, (w, b) in enumerate(zip(ws, bs)):
        z = (h.astype(dt) @ r(np.asarray(w, dtype=dt)).astype(dt)).astype(dt) + np.asarray(b, dtype=dt)
        zs.append(z)
        if l < len(ws) - 1:
            h = r(silu(z).astype(dt))
    return zs


def loss_and_grads(ws, bs, x, t, bf16=True):
    """Mean square error over all outputs and its gradient with respect to every weight and bias."""
    dt = np.float32 if bf16 else np.float64
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    zs = forward(ws, bs, x, bf16)
    hs = [r(np.asarray(x, dtype=dt))] + [r(silu(z).astype(dt)) for z in zs[:-1]]
    y = zs[-1]
    n = y.size
    d = (y - np.asarray(t, dtype=dt)).astype(dt)
    loss = float(np.sum(d.astype(np.float64) ** 2) / n)
    dzf = (2.0 / n * d).astype(dt)
    gw, gb = [None] * len(ws), [None] * len(ws)
    for l in range(len(ws) - 1, -1, -1):
        dzb = r(dzf)
        gw[l] = (hs[l].T.astype(dt) @ dzb.astype(dt)).astype(dt)
        gb[l] = dzf.sum(axis=0).astype(dt)
        if l > 0:
            dh = (dzb.astype(dt) @ r(np.asarray(ws[l], dtype=dt)).T.astype(dt)).astype(dt)
            dzf = (dh * silu_grad(zs[l - 1])).astype(dt)
    return loss, gw, gb
