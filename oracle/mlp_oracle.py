"""CPU ORACLE for the order model -- test infrastructure only.

Numpy restatement of the forward pass, softmax cross entropy and backward
pass that the MFMA kernels compute (no counterpart in the reference).  With
``bf16=True`` every operand the GPU rounds to BF16 is rounded the same way
(round to nearest even) and products are accumulated in float64, so the
remaining difference to the GPU is the FP32 accumulation order inside the
MFMA units.  With ``bf16=False`` the same graph runs in float64, which the
tests pin with central finite differences.

Network: x[16] -> relu(x W1 + b1)[64] -> relu(. W2 + b2)[64] -> . W3 + b3[16],
loss = mean over rows of -log softmax(z[:13])[y].
"""
from __future__ import annotations

import numpy as np

NCLASSES = 13
LAYOUT = (("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 16)), ("b3", (16,)))


def unpack(params):
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name]).ravel() for name, _ in LAYOUT])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def forward(params, x, bf=True):
    """Logits [n][16] and the activations the backward pass needs."""
    P = unpack(params)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    f = _f32 if bf else (lambda v: v)
    xb = r(x)
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    h1 = r(np.maximum(f(f(xb @ W1) + P["b1"]), 0.0))
    h2 = r(np.maximum(f(f(h1 @ W2) + P["b2"]), 0.0))
    z = f(f(h2 @ W3) + P["b3"])
    return z, (xb, h1, h2, W1, W2, W3)


def loss_and_grad(params, x, y, bf=True):
    """Mean cross entropy over the rows and its gradient (flat, LAYOUT order)."""
    n = x.shape[0]
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    z, (xb, h1, h2, W1, W2, W3) = forward(params, x, bf)
    zc = z[:, :NCLASSES]
    m = zc.max(axis=1, keepdims=True)
    e = np.exp(zc - m)
    s = e.sum(axis=1, keepdims=True)
    loss = float(np.mean(-(zc[np.arange(n), y] - m[:, 0] - np.log(s[:, 0]))))
    pr = e / s
    dz = np.zeros((n, 16))
    dz[:, :NCLASSES] = pr
    dz[np.arange(n), y] -= 1.0
    dz = r(dz / n)
    g = {"W3": h2.T @ dz, "b3": dz.sum(0)}
    d2 = r((dz @ W3.T) * (h2 > 0))
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = r((d2 @ W2.T) * (h1 > 0))
    g["W1"] = xb.T @ d1
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def numeric_grad(params, x, y, idx, h=1e-6):
    """Central differences of the float64 loss at parameter positions idx."""
    out = np.empty(len(idx))
    p = np.asarray(params, np.float64).copy()
    for k, i in enumerate(idx):
        old = p[i]
        p[i] = old + h
        lp, _ = loss_and_grad(p, x, y, bf=False)
        p[i] = old - h
        lm, _ = loss_and_grad(p, x, y, bf=False)
        p[i] = old
        out[k] = (lp - lm) / (2 * h)
    return out
