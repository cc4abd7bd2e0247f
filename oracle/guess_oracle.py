"""CPU restatement of the learned-initial-iterate network (test infrastructure).

Only tests/ use this module; the product path runs the HIP kernels
(pycathy_wrapper_amd/csrc/guess.hip).

The per-column 1-D convolutional network: layers 16->64, 64->64 (x3) with
ReLU, then 64->1, all with kernel 3 and zero padding outside the column's
L live levels; tensors are [column][Lp][channel] with Lp >= L.  Parameters
are one float32 vector, per layer W[k][ci][co] then b[co] (40385 values).

Rounding follows the device (``exact=False``): inputs, weights, hidden
activations and backward gradients are rounded to bfloat16 (round to
nearest even) where the device stores them in bf16; biases, the network
output and the loss stay in wider precision.  ``exact=True`` keeps float64
throughout, for finite-difference checks of the backward formulas.
"""
from __future__ import annotations

import numpy as np

CIN = (16, 64, 64, 64, 64)
COUT = (64, 64, 64, 64, 1)
NPARAM = sum(3 * ci * co + co for ci, co in zip(CIN, COUT))


def offsets():
    """(offset of W, offset of b) per layer in the parameter vector."""
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        out.append((o, o + 3 * ci * co))
        o += 3 * ci * co + co
    return out


def bf16_bits(a) -> np.ndarray:
    """float32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(a) -> np.ndarray:
    """Round to the nearest bfloat16, returned as float64."""
    return from_bits(bf16_bits(a)).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)  # float32 parameters convert exactly
    assert p.size == NPARAM
    out = []
    for (ow, ob), ci, co in zip(offsets(), CIN, COUT):
        out.append((p[ow:ob].reshape(3, ci, co), p[ob:ob + co]))
    return out


def _pad(x, L):
    """[B][Lp][c] -> [B][Lp+2][c]: level l at row l+1, zero outside 0..L-1."""
    B, Lp, c = x.shape
    xp = np.zeros((B, Lp + 2, c))
    xp[:, 1:L + 1] = x[:, :L]
    return xp


def conv(x, W, L):
    """z[b, l, co] = sum_k sum_ci x[b, l+k-1, ci] W[k, ci, co] (x zero
    outside levels 0..L-1)."""
    Lp = x.shape[1]
    xp = _pad(x, L)
    return sum(xp[:, k:k + Lp] @ W[k] for k in range(3))


def forward(params, x, L, exact=False):
    """Network output y [B][Lp] (zero at levels >= L) and the inputs of
    every layer (acts[0] = x, acts[i] = output of hidden layer i)."""
    r = (lambda a: np.asarray(a, dtype=np.float64)) if exact else bf16
    layers = unpack(params)
    h = r(x)
    acts = [h]
    for W, b in layers[:4]:
        z = conv(h, r(W), L) + b
        a = np.maximum(z, 0.0)
        a[:, L:] = 0.0
        h = r(a)
        acts.append(h)
    W, b = layers[4]
    y = conv(h, r(W), L)[..., 0] + b[0]
    y[:, L:] = 0.0
    return y, acts


def backward(params, x, t, L, exact=False):
    """Mean squared error over live levels and its gradient (float64),
    with the device's rounding points unless exact."""
    r = (lambda a: np.asarray(a, dtype=np.float64)) if exact else bf16
    layers = unpack(params)
    y, acts = forward(params, x, L, exact)
    B, Lp = y.shape
    n = B * L
    d = y - np.asarray(t, dtype=np.float64)
    d[:, L:] = 0.0
    loss = float(np.sum(d[:, :L] ** 2) / n)
    g = r(2.0 * d / n)[..., None]                  # [B][Lp][1]
    grad = np.zeros(NPARAM)
    offs = offsets()
    for li in range(4, -1, -1):
        W, _ = layers[li]
        Wr = r(W)
        xin = acts[li]
        xp = _pad(xin, L)
        dW = np.stack([np.einsum("bli,blo->io", xp[:, k:k + Lp], g) for k in range(3)])
        db = g[:, :L].sum(axis=(0, 1))
        ow, ob = offs[li]
        grad[ow:ob] = dW.ravel()
        grad[ob:ob + COUT[li]] = db
        if li == 0:
            break
        # input gradient: conv of g with W'[k'] = W[2-k']^T, masked by ReLU
        Wt = np.stack([Wr[2 - k].T for k in range(3)])
        da = conv(g, Wt, L)
        da[xin <= 0.0] = 0.0
        da[:, L:] = 0.0
        g = r(da)
    return loss, grad


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    for (ow, ob), ci, co in zip(offsets(), CIN, COUT):
        p[ow:ob] = rng.standard_normal(3 * ci * co) * np.sqrt(2.0 / (3 * ci))
    return p
