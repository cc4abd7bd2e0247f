"""The CPU restatement of the initial-iterate network (oracle/guess_oracle.py).

Pinned by what the mathematics fixes, so that a wrong tap index, a dropped
padding row or a transposed weight fails: the convolution against an
explicit loop, the backward pass against central finite differences of the
forward pass (float64), and the bfloat16 rounding against known bit
patterns (round to nearest, ties to even).
"""
from __future__ import annotations

import numpy as np
import pytest

import guess_oracle as GO


def test_bf16_rounding_known_bits():
    vals = np.array([1.0, -2.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, 3.0e38, 1e-30, 0.0],
                    dtype=np.float32)
    bits = GO.bf16_bits(vals)
    # 1+2^-8 is halfway between 1 and 1+2^-7: ties to even -> 1.0;
    # 1+3*2^-8 is halfway between 1+2^-7 and 1+2^-6 -> 1+2^-6 (even)
    assert list(bits[:5]) == [0x3F80, 0xC000, 0x3F80, 0x3F82, 0x3F80]
    assert GO.from_bits(bits[5]) == pytest.approx(3.0e38, rel=4e-3)
    assert GO.from_bits(bits[6]) == pytest.approx(1e-30, rel=4e-3)
    assert bits[7] == 0


def test_conv_matches_explicit_loop():
    rng = np.random.default_rng(1)
    B, Lp, L, ci, co = 2, 16, 11, 3, 4
    x = rng.standard_normal((B, Lp, ci))
    W = rng.standard_normal((3, ci, co))
    z = GO.conv(x, W, L)
    ref = np.zeros((B, Lp, co))
    for b in range(B):
        for l in range(Lp):
            for k in range(3):
                lv = l + k - 1
                if 0 <= lv < L:
                    ref[b, l] += x[b, lv] @ W[k]
    assert np.allclose(z, ref, rtol=1e-13, atol=1e-13)


def test_forward_zero_beyond_live_levels_and_receptive_field():
    """Output is zero at padded levels; a change at one input level moves
    outputs only within 5 levels of it (5 layers of kernel 3)."""
    rng = np.random.default_rng(2)
    p = GO.init_params(3)
    L = 20
    x = rng.standard_normal((1, 32, 16))
    y, _ = GO.forward(p, x, L, exact=True)
    assert np.all(y[:, L:] == 0)
    x2 = x.copy()
    x2[0, 10] += 1.0
    y2, _ = GO.forward(p, x2, L, exact=True)
    changed = np.nonzero(np.abs(y2 - y)[0] > 0)[0]
    assert changed.min() >= 5 and changed.max() <= 15 and 10 in changed


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = GO.init_params(5).astype(np.float64) * 0.5
    p += 0.05 * rng.standard_normal(p.size)   # nonzero biases
    L, Lp, B = 9, 16, 3
    x = rng.standard_normal((B, Lp, 16))
    x[:, L:] = 0
    t = rng.standard_normal((B, Lp))
    loss, grad = GO.backward(p, x, t, L, exact=True)
    offs = GO.offsets()
    picks = []
    for (ow, ob), co in zip(offs, GO.COUT):
        picks += list(rng.integers(ow, ob, 6)) + list(ob + rng.integers(0, co, 2))
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = GO.backward(pp, x, t, L, exact=True)[0]
        lm = GO.backward(pm, x, t, L, exact=True)[0]
        fd = (lp - lm) / (2 * h)
        assert grad[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_rounded_backward_close_to_exact():
    """The bf16 rounding points change the gradient by bf16-sized amounts."""
    rng = np.random.default_rng(6)
    p = GO.init_params(7)
    L = 11
    x = GO.bf16(rng.standard_normal((4, 16, 16)))
    x[:, L:] = 0
    t = rng.standard_normal((4, 16))
    _, g1 = GO.backward(p, x, t, L, exact=False)
    _, g2 = GO.backward(GO.bf16(p), x, t, L, exact=True)
    assert np.linalg.norm(g1 - g2) <= 0.1 * np.linalg.norm(g2)
