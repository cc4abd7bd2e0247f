"""The learned-initial-iterate network on the device against its CPU
restatement (oracle/guess_oracle.py), and its effect on a flow run.

bf16 tolerances: the device and the restatement round inputs, weights,
activations and gradients to bf16 at the same points and differ only in
the order of the fp32 (device) / float64 (restatement) sums, which can flip
an occasional bf16 rounding; the bounds below are a few bf16 ulps of the
largest output.
"""
from __future__ import annotations

import numpy as np
import pytest

import guess_oracle as GO
from pycathy_wrapper_amd import configs as CF
from pycathy_wrapper_amd import guess as GU

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    p = GO.init_params(seed)
    for (ow, ob), co in zip(GO.offsets(), GO.COUT):
        p[ob:ob + co] = 0.1 * rng.standard_normal(co)
    return p.astype(np.float32)


def _inputs(seed, ncol, L):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((ncol, GU.padded(L), GU.NFEAT)).astype(np.float32)
    x[:, L:] = 0
    return GU.to_bf16_bits(x)


@pytest.mark.parametrize("L", [11, 31])
def test_forward_matches_oracle(L):
    p = _params(1)
    x = _inputs(2, 37, L)
    net = GU.GuessNet(L, p)
    try:
        y = net.forward(x)
    finally:
        net.close()
    yo, _ = GO.forward(p, GO.from_bits(x), L)
    assert np.all(y[:, L:] == 0)
    d = np.abs(y - yo)
    top = np.max(np.abs(yo))
    assert np.max(d) <= 1e-2 * top and np.mean(d) <= 1e-3 * top, (np.max(d), top)


@pytest.mark.parametrize("L", [11, 31])
def test_gradient_matches_oracle(L):
    p = _params(3)
    x = _inputs(4, 29, L)
    rng = np.random.default_rng(5)
    t = rng.standard_normal((29, GU.padded(L))).astype(np.float32)
    net = GU.GuessNet(L, p)
    try:
        loss, g = net.gradient(x, t)
        loss2, g2 = net.gradient(x, t)
    finally:
        net.close()
    assert loss == loss2 and np.array_equal(g, g2)  # deterministic reductions
    lo, go = GO.backward(p, GO.from_bits(x), t, L)
    assert loss == pytest.approx(lo, rel=1e-3)
    for li, ((ow, ob), co) in enumerate(zip(GO.offsets(), GO.COUT)):
        for a, b in ((g[ow:ob], go[ow:ob]), (g[ob:ob + co], go[ob:ob + co])):
            assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-12, li


def test_training_fits_a_column_operator():
    """Adam on a target the network can represent (a smooth nonlinear
    function of neighbouring levels): the loss falls tenfold, and the run is
    reproducible."""
    L, n = 11, 4096
    rng = np.random.default_rng(6)
    xf = rng.standard_normal((n, 16, GU.NFEAT)).astype(np.float32)
    xf[:, L:] = 0
    x = GU.to_bf16_bits(xf)
    xr = GO.from_bits(x)
    up = np.zeros_like(xr[..., 1])
    up[:, 1:L] = xr[:, 0:L - 1, 1]
    t = np.tanh(xr[..., 0] + 0.5 * up) + 0.3 * xr[..., 2] * xr[..., 3]
    t[:, L:] = 0
    runs = []
    for _ in range(2):
        net = GU.GuessNet(L, GU.init_params(7))
        try:
            net.load(x, t)
            losses = net.train(600, 256, 3e-3, seed=8)
            runs.append((losses, net.get()[0]))
        finally:
            net.close()
    losses = runs[0][0]
    assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:5]), (losses[:5], losses[-20:])
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


def test_untrained_guess_is_the_plain_run():
    """The initial network predicts zero (output layer zero); with the
    extrapolation weight beta = 0 a run with it attached is bit-identical
    to the plain run: the guess plumbing changes nothing but the initial
    iterate."""
    from pycathy_wrapper_amd import gpu as G
    pr = CF.hillslope_small()
    out = []
    for attach in (False, True):
        f = G.FlowGPU(pr, G.ORDER_NATURAL)
        net = GU.GuessNet(pr.mesh.nstr + 1, scale=GU.default_scale(beta=0.0)) if attach else None
        try:
            if attach:
                f.guess(net, GU.static_features(pr))
            st = [f.advance()[1] for _ in range(25)]
            out.append((st, f.psi()))
        finally:
            f.close()
            if net is not None:
                net.close()
    (s0, p0), (s1, p1) = out
    assert np.array_equal(p0, p1)
    for a, b in zip(s0, s1):
        assert a == b


def test_trained_guess_against_plain_at_identical_steps():
    """The learned initial iterate end to end: training pairs from 24
    perturbed-parameter members of the small hillslope, a network trained
    on the device, then a held-out member replayed at the plain run's step
    sizes with the plain, the extrapolated (untrained network) and the
    learned initial iterate.  The learned one must converge every step
    without back-steps, with fewer Picard and PCG iterations than the plain
    one and than the extrapolation, and with psi within 2e-6 (relative
    L-inf) of the plain run at every step."""
    from pycathy_wrapper_amd import gpu as G
    L = 11
    raws, tgts = [], []
    for sd in range(24):
        r, t, _ = GU.collect(CF.hillslope_small_member(sd), G.ORDER_COLUMN3)
        raws.append(r)
        tgts.append(t)
    net, losses = GU.train_guess(np.concatenate(raws), np.concatenate(tgts), L, steps=8000, batch=2048, lr=1e-3)
    extrap = GU.GuessNet(L, GU.init_params(0), net.get()[1])
    try:
        assert np.mean(losses[-200:]) < 0.5 * np.mean(losses[:10])
        res = GU.compare(CF.hillslope_small_member(100), dict(extrap=extrap, learned=net), G.ORDER_COLUMN3)
    finally:
        net.close()
        extrap.close()
    P, X, Lr = res["plain"], res["extrap"], res["learned"]
    print("picard plain/extrap/learned", P["picard"].sum(), X["picard"].sum(), Lr["picard"].sum(),
          "pcg", P["pcg"].sum(), X["pcg"].sum(), Lr["pcg"].sum(), "psi rel", Lr["psi_rel"].max())
    assert Lr["back"].sum() == 0
    assert Lr["picard"].sum() < X["picard"].sum() < P["picard"].sum()
    assert Lr["pcg"].sum() < P["pcg"].sum()
    assert Lr["psi_rel"].max() <= 2e-6
