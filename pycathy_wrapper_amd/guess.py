"""Learned Picard initial iterate: network handle, column features, training
sets collected from the solver's own runs, and training on the device.

The network (include/cathy_gpu.h, ``cathy_gpu_guess_*``) corrects the
extrapolated increment of each soil column.  At the start of every attempt
after the first step the flow run starts the Picard loop from

    psi_n + beta E + s_col * out * clamp(y, -10, 10),
    E = dt_n (psi_n - psi_{n-1}) / dt_{n-1},   s_col = rms over the column of E + floor,

instead of psi_n; the loop, its tolerance and its convergence test are
those of the plain method, so the prediction can only change how many
Picard (and PCG) iterations a step takes.  The output is relative to the
column's own rate of change, so a step whose head moves by 1e-4 m is
predicted as accurately, relatively, as one that moves by 1 m; the floor
(default 1e-5 m, ten times the Picard tolerance) keeps the correction
below the tolerance where the column barely moves.  Nodes whose psi_n or
guess lies above -band (default 0.1 m: at or near saturation) keep psi_n:
there the Picard linearisation of the storage term changes branch, and a
moved start makes the loop oscillate (measured: back-steps at the plain
run's step sizes without the band, none with it).

Columns: surface node s, levels l = 0 (surface) .. NSTR; CATHY node id
``l * NNOD + s``.  Feature channels per level (16 stored, 9 used):
  0 psi_n  1 theta_n = Sw * porosity  2 log10 Ks  3 alpha = 1/VGPSAT
  4 VGN    5 rainfall rate ATMPOT/ARENOD (m/s, the column's surface value)
  6 log10 dt  7 depth below the surface (m)  8 E (m)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gpu as G
from . import soil as SO

NPARAM = 40385
NFEAT = 16
NLIVE = 9
NSCALE = 2 * NFEAT + 4
CIN = (16, 64, 64, 64, 64)
COUT = (64, 64, 64, 64, 1)


def padded(levels: int) -> int:
    return (levels + 15) // 16 * 16


def to_bf16_bits(a) -> np.ndarray:
    """float32 -> bfloat16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases; the output layer starts at zero so
    an untrained network predicts no change (the plain initial iterate)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    o = 0
    for li, (ci, co) in enumerate(zip(CIN, COUT)):
        nw = 3 * ci * co
        if li < len(CIN) - 1:
            p[o:o + nw] = rng.standard_normal(nw) * np.sqrt(2.0 / (3 * ci))
        o += nw + co
    return p


class GuessNet:
    """Device handle of one network for columns of ``levels`` nodes."""

    def __init__(self, levels: int, params=None, scale=None, seed: int = 0):
        self.levels = int(levels)
        self.Lp = padded(self.levels)
        h = C.c_int64(0)
        npar = C.c_int(0)
        G._check(G.lib().cathy_gpu_guess_create(C.byref(h), G._i(self.levels), C.byref(npar)), "guess_create")
        self.h = h
        assert npar.value == NPARAM
        self.set(init_params(seed) if params is None else params,
                 scale if scale is not None else default_scale())

    def set(self, params=None, scale=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        s = None if scale is None else np.ascontiguousarray(scale, dtype=np.float32)
        if p is not None and p.size != NPARAM:
            raise ValueError(f"expected {NPARAM} parameters")
        if s is not None and s.size != NSCALE:
            raise ValueError("scale holds 16 means, 16 standard deviations, out, floor, beta and band")
        G._check(G.lib().cathy_gpu_guess_set(C.byref(self.h), None if p is None else G._p(p),
                                             None if s is None else G._p(s)), "guess_set")

    def get(self):
        p = np.zeros(NPARAM, dtype=np.float32)
        s = np.zeros(NSCALE, dtype=np.float32)
        G._check(G.lib().cathy_gpu_guess_get(C.byref(self.h), G._p(p), G._p(s)), "guess_get")
        return p, s

    def _x(self, x):
        x = np.ascontiguousarray(x, dtype=np.uint16)
        if x.ndim != 3 or x.shape[1:] != (self.Lp, NFEAT):
            raise ValueError(f"inputs are [ncol][{self.Lp}][{NFEAT}] bf16 bit patterns")
        return x

    def forward(self, x) -> np.ndarray:
        x = self._x(x)
        y = np.zeros((x.shape[0], self.Lp), dtype=np.float32)
        G._check(G.lib().cathy_gpu_guess_forward(C.byref(self.h), G._i(x.shape[0]), G._p(x), G._p(y)),
                 "guess_forward")
        return y

    def gradient(self, x, t):
        x = self._x(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        if t.shape != (x.shape[0], self.Lp):
            raise ValueError("targets are [ncol][Lp]")
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(NPARAM, dtype=np.float32)
        G._check(G.lib().cathy_gpu_guess_gradient(C.byref(self.h), G._i(x.shape[0]), G._p(x), G._p(t),
                                                  G._p(loss), G._p(g)), "guess_gradient")
        return float(loss[0]), g

    def load(self, x, t):
        x = self._x(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        if t.shape != (x.shape[0], self.Lp):
            raise ValueError("targets are [nsamp][Lp]")
        self.nsamp = x.shape[0]
        G._check(G.lib().cathy_gpu_guess_load(C.byref(self.h), C.byref(C.c_int64(x.shape[0])), G._p(x), G._p(t)),
                 "guess_load")

    def train(self, nsteps: int, batch: int, lr: float = 1e-3, seed: int = 0, chunk: int = 500) -> np.ndarray:
        """Adam on random batches of the loaded set (without replacement
        within an epoch); returns the loss of every step."""
        rng = np.random.default_rng(seed)
        n = self.nsamp
        order = np.zeros(0, dtype=np.int64)
        losses = []
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            need = k * batch
            while order.size < need:
                order = np.concatenate([order, rng.permutation(n)])
            idx = np.ascontiguousarray(order[:need], dtype=np.int32)
            order = order[need:]
            out = np.zeros(k, dtype=np.float32)
            G._check(G.lib().cathy_gpu_guess_train(C.byref(self.h), G._i(k), G._i(batch), G._p(idx),
                                                   C.byref(C.c_float(lr)), G._p(out)), "guess_train")
            losses.append(out)
            done += k
        return np.concatenate(losses)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            G.lib().cathy_gpu_guess_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_scale(floor: float = 1e-5, beta: float = 1.0, band: float = 0.1) -> np.ndarray:
    """Identity standardisation, out = 1, the given floor (m), beta and
    saturation band (m; < 0 applies the guess at every node)."""
    s = np.zeros(NSCALE, dtype=np.float32)
    s[NFEAT:2 * NFEAT] = 1.0
    s[2 * NFEAT] = 1.0
    s[2 * NFEAT + 1] = floor
    s[2 * NFEAT + 2] = beta
    s[2 * NFEAT + 3] = band
    return s


def static_features(problem) -> np.ndarray:
    """[N][5] float32 in CATHY node order: log10 Ks (KZNOD), alpha =
    1/VGPSAT, VGN, depth below the column's surface node, porosity."""
    m, s = problem.mesh, problem.soil
    nodep = SO.node_params(m, s)
    kz = SO.kz_nodes(m, s)
    n, nnod = m.n, m.nnod
    surf = np.tile(np.asarray(m.z)[:nnod], n // nnod)
    out = np.zeros((n, 5), dtype=np.float32)
    out[:, 0] = np.log10(kz)
    out[:, 1] = 1.0 / nodep["vgpsat"]
    out[:, 2] = nodep["vgn"]
    out[:, 3] = surf - np.asarray(m.z)
    out[:, 4] = nodep["pnodi"]
    return out


def collect(problem, ordering: int = G.ORDER_COLUMN3, max_steps: int | None = None, col_frac: float = 1.0,
            seed: int = 0):
    """Run the plain solver over ``problem`` and return the training pairs
    of every accepted step: raw features [S][Lp][8] (float32) and head
    increments [S][Lp], S = columns kept per step x steps, plus the run's
    step statistics."""
    rng = np.random.default_rng(seed)
    f = G.FlowGPU(problem, ordering)
    feats, tgts, stats = [], [], []
    try:
        f.guess(None, static_features(problem))
        f.guess_record(True)
        k = 0
        while True:
            rc, st = f.advance()
            raw, tgt = f.guess_sample()
            keep = np.arange(raw.shape[0]) if col_frac >= 1.0 else np.sort(
                rng.choice(raw.shape[0], max(1, int(col_frac * raw.shape[0])), replace=False))
            feats.append(raw[keep])
            tgts.append(tgt[keep])
            stats.append(st)
            k += 1
            if rc or (max_steps is not None and k >= max_steps):
                break
    finally:
        f.close()
    return np.concatenate(feats), np.concatenate(tgts), stats


def column_scale(raw, levels: int, floor: float) -> np.ndarray:
    """s_col per sample: rms over live levels of E plus the floor."""
    E = np.asarray(raw[:, :levels, 8], dtype=np.float64)
    return np.sqrt(np.mean(E * E, axis=1)) + floor


def fit_scale(raw, tgt, levels: int, floor: float = 1e-5, beta: float = 1.0):
    """Standardisation of the used channels over live levels, and the
    output scale: the spread of the relative residual (tgt - beta E)/s_col
    over the samples that have a previous step (its 99th-percentile
    magnitude over 3, so outliers do not set it)."""
    live = np.asarray(raw[:, :levels, :NLIVE], dtype=np.float64).reshape(-1, NLIVE)
    s = default_scale(floor, beta)
    mu = live.mean(axis=0)
    sd = live.std(axis=0)
    s[:NLIVE] = mu
    s[NFEAT:NFEAT + NLIVE] = np.where(sd > 1e-6 * np.maximum(1.0, np.abs(mu)), sd, 1.0)
    has_prev = np.any(np.asarray(raw[:, :levels, 8]) != 0.0, axis=1)
    r = relative_residual(raw[has_prev], tgt[has_prev], levels, floor, beta)
    q = float(np.percentile(np.abs(r[:, :levels]), 99.0)) if r.size else 0.0
    s[2 * NFEAT] = q / 3.0 if q > 0 else 1.0
    return s


def relative_residual(raw, tgt, levels: int, floor: float, beta: float) -> np.ndarray:
    E = np.asarray(raw[..., 8], dtype=np.float64)
    sc = column_scale(raw, levels, floor)
    r = (np.asarray(tgt, dtype=np.float64) - beta * E) / sc[:, None]
    r[:, levels:] = 0.0
    return r


def training_set(raw, tgt, levels: int, scale, clip: float = 10.0):
    """Network inputs (bf16 bits) and targets y = residual / out, clipped
    to +-clip, of the samples that have a previous step (E defined)."""
    has_prev = np.any(np.asarray(raw[:, :levels, 8]) != 0.0, axis=1)
    raw, tgt = raw[has_prev], tgt[has_prev]
    floor, beta, out = float(scale[2 * NFEAT + 1]), float(scale[2 * NFEAT + 2]), float(scale[2 * NFEAT])
    y = np.clip(relative_residual(raw, tgt, levels, floor, beta) / out, -clip, clip).astype(np.float32)
    return normalize(raw, scale, levels), y


def normalize(raw, scale, levels: int) -> np.ndarray:
    """Standardised bf16 inputs [S][Lp][16]; padded levels and unused
    channels are zero."""
    mu = np.asarray(scale[:NFEAT], dtype=np.float32)
    sd = np.asarray(scale[NFEAT:2 * NFEAT], dtype=np.float32)
    z = (np.asarray(raw, dtype=np.float32) - mu) / sd
    z[:, levels:] = 0.0
    z[..., NLIVE:] = 0.0
    return to_bf16_bits(z)


def train_guess(raw, tgt, levels: int, steps: int = 2000, batch: int = 1024, lr: float = 1e-3, seed: int = 0,
                floor: float = 1e-5, beta: float = 1.0):
    """A network trained on the pairs (raw, tgt); returns (net, losses)."""
    scale = fit_scale(raw, tgt, levels, floor, beta)
    x, t = training_set(raw, tgt, levels, scale)
    net = GuessNet(levels, init_params(seed), scale)
    net.load(x, t)
    losses = net.train(steps, min(batch, x.shape[0]), lr, seed)
    return net, losses

def compare(problem, nets: dict, ordering: int = G.ORDER_COLUMN3, max_steps: int | None = None) -> dict:
    """The plain initial iterate against networks (name -> GuessNet) at
    identical step sizes.

    A free plain run fixes the accepted DELTAT sequence; the plain method
    and every network then replay it side by side (so none back-steps where
    the free run did not), and each step's Picard and PCG counts, time and
    the largest relative psi difference from the plain replay are kept."""
    f = G.FlowGPU(problem, ordering)
    dts = []
    try:
        while True:
            rc, st = f.advance()
            dts.append(st["DELTAT"])
            if rc or (max_steps is not None and len(dts) >= max_steps):
                break
    finally:
        f.close()
    stat = static_features(problem)
    modes = dict(plain=None, **nets)
    runs = {}
    try:
        for mode, net in modes.items():
            g = G.FlowGPU(problem, ordering)
            g.set_dt_schedule(dts)
            g.guess(net, stat)
            runs[mode] = g
        out = {k: dict(picard=[], pcg=[], back=[], ms=[], psi_rel=[]) for k in runs}
        for _ in dts:
            ref = None
            for mode, g in runs.items():
                rc, st = g.advance()
                o = out[mode]
                o["picard"].append(st["ITER"])
                o["pcg"].append(st["NITERT"])
                o["back"].append(st["BACK"])
                o["ms"].append(g.counters()["step_ms"])
                psi = g.psi()
                if ref is None:
                    ref = psi
                o["psi_rel"].append(float(np.max(np.abs(psi - ref)) / np.max(np.abs(ref))))
    finally:
        for g in runs.values():
            g.close()
    res = dict(steps=len(dts), dts=np.asarray(dts))
    for mode, o in out.items():
        res[mode] = {k: np.asarray(v) for k, v in o.items()}
    return res
