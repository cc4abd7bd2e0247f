"""Autotuner of the ensemble launch shapes, per mesh size.

The step's kernels have fixed shapes where the hardware sets them (SELL
slices of 64 rows = one wavefront, 256-thread work-groups, gather windows
picked per colour from the pattern at set-up).  What depends on the mesh
size and the ensemble is tuned here, by timing steps on the device:

* members per lockstep group: how many members share one launch per kernel
  (cathy_gpu_ensemble_*); groups run on streams of their own.  Small meshes
  fill the GPU only together; a 1.25M-node member fills it alone, and
  coupling such members in lockstep makes each wait for the slowest;
* the PCG launch batches between convergence polls (first batch as a
  percentage of the previous solve's iterations, follow-up batch size;
  cathy_gpu_flow3d_set_pcg_batches);
* the transposed-read SpMV (cathy_gpu_flow3d_set_spmv_transposed);
* HIP graphs for the lockstep rounds (cathy_gpu_ensemble_set_graphs).

``python -m pycathy_wrapper_amd.autotune`` runs the sweep on the GPU and
writes the table pycathy_wrapper_amd/tuning.json, one entry per mesh size
class of the ensemble generator (configs.ensemble_member S / M / L);
``load``, ``entry_for`` and ``groups`` read it (bench.py uses them).  None of
the knobs changes results: every configuration computes the same bits.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

DEFAULT_ENTRY = dict(members_per_group=1, pcg_first_pct=100, pcg_followup=2, spmv_transposed=0, graphs=0)
MARGIN = 0.01  # a knob leaves its default only for a gain above this fraction


def load(path: str = TABLE) -> dict:
    """The tuning table, or the defaults (every member on its own stream,
    the library's default PCG batches) when none has been written."""
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return dict(classes=[], default=dict(DEFAULT_ENTRY))


def entry_for(nodes: int, table: dict | None = None) -> dict:
    """The entry of the size class nearest to ``nodes`` (by log ratio)."""
    import math
    t = table or load()
    best, dist = None, None
    for c in t.get("classes", []):
        d = abs(math.log(max(nodes, 1) / max(c["nodes"], 1)))
        if dist is None or d < dist:
            best, dist = c, d
    return dict(t.get("default", DEFAULT_ENTRY), **(best or {}))


def groups(nodes: list[int], table: dict | None = None) -> list[list[int]]:
    """Lockstep groups of member indices: members of one mesh size together,
    members_per_group of them per group (from the size's table entry)."""
    t = table or load()
    by_n: dict[int, list[int]] = {}
    for j, n in enumerate(nodes):
        by_n.setdefault(n, []).append(j)
    out = []
    for n in sorted(by_n):
        k = max(1, int(entry_for(n, t)["members_per_group"]))
        js = by_n[n]
        out.extend(js[i:i + k] for i in range(0, len(js), k))
    return out


def apply(flow, entry: dict) -> None:
    """Set a flow handle's PCG batches and SpMV variant from a table entry."""
    flow.pcg_batches(int(entry["pcg_first_pct"]), int(entry["pcg_followup"]))
    flow.spmv_transposed(bool(entry["spmv_transposed"]))


# ------------------------------------------------------------------ sweep --
def _timed(problems, flows, grp: list[list[int]], steps: int, graphs: bool = False) -> float:
    """member-steps/s of ``steps`` steps of every member from its initial
    state, lockstep groups ``grp`` concurrently (one host thread each)."""
    from . import gpu
    for pr, f in zip(problems, flows):
        f.init(pr.psi0)
    ens = [gpu.EnsembleGPU([flows[j] for j in g]) for g in grp]
    for e in ens:
        e.graphs(graphs)
    done = [0] * len(ens)

    def run(i):
        _, taken, _ = ens[i].advance(steps)
        done[i] = int(taken.sum())

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(ens))]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    for e in ens:
        e.close()
    return sum(done) / wall


def sweep_class(size: str, count: int, steps: int, mpg_candidates, log=print) -> dict:
    """Tune one size class on ``count`` members of configs.ensemble_member."""
    from . import configs as CF
    from . import gpu
    problems = [CF.ensemble_member(size, 100 + i, params=CF.hillslope_small_params(tmax=86400.0))
                for i in range(count)]
    flows = [gpu.FlowGPU(p, gpu.ORDER_COLUMN3) for p in problems]
    nodes = problems[0].mesh.n
    trials = []

    def trial(mpg, first, follow, transposed, graphs=0):
        for f in flows:
            apply(f, dict(pcg_first_pct=first, pcg_followup=follow, spmv_transposed=transposed))
        grp = [list(range(i, min(i + mpg, count))) for i in range(0, count, mpg)]
        rate = _timed(problems, flows, grp, steps, bool(graphs))
        trials.append(dict(members_per_group=mpg, pcg_first_pct=first, pcg_followup=follow,
                           spmv_transposed=transposed, graphs=graphs, member_steps_per_s=rate))
        log(f"[autotune] {size} ({nodes} nodes, {count} members): {mpg} per group, PCG batches {first}% + "
            f"{follow}, transposed SpMV {transposed}, graphs {graphs}: {rate:.2f} member-steps/s")
        return rate

    def pick(rate: dict, default):
        """The fastest value, but the default unless another beats it by
        more than the noise between repeated trials (1%)."""
        top = max(rate, key=rate.get)
        return top if rate[top] > (1.0 + MARGIN) * rate[default] else default

    _timed(problems, flows, [[j] for j in range(count)], 1)  # warm-up
    best = dict(DEFAULT_ENTRY)
    rate = {m: trial(m, 100, 2, 0) for m in mpg_candidates if m <= count}
    best["members_per_group"] = pick(rate, 1)
    mpg = best["members_per_group"]
    rate = {f: trial(mpg, 100, f, 0) for f in (1, 4)}
    rate[2] = max(t["member_steps_per_s"] for t in trials if t["members_per_group"] == mpg and
                  t["pcg_followup"] == 2)
    best["pcg_followup"] = pick(rate, 2)
    rate = {p: trial(mpg, p, best["pcg_followup"], 0) for p in (75, 125)}
    rate[100] = max(t["member_steps_per_s"] for t in trials if t["members_per_group"] == mpg and
                    t["pcg_followup"] == best["pcg_followup"] and t["pcg_first_pct"] == 100)
    best["pcg_first_pct"] = pick(rate, 100)
    r1 = trial(mpg, best["pcg_first_pct"], best["pcg_followup"], 1)
    r0 = max(t["member_steps_per_s"] for t in trials if t["members_per_group"] == mpg and
             t["pcg_followup"] == best["pcg_followup"] and t["pcg_first_pct"] == best["pcg_first_pct"] and
             not t["spmv_transposed"])
    best["spmv_transposed"] = pick({0: r0, 1: r1}, 0)
    rg = trial(mpg, best["pcg_first_pct"], best["pcg_followup"], best["spmv_transposed"], 1)
    best["graphs"] = pick({0: max(r0, r1), 1: rg}, 0)
    for f in flows:
        f.close()
    return dict(size=size, nodes=nodes, layers=problems[0].mesh.nstr, members=count, steps=steps,
                member_steps_per_s=max(r0, r1, rg), **best, trials=trials)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Tune the ensemble launch shapes per mesh size on the GPU.")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--sizes", default="S,M,L")
    ap.add_argument("--counts", default="32,8,3", help="members per size class in the sweep")
    args = ap.parse_args(argv)
    from . import gpu
    name, cus = gpu.device_info()
    cand = {"S": (1, 4, 8, 16, 32), "M": (1, 2, 4, 8), "L": (1, 2, 3)}
    sizes = args.sizes.split(",")
    counts = [int(c) for c in args.counts.split(",")]
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    classes = [sweep_class(sz, c, args.steps, cand[sz], log) for sz, c in zip(sizes, counts)]
    table = dict(device=name, compute_units=cus, steps=args.steps, default=dict(DEFAULT_ENTRY), classes=classes,
                 command="python -m pycathy_wrapper_amd.autotune " + " ".join(argv or sys.argv[1:]))
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log(f"[autotune] table written to {args.out}")
    print(json.dumps({c["size"]: {k: c[k] for k in ("nodes", "members_per_group", "pcg_first_pct", "pcg_followup",
                                                      "spmv_transposed", "graphs", "member_steps_per_s")}
                      for c in classes}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
