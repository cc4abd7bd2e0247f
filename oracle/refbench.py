"""The reference's own FLOW3D timed on the host (BASELINE INFRASTRUCTURE:
bench.py's cpu_baseline leg runs it as a child process; nothing here is
product code).

One process = one ensemble member, as pyCATHY's data-assimilation forecast
runs one ./cathy per member (pyCATHY/DA/cathy_DA.py:1389-1390).  The member
is the hillslope_large generator at the given size and seed; the reference
is the build sized for it (oracle/_ref/large) or the weil build when the
mesh fits that.  Steps 1..K of the run are taken, stopping after K steps or
once the time budget is spent; the mesh and harness set-up are not timed.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=200)
    ap.add_argument("--nstr", type=int, default=30)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--tmax", type=float, default=172800.0)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--seconds", type=float, default=60.0)
    ap.add_argument("--wait", action="store_true", help="print 'ready' after the set-up, then wait for a line on "
                                                        "stdin before the timed steps")
    a = ap.parse_args()
    from pycathy_wrapper_amd import configs as CF
    import pyoracle as P
    import refrun as R
    t0 = time.perf_counter()
    pr = CF.hillslope_large(seed=a.seed, cells=a.cells, nstr=a.nstr, params=CF.hillslope_small_params(tmax=a.tmax))
    m = pr.mesh
    lib = P.REF_LARGE_SO if os.path.exists(P.REF_LARGE_SO) else P.REF_SO
    ref = P.RefLib(lib)
    if m.n > ref.nmax or pr.soil["nzone"] > ref.maxzon or m.nstr > ref.maxstr:
        if a.wait:
            print("ready", flush=True)
        print(json.dumps(dict(error=f"mesh larger than the reference build {os.path.basename(lib)}")))
        return 1
    rr = R.RefFlowRun(ref, m, pr.soil, pr.psi0, float(pr.atm_vals[0]), pr.params.as_dict())
    setup = time.perf_counter() - t0
    if a.wait:
        print("ready", flush=True)
        sys.stdin.readline()
    steps, step_s, picard, pcg, backs = 0, [], 0, 0, 0
    tstart = time.time()
    t1 = time.perf_counter()
    while steps < a.steps:
        ts = time.perf_counter()
        rc, s = rr.step()
        step_s.append(time.perf_counter() - ts)
        steps += 1
        picard += s["ITER"]   # the accepted attempts, as the port's figure counts them
        pcg += s["NITERT"]
        backs += s["BACK"]
        if rc or time.perf_counter() - t1 > a.seconds:
            break
    wall = time.perf_counter() - t1
    print(json.dumps(dict(kind="reference", library=os.path.basename(lib), nodes=m.n, seed=a.seed, setup_s=setup,
                          steps=steps, seconds=wall, step_s=step_s, picard=picard, pcg=pcg, back_steps=backs,
                          t_start=tstart, t_end=time.time())))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
