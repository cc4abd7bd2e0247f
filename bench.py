#!/usr/bin/env python3
"""Throughput of the CATHY subsurface Picard step on one MI355X.

Workload: the 1.25M-node heterogeneous hillslope (201x201-node DEM, 2 m
cells, 0-20 m spectral relief, 30 graded layers to 10 m, lognormal Ks per
(layer, cell), +-20% vG alpha/n, 6 h of 5e-7 m/s rain then drainage, free
drainage at the base), FP64 Picard with IC(0)-PCG (FP32 factor, 3-colour
column ordering).  A step = one accepted implicit time step of one member
(FLOW3D returning converged, back-steps included) -- the metric's
"member Picard step".

The W warm-up steps run on the device and the state is then reset to the
initial condition, so the K timed steps are steps 1..K of the run; the CPU
baseline (the C oracle, one core) times the first steps of the same run.
The library call is synchronous, so the wall clock around the K calls
brackets all device work.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_params():
    from pycathy_wrapper_amd import configs as CF
    return CF.hillslope_small_params(tmax=172800.0)


def build_problem(cells: int, nstr: int):
    from pycathy_wrapper_amd import configs as CF
    return CF.hillslope_large(cells=cells, nstr=nstr, params=bench_params())


def cpu_baseline(problem, seconds: float) -> dict:
    """The C oracle (one core) on the first steps of the same run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    m = problem.mesh
    orc = P.Oracle()
    run = orc.run_create(dict(tetra=m.tetra, x=m.x, y=m.y, z=m.z, arenod=m.arenod, nnod=m.nnod, nstr=m.nstr,
                              ntri=m.ntri), problem.soil, problem.psi0, problem.atm_times, problem.atm_vals, 1,
                         P.OrcParams(**problem.params.as_dict()))
    steps = picard = pcg = 0
    t0 = time.perf_counter()
    while True:
        rc, s = run.step()
        steps += 1
        picard += s["ITER"]
        pcg += s["NITERT"]
        el = time.perf_counter() - t0
        log(f"[cpu] step {steps} dt={s['DELTAT']:.4g} iter={s['ITER']:.0f} pcg={s['NITERT']:.0f} t={el:.1f}s")
        if rc or el >= seconds:
            break
    run.close()
    return dict(value=steps / el, unit="member-steps/s", cores=1, kind="port",
                sample=f"oracle/cathy_oracle.c, steps 1..{steps} of the same run ({picard:.0f} Picard, "
                       f"{pcg:.0f} PCG iterations, natural-order IC(0)) in {el:.1f} s on one host core")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=200)
    ap.add_argument("--nstr", type=int, default=30)
    ap.add_argument("--ordering", choices=["column3", "natural"], default="column3")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="0 skips the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one member per GPU on a single MI355X; --gpus {args.gpus} "
              f"is not supported (ensemble members shard across GPUs without communication, see DESIGN.md)",
              file=sys.stderr)
        return 2

    from pycathy_wrapper_amd import gpu

    t0 = time.perf_counter()
    pr = build_problem(args.cells, args.nstr)
    m = pr.mesh
    log(f"[setup] mesh n={m.n} nt={m.nt} built in {time.perf_counter() - t0:.1f}s")
    name, cus = gpu.device_info()
    order = gpu.ORDER_COLUMN3 if args.ordering == "column3" else gpu.ORDER_NATURAL
    t0 = time.perf_counter()
    g = gpu.FlowGPU(pr, order)
    log(f"[setup] device {name} ({cus} CUs), flow created in {time.perf_counter() - t0:.1f}s")

    for w in range(args.warmup):
        rc, s = g.advance()
        log(f"[warmup] step {w + 1} dt={s['DELTAT']:.4g} iter={s['ITER']:.0f} pcg={s['NITERT']:.0f}")
    g.init(pr.psi0)  # timed steps restart from the initial condition

    g.profile(True)
    steps = picard = pcg = backs = 0
    sim0 = 0.0
    sim1 = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        rc, s = g.advance()
        steps += 1
        picard += s["ITER"]
        pcg += s["NITERT"]
        backs += s["BACK"]
        sim1 = s["TIME"]
        if rc:
            break
    t1 = time.perf_counter()
    spmv_ms, spmv_n = g.profile(None)
    g.profile(False)
    wall = t1 - t0
    log(f"[timed] {steps} steps in {wall:.3f}s: {picard:.0f} Picard, {pcg:.0f} PCG, {backs:.0f} back-steps, "
        f"sim {sim0:.1f}->{sim1:.1f}s")
    spmv_bytes = g.spmv_bytes()
    avg_ms = spmv_ms / max(1, spmv_n)
    achieved = spmv_bytes / (avg_ms * 1e-3) / 1e9 if spmv_n else 0.0
    log(f"[roofline] k_spmv_dot2: {spmv_n} launches, {avg_ms * 1e3:.2f} us avg, {spmv_bytes / 1e6:.1f} MB "
        f"-> {achieved:.0f} GB/s")
    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(pr, args.cpu_seconds)
    out = {
        "metric": "ensemble member Picard steps/s at 1.2M nodes, learned vs plain; SpMV % HBM peak",
        "value": steps / wall,
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"hillslope_large: {m.nrow}x{m.ncol} cells, {m.nstr} layers, N={m.n} nodes, "
                        f"NT={m.nt} tets, plain Picard + IC(0)-PCG ({args.ordering}, FP32 factor), one member",
            "nodes": m.n,
            "tets": m.nt,
            "picard_per_s": picard / wall,
            "pcg_per_s": pcg / wall,
            "sim_s_per_wall_s": (sim1 - sim0) / wall,
            "picard_iterations": picard,
            "pcg_iterations": pcg,
            "back_steps": backs,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_spmv_dot2",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_us": avg_ms * 1e3,
            "launches": spmv_n,
            "bytes_per_launch": spmv_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
