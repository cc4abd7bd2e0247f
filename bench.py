#!/usr/bin/env python3
"""Throughput of the CATHY subsurface Picard step on one MI355X.

Workload: an ensemble of M members (default 3) of the 1.25M-node
heterogeneous hillslope generator (201x201-node DEM, 2 m cells, 0-20 m
spectral relief, 30 graded layers to 10 m, lognormal Ks per (layer, cell),
+-20% vG alpha/n, 6 h of 5e-7 m/s rain then drainage, free drainage at the
base; member i uses seed 2+i, member 0 is the configuration's own run),
FP64 Picard with IC(0)-PCG (FP32 factor, 3-colour column ordering).  A step
= one accepted implicit time step of one member (FLOW3D returning
converged, back-steps included) -- the metric's "member Picard step".
``value`` = member-steps/s of the whole ensemble: the members run steps
1..K in lockstep (cathy_gpu_ensemble_*): each round every member still
iterating takes one Picard iteration, and each kernel of it is one launch
over all members, with a PCG convergence mask per member.  The same steps
with one HIP stream and host thread per member are reported beside it
(``config.streams``).

Member 0 first runs the same K steps alone (``config.single_member``):
that pass carries the SpMV roofline, so no other stream shares the GPU
while the kernel is timed.  The W warm-up steps run on the device and each
state is then reset to the initial condition; the CPU baseline (the C
oracle, one core) times the first steps of member 0's run.  Every library
call is synchronous, so the wall clock around the calls brackets all
device work.

The roofline entry is the PCG SpMV (k_spmv_dot2): its average kernel
duration comes from the dispatch packets' own timestamps
(hipExtLaunchKernelGGL start/stop events on the library's stream) over the
timed region; the bytes are the algorithmic CSR SpMV bytes
12*nnz + 4*(N+1) + 16*N.  ``traffic`` is the measured HBM bytes per launch
from rocprofv3 FETCH_SIZE / WRITE_SIZE passes, read from
profiles/k_spmv_dot2_hillslope_large_traffic.txt when that summary exists
(made by --pmc-summary from two --pmc runs of this script with --calib-mb).

The learned leg (``config.learned``; skipped with --learned-members 0)
trains the initial-iterate network on the device from the plain solver's
own first steps on independent small members of the same generator (41x41
nodes, same layers, forcing and settings, other seeds), then (1) times the
same K steps of the 1.25M-node run with the learned initial iterate under
its own step control, and (2) replays the plain run's accepted step sizes
with the plain and the learned initial iterate, comparing Picard/PCG
counts and psi step by step.  ``value`` stays the plain method's rate.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "ensemble member Picard steps/s at 1.2M nodes, learned vs plain; SpMV % HBM peak"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_spmv_dot2_hillslope_large_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_params():
    from pycathy_wrapper_amd import configs as CF
    return CF.hillslope_small_params(tmax=172800.0)


def build_problem(cells: int, nstr: int, seed: int = 2):
    from pycathy_wrapper_amd import configs as CF
    return CF.hillslope_large(seed=seed, cells=cells, nstr=nstr, params=bench_params())


def full_run(pr, g, seconds: float) -> dict:
    """Member 0 alone from its initial state through the whole 2-day run
    (or until `seconds` of wall time), timed as one block: the later windows
    -- the 6 h rain pulse, the drainage after it -- next to the first steps
    the default line times."""
    g.init(pr.psi0)
    n = pic = pcg = backs = 0
    marks = {}
    t0 = time.perf_counter()
    rc, s = 0, {}
    while True:
        rc, s = g.advance()
        c = g.counters()
        n += 1
        pic += c["picard"]
        pcg += c["pcg"]
        backs += s["BACK"]
        for hrs in (6, 24, 48):   # wall time and steps when the run passes 6 h, 1 day, 2 days
            if hrs not in marks and s["TIME"] >= hrs * 3600.0 - 1e-6:
                marks[hrs] = dict(steps=n, wall_s=time.perf_counter() - t0)
        if rc or time.perf_counter() - t0 > seconds:
            break
    wall = time.perf_counter() - t0
    return dict(steps=n, wall_s=wall, steps_per_s=n / wall, sim_s=s.get("TIME", 0.0),
                sim_s_per_wall_s=s.get("TIME", 0.0) / wall, picard=pic, pcg=pcg, back_steps=backs, rc=rc,
                last_dt=s.get("DELTAT", 0.0), marks_h=marks)


def ensemble_run(members, steps: int) -> dict:
    """Steps 1..K of every member at once, one host thread and HIP stream
    per member (the library releases nothing between calls; ctypes drops
    the interpreter lock during each call)."""
    import threading
    res = [None] * len(members)

    for pr, g in members:  # every member from its initial state, set before the clock starts
        g.init(pr.psi0)

    def run(j):
        pr, g = members[j]
        n = pic = pc = bk = 0
        sim = mbe = 0.0
        lat = []
        for _ in range(steps):
            ts = time.perf_counter()
            rc, s = g.advance()
            lat.append(time.perf_counter() - ts)
            c = g.counters()
            n += 1
            pic += c["picard"]
            pc += c["pcg"]
            bk += s["BACK"]
            sim = s["TIME"]
            mbe = max(mbe, abs(s["ERREL"]))
            if rc:
                break
        res[j] = dict(steps=n, picard=pic, pcg=pc, back_steps=bk, sim_s=sim, max_rel_mbe_pct=mbe, step_s=lat)

    threads = [threading.Thread(target=run, args=(j,)) for j in range(len(members))]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    if any(r is None for r in res):
        raise RuntimeError("an ensemble member failed")
    return dict(wall=wall, members=res)


def lockstep_prepare(members, graphs: bool | None = None):
    """An ensemble handle over the members, each reset to its initial state
    (set-up outside any timed region)."""
    from pycathy_wrapper_amd import gpu
    for pr, g in members:
        g.init(pr.psi0)
    e = gpu.EnsembleGPU([g for _, g in members])
    if graphs is None:  # the autotuner's choice for the members' size
        from pycathy_wrapper_amd import autotune
        graphs = bool(autotune.entry_for(members[0][0].mesh.n)["graphs"])
    e.graphs(graphs)
    return e


def lockstep_run(members, steps: int, graphs: bool | None = None, e=None) -> dict:
    """Steps 1..K of every member in lockstep (cathy_gpu_ensemble_*): each
    round every member still iterating takes one Picard iteration, one
    launch per kernel for all of them; members leave the lockstep when they
    have taken K steps."""
    if e is None:
        e = lockstep_prepare(members, graphs)
    t0 = time.perf_counter()
    rc, taken, stats = e.advance(steps)
    wall = time.perf_counter() - t0
    info = e.info()
    res = []
    for j in range(len(members)):
        st = stats[j]
        res.append(dict(steps=int(taken[j]), picard=sum(s["PICARD"] for s in st), pcg=sum(s["PCG"] for s in st),
                        back_steps=sum(s["BACK"] for s in st), sim_s=st[-1]["TIME"] if st else 0.0,
                        max_rel_mbe_pct=max((abs(s["ERREL"]) for s in st), default=0.0),
                        step_s=[v / 1e3 for v in e.wall_ms[j]]))
    e.close()
    return dict(wall=wall, members=res, info=info)


def make_groups(members, group_rows: int) -> list[list[int]]:
    """Lockstep groups: members of equal node count together; with
    group_rows < 0 as many per group as the autotuner's table gives for the
    size (pycathy_wrapper_amd/tuning.json), else each group filled up to
    group_rows rows (at least one member).  Returns member indices per
    group."""
    if group_rows < 0:
        from pycathy_wrapper_amd import autotune
        return autotune.groups([pr.mesh.n for pr, _ in members])
    order = sorted(range(len(members)), key=lambda j: members[j][0].mesh.n)
    groups, cur, rows, size = [], [], 0, None
    for j in order:
        n = members[j][0].mesh.n
        if cur and (n != size or rows + n > group_rows):
            groups.append(cur)
            cur, rows = [], 0
        cur.append(j)
        rows += n
        size = n
    if cur:
        groups.append(cur)
    return groups


def grouped_run(members, steps: int, group_rows: int) -> dict:
    """Steps 1..K of every member: lockstep groups (make_groups), each group
    on its own HIP stream and host thread, the groups concurrently."""
    import threading
    groups = make_groups(members, group_rows)
    out = [None] * len(groups)
    # a group of one member is that member's own run on its own stream
    ens = [lockstep_prepare([members[j] for j in g]) if len(g) > 1 else None for g in groups]
    for g, e in zip(groups, ens):
        if e is None:
            pr, fl = members[g[0]]
            fl.init(pr.psi0)

    def solo(j):
        pr, g = members[j]
        n = pic = pc = bk = 0
        sim = mbe = 0.0
        lat = []
        for _ in range(steps):
            rc, s = g.advance()
            lat.append(g.last_wall_ms / 1e3)
            n += 1
            pic += s["PICARD"]
            pc += s["PCG"]
            bk += s["BACK"]
            sim = s["TIME"]
            mbe = max(mbe, abs(s["ERREL"]))
            if rc:
                break
        return dict(steps=n, picard=pic, pcg=pc, back_steps=bk, sim_s=sim, max_rel_mbe_pct=mbe, step_s=lat)

    def run(i):
        if ens[i] is None:
            out[i] = dict(members=[solo(groups[i][0])], info=dict(rounds=0))
        else:
            out[i] = lockstep_run([members[j] for j in groups[i]], steps, e=ens[i])

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(groups))]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    if any(o is None for o in out):
        raise RuntimeError("a lockstep group failed")
    res = [None] * len(members)
    for g, o in zip(groups, out):
        for j, r in zip(g, o["members"]):
            res[j] = r
    return dict(wall=wall, members=res, groups=[len(g) for g in groups],
                rounds=sum(o["info"]["rounds"] for o in out))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def reference_baseline(args, members: int, max_steps: int, seconds: float) -> dict | None:
    """The reference's own FLOW3D (Fortran, flang -O2, sized by pyCATHY's
    update_cathyH rules) on the host, one process per member as pyCATHY's
    DA forecast runs ./cathy (oracle/refbench.py): first member 0 alone
    (the baseline value), then min(members, cores - 5) members at once, the
    forecast pool of pyCATHY/DA/cathy_DA.py:83,1389."""
    import subprocess
    script = os.path.join(ROOT, "oracle", "refbench.py")
    base = [sys.executable, "-u", script, "--cells", str(args.cells), "--nstr", str(args.nstr), "--steps",
            str(max_steps), "--seconds", str(seconds)]

    def run(seeds):
        # every process builds its member first and reports ready; the timed
        # steps of all start together
        procs = [subprocess.Popen(base + ["--seed", str(s), "--wait"], stdin=subprocess.PIPE,
                                  stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True) for s in seeds]
        for p in procs:
            p.stdout.readline()
        for p in procs:
            p.stdin.write("go\n")
            p.stdin.flush()
        out = []
        for p in procs:
            o, _ = p.communicate(timeout=900)
            line = [l for l in o.splitlines() if l.startswith("{")]
            out.append(json.loads(line[-1]) if line else dict(error=f"no output (rc {p.returncode})"))
        return out

    log("[cpu] the reference FLOW3D, member 0 alone (set-up untimed)")
    one = run([2])[0]
    if "error" in one:
        log(f"[cpu] reference: {one['error']}")
        return None
    log(f"[cpu] reference: {one['steps']} step(s) in {one['seconds']:.2f} s ({one['picard']} Picard, "
        f"{one['pcg']} PCG)")
    share = min(16, os.cpu_count() or 1)  # this process's share of the host
    nproc = max(1, min(members, share - 5))
    log(f"[cpu] the reference, {nproc} member processes at once")
    pool = run([2 + i for i in range(nproc)])
    ok = [r for r in pool if "error" not in r]
    pool_rate = None
    if ok:
        span = max(r["t_end"] for r in ok) - min(r["t_start"] for r in ok)
        pool_rate = sum(r["steps"] for r in ok) / span
        log(f"[cpu] reference pool: {sum(r['steps'] for r in ok)} member-steps in {span:.2f} s -> "
            f"{pool_rate:.3f} member-steps/s")
    return dict(value=one["steps"] / one["seconds"], unit="member-steps/s", cores=1, kind="reference",
                steps=one["steps"], seconds=one["seconds"], cpu_model=cpu_model(),
                sample=f"the reference's own FLOW3D and subsurface time loop (Fortran sources built with flang -O2 "
                       f"-fno-automatic against a CATHY.H sized by pyCATHY's update_cathyH rules, "
                       f"oracle/_ref/large; natural-order FP64 IC(0)-PCG): steps 1..{one['steps']} of member 0 "
                       f"in {one['seconds']:.2f} s on one host core ({one['picard']} Picard, {one['pcg']} PCG); "
                       f"set-up untimed",
                process_pool=dict(processes=nproc, member_steps_per_s=pool_rate,
                                  members=[dict((k, r.get(k)) for k in ("seed", "steps", "seconds", "picard", "pcg"))
                                           for r in pool],
                                  note="one process per member, min(members, cores - 5) at a time, as "
                                       "pyCATHY's DA forecast pool"))


def cpu_baseline(problem, max_steps: int, seconds: float) -> dict:
    """The C oracle (one core) on the first steps of the same run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    m = problem.mesh
    orc = P.Oracle()
    run = orc.run_create(dict(tetra=m.tetra, x=m.x, y=m.y, z=m.z, arenod=m.arenod, nnod=m.nnod, nstr=m.nstr,
                              ntri=m.ntri), problem.soil, problem.psi0, problem.atm_times, problem.atm_vals, problem.hspatm,
                         P.OrcParams(**problem.params.as_dict()))
    steps = picard = pcg = 0
    t0 = time.perf_counter()
    while True:
        rc, s = run.step()
        steps += 1
        picard += s["ITER"]
        pcg += s["NITERT"]
        el = time.perf_counter() - t0
        log(f"[cpu] step {steps} dt={s['DELTAT']:.4g} iter={s['ITER']:.0f} pcg={s['NITERT']:.0f} "
            f"back={s['BACK']:.0f} t={el:.1f}s")
        if rc or steps >= max_steps or el >= seconds:
            break
    run.close()
    return dict(value=steps / el, unit="member-steps/s", cores=1, kind="port", steps=steps, seconds=el,
                sample=f"oracle/cathy_oracle.c (C restatement of the reference FLOW3D, natural-order FP64 IC(0)), "
                       f"steps 1..{steps} of the same run in {el:.1f} s on one host core; ITER/NITERT of the "
                       f"accepted attempts: {picard:.0f} Picard, {pcg:.0f} PCG")


def read_traffic():
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    for ln in open(TRAFFIC_FILE):
        if ":" in ln and not ln.lstrip().startswith("#"):
            k, v = ln.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def _pmc_rows(paths: list[str]):
    for pat in paths:
        for f in glob.glob(os.path.join(pat, "**", "*counter_collection.csv"), recursive=True) + \
                 (glob.glob(pat) if pat.endswith(".csv") else []):
            with open(f) as fh:
                yield from csv.DictReader(fh)


def trace_summary(paths: list[str], skip: float = 1.0 / 3.0) -> int:
    """Where the GPU's time goes in a rocprofv3 --kernel-trace CSV: after the
    first `skip` of the dispatches (set-up), every instant is shared equally
    among the kernels running then (so concurrent streams are not counted
    twice); idle time is split by the kernel pair around each gap."""
    import collections
    import csv
    import re

    def short(name):
        name = re.sub(r"^void ", "", name).split("(")[0]
        return re.sub(r"cg::", "", name)

    ev = []
    for p in paths:
        with open(p, newline="") as fh:
            for r in csv.DictReader(fh):
                ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"])))
    ev.sort()
    ev = ev[int(len(ev) * skip):]
    pts = sorted([(s, 1, i) for i, (s, _, _) in enumerate(ev)] + [(e, -1, i) for i, (_, e, _) in enumerate(ev)])
    active, share, conc, gaps = set(), collections.Counter(), collections.Counter(), collections.Counter()
    last, last_end_kernel, idle = pts[0][0], None, 0
    for t, typ, i in pts:
        dt = t - last
        if dt > 0:
            if active:
                for j in active:
                    share[ev[j][2]] += dt / len(active)
                conc[len(active)] += dt
            else:
                idle += dt
                gaps[(last_end_kernel, ev[i][2] if typ == 1 else None)] += dt
        last = t
        if typ == 1:
            active.add(i)
        else:
            active.discard(i)
            last_end_kernel = ev[i][2]
    tot = last - pts[0][0]
    print(f"# window {tot / 1e6:.1f} ms after the first {skip:.0%} of {len(ev) / (1 - skip):.0f} dispatches; "
          f"GPU idle {idle / 1e6:.1f} ms ({idle / tot:.1%}); time with k kernels running: " +
          ", ".join(f"{k}: {v / tot:.1%}" for k, v in sorted(conc.items())))
    print(f"{'kernel':60s} {'shared_ms':>10s} {'share':>7s}")
    for k, v in share.most_common(25):
        print(f"{k[:60]:60s} {v / 1e6:10.2f} {v / tot:7.1%}")
    print(f"# idle gaps by the kernels around them (ended -> started)")
    for (a, b), v in gaps.most_common(8):
        print(f"{str(a)[:40]:40s} -> {str(b)[:40]:40s} {v / 1e6:8.2f} ms")
    return 0


def pmc_summary(paths: list[str], kernel: str, copy_bytes: float, all_kernels: bool = False) -> int:
    """Average per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of ``kernel`` and of
    the calibration copy kernel from rocprofv3 --pmc CSV output, and the HBM
    bytes per launch each implies after scaling by copy_bytes / counter.
    With all_kernels, the same per-launch figures for every kernel seen."""
    acc = {}
    for row in _pmc_rows(paths):
        name = row.get("Kernel_Name", "")
        if "k_copy" in name:
            key = "copy"
        elif kernel in name:
            key = "spmv"
        elif all_kernels:
            key = name.split("(")[0].replace("void ", "").strip()
        else:
            continue
        cn = row.get("Counter_Name", "")
        v = float(row.get("Counter_Value", "nan"))
        acc.setdefault((key, cn), []).append(v)
    out = {}
    for (key, cn), vs in sorted(acc.items()):
        # launches queued after PCG convergence exit at once and move
        # (almost) nothing; only the live launches are averaged
        live = [v for v in vs if v >= 0.05 * max(vs)] or vs
        out[(key, cn)] = (sum(live) / len(live), len(live), len(vs))
        if key in ("spmv", "copy"):
            print(f"{key} {cn}: {len(live)} live of {len(vs)} dispatches, mean {out[(key, cn)][0]:.1f} KiB")
    total = 0.0
    ok = True
    factor = {}
    for cn in ("FETCH_SIZE", "WRITE_SIZE"):
        if ("copy", cn) not in out:
            ok = False
            continue
        factor[cn] = copy_bytes / (out[("copy", cn)][0] * 1024.0)
        if ("spmv", cn) not in out:
            ok = False
            continue
        b = out[("spmv", cn)][0] * 1024.0 * factor[cn]
        print(f"{cn}_calibration_factor: {factor[cn]:.4f}")
        print(f"{cn}_bytes_per_launch: {b:.0f}")
        total += b
    if ok:
        print(f"traffic_bytes_per_launch: {total:.0f}")
    if all_kernels and len(factor) == 2:
        print(f"{'kernel':58s} {'live':>6s} {'fetch MB':>9s} {'write MB':>9s} {'HBM MB':>8s}")
        names = sorted({k for (k, _) in out if k not in ("copy",)})
        for k in names:
            f = out.get((k, "FETCH_SIZE"))
            w = out.get((k, "WRITE_SIZE"))
            if not f or not w:
                continue
            fb = f[0] * 1024.0 * factor["FETCH_SIZE"] / 1e6
            wb = w[0] * 1024.0 * factor["WRITE_SIZE"] / 1e6
            print(f"{('cg::k_spmv_dot2' if k == 'spmv' else k)[:58]:58s} {f[1]:6d} {fb:9.2f} {wb:9.2f} {fb + wb:8.2f}")
    return 0 if ok else 1


def learned_leg(g, pr, args, dts, steps: int) -> dict:
    """Train the initial-iterate network on small members, then time and
    compare it on this run (see the module docstring)."""
    import numpy as np
    from pycathy_wrapper_amd import configs as CF
    from pycathy_wrapper_amd import gpu
    from pycathy_wrapper_amd import guess as GU
    L = pr.mesh.nstr + 1
    t0 = time.perf_counter()
    raws, tgts = [], []
    for i in range(args.learned_members):
        mem = CF.hillslope_large(seed=1000 + i, cells=40, nstr=args.nstr, params=bench_params())
        r, t, st = GU.collect(mem, gpu.ORDER_COLUMN3, max_steps=steps + 10)
        raws.append(r)
        tgts.append(t)
        log(f"[learned] member {i}: {len(st)} steps, {sum(x['ITER'] for x in st):.0f} Picard")
    raw, tgt = np.concatenate(raws), np.concatenate(tgts)
    t1 = time.perf_counter()
    net, losses = GU.train_guess(raw, tgt, L, steps=args.learned_train_steps, batch=2048, lr=1e-3)
    t2 = time.perf_counter()
    log(f"[learned] {raw.shape[0]} column-step pairs in {t1 - t0:.1f}s; {args.learned_train_steps} Adam steps in "
        f"{t2 - t1:.1f}s, loss {np.mean(losses[:10]):.3f} -> {np.mean(losses[-100:]):.3f}")
    stat = GU.static_features(pr)
    out = dict(trained_on=f"{args.learned_members} members hillslope_large(cells=40, seeds 1000..), "
                          f"steps 1..{steps + 10} each", pairs=int(raw.shape[0]),
               collect_s=t1 - t0, train_s=t2 - t1, loss_first=float(np.mean(losses[:10])),
               loss_last=float(np.mean(losses[-100:])))
    # (1) own step control, same number of steps
    g.init(pr.psi0)
    g.guess(net, stat)
    n = pic = pc = bk = 0
    sim = 0.0
    ta = time.perf_counter()
    for _ in range(steps):
        rc, s = g.advance()
        c = g.counters()
        n += 1
        pic += c["picard"]
        pc += c["pcg"]
        bk += s["BACK"]
        sim = s["TIME"]
        if rc:
            break
    tb = time.perf_counter()
    out.update(steps_per_s=n / (tb - ta), picard=pic, pcg=pc, back_steps=bk, sim_s_per_wall_s=sim / (tb - ta))
    log(f"[learned] own control: {n} steps in {tb - ta:.3f}s, {pic:.0f} Picard, {pc:.0f} PCG, {bk:.0f} back-steps, "
        f"sim -> {sim:.2f} s")
    # (2) replay of the plain run's step sizes, plain then learned
    rep = {}
    for mode in ("plain", "learned"):
        g.init(pr.psi0)
        g.guess(net if mode == "learned" else None, detach=mode == "plain")
        g.set_dt_schedule(dts)
        pic = pc = bk = 0
        ta = time.perf_counter()
        for _ in range(len(dts)):
            rc, s = g.advance()
            c = g.counters()
            pic += c["picard"]
            pc += c["pcg"]
            bk += s["BACK"]
            if rc:
                break
        tb = time.perf_counter()
        rep[mode] = dict(picard=pic, pcg=pc, back_steps=bk, seconds=tb - ta, psi=g.psi())
    a, b = rep["plain"], rep["learned"]
    rel = float(np.max(np.abs(b["psi"] - a["psi"])) / np.max(np.abs(a["psi"])))
    out["replay"] = dict(steps=len(dts), plain={k: v for k, v in a.items() if k != "psi"},
                         learned={k: v for k, v in b.items() if k != "psi"}, psi_rel_diff_last=rel)
    log(f"[learned] replay of {len(dts)} steps: Picard {a['picard']:.0f} -> {b['picard']:.0f}, PCG {a['pcg']:.0f} -> "
        f"{b['pcg']:.0f}, back-steps {a['back_steps']:.0f} / {b['back_steps']:.0f}, time {a['seconds']:.3f} -> "
        f"{b['seconds']:.3f} s, psi rel diff {rel:.2e}")
    g.guess(None, detach=True)
    net.close()
    return out


def mixed_ensemble(args) -> int:
    """Steps 1..K of every member of the mixed-size ensemble at once (one
    host thread and HIP stream per member); prints the JSON line with
    member-steps/s of the whole ensemble."""
    from pycathy_wrapper_amd import configs as CF
    from pycathy_wrapper_amd import gpu
    import numpy as np
    stress = args.stress > 0
    f = args.stress if stress else args.mixed
    base = (("S", 64), ("M", 48), ("L", 16)) if stress else (("S", 32), ("M", 24), ("L", 8))
    counts = {sz: max(1, int(round(c * f))) for sz, c in base}
    members, sizes, iopts = [], [], []
    t0 = time.perf_counter()
    seed = 100
    for sz, c in counts.items():
        for i in range(c):
            iopt = 2 if (stress and i % 2 == 0) else 1
            pr = CF.ensemble_member(sz, seed, params=CF.hillslope_small_params(tmax=86400.0, iopt=iopt))
            if stress:  # dry start, 2 h storm, then nothing
                pr.psi0 = np.full(pr.mesh.n, -5.0)
                pr.atm_times = np.array([0.0, 7200.0, 7201.0])
                pr.atm_vals = np.array([2e-6, 2e-6, 0.0])
            fl = gpu.FlowGPU(pr, gpu.ORDER_COLUMN3)
            if args.group_rows < 0:
                from pycathy_wrapper_amd import autotune
                autotune.apply(fl, autotune.entry_for(pr.mesh.n))
            members.append((pr, fl))
            sizes.append(sz)
            iopts.append(iopt)
            seed += 1
        log(f"[setup] {c} members of size {sz} ({members[-1][0].mesh.n} nodes) after {time.perf_counter() - t0:.1f}s")
    for pr, g in members:  # one warm-up step each, then back to the initial state
        for _ in range(args.warmup):
            g.advance()
    # Picard members in lockstep groups (one launch per kernel for a
    # group's members, groups on streams of their own); Newton members,
    # which the lockstep does not take, on a stream each beside them
    pic = [j for j, io in enumerate(iopts) if io == 1]
    newt = [j for j, io in enumerate(iopts) if io == 2]
    if args.group_rows == 0:   # every member on its own stream
        newt, pic = newt + pic, []
    import threading
    out_lock, out_streams = {}, {}

    def run_lock():
        out_lock["r"] = grouped_run([members[j] for j in pic], args.steps, args.group_rows)

    def run_streams():
        out_streams["r"] = ensemble_run([members[j] for j in newt], args.steps)

    threads = [threading.Thread(target=run_lock)] if pic else []
    if newt:
        threads.append(threading.Thread(target=run_streams))
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    res = [None] * len(members)
    if pic:
        for j, r in zip(pic, out_lock["r"]["members"]):
            res[j] = r
    if newt:
        for j, r in zip(newt, out_streams["r"]["members"]):
            res[j] = r
    ens = dict(wall=wall, members=res, groups=out_lock["r"]["groups"] if pic else [])
    streams = None
    if pic:  # the same steps with one stream and host thread per member
        sr = ensemble_run(members, args.steps)
        st_steps = sum(r["steps"] for r in sr["members"])
        streams = dict(member_steps_per_s=st_steps / sr["wall"], wall_s=sr["wall"], member_steps=st_steps)
        log(f"[streams] {len(members)} members, one stream each: {st_steps} member-steps in {sr['wall']:.3f}s -> "
            f"{st_steps / sr['wall']:.2f} member-steps/s")
    total = sum(r["steps"] for r in ens["members"])
    per = {sz: dict(members=counts[sz], nodes=int(next(pr.mesh.n for (pr, _), s2 in zip(members, sizes) if s2 == sz)),
                    steps=sum(r["steps"] for r, s2 in zip(ens["members"], sizes) if s2 == sz),
                    picard=sum(r["picard"] for r, s2 in zip(ens["members"], sizes) if s2 == sz),
                    pcg=sum(r["pcg"] for r, s2 in zip(ens["members"], sizes) if s2 == sz)) for sz in counts}
    lat = np.concatenate([r["step_s"] for r in ens["members"]])
    by_iopt = {("newton" if io == 2 else "picard"): dict(
        members=sum(1 for x in iopts if x == io),
        steps=sum(r["steps"] for r, x in zip(ens["members"], iopts) if x == io),
        max_rel_mbe_pct=max((r["max_rel_mbe_pct"] for r, x in zip(ens["members"], iopts) if x == io),
                            default=0.0)) for io in sorted(set(iopts))}
    lat_ms = dict(p50=1e3 * float(np.percentile(lat, 50)), p99=1e3 * float(np.percentile(lat, 99)))
    for r in ens["members"]:
        del r["step_s"]
    log(f"[{'stress' if stress else 'mixed'}] {len(members)} members: {total} member-steps in {ens['wall']:.3f}s -> "
        f"{total / ens['wall']:.2f} member-steps/s; step latency p50 {lat_ms['p50']:.1f} ms, p99 {lat_ms['p99']:.1f} "
        f"ms; {per}; {by_iopt}")
    out = {
        "metric": METRIC, "value": total / ens["wall"], "unit": "member-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * ens["wall"] / total,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("stress" if stress else "mixed-size") + " ensemble: " +
                               ", ".join(f"{c} x {sz}" for sz, c in counts.items()) +
                               f" members (ensemble_member, seeds 100..), steps 1..{args.steps} each, Picard "
                               f"members in lockstep (one launch per kernel), Newton members on a HIP stream each, "
                               + ("dry start psi = -5 m, 2 h of 2e-6 m/s rain, Newton "
                               "(ILU-BiCGSTAB) for every other member, Picard + IC(0)-PCG for the rest"
                               if stress else "plain Picard + IC(0)-PCG") + " (column3, FP32 factor)",
                   "per_size": per, "by_linearisation": by_iopt, "step_latency_ms": lat_ms,
                   "lockstep_group_rows": args.group_rows, "lockstep_groups": ens["groups"],
                   "streams": streams,
                   "learned_guess": False, "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    for _, g in members:
        g.close()
    return 0


def route_bench(members: int, steps: int, warmup: int) -> int:
    """Surface routing substeps (ROUTE + ALTEZZE, one launch for all
    members) on the synthetic 201x201-cell network, through the C-ABI with
    host buffers (each call copies the state in and out).  The oracle times
    one member's substep on one core beside it."""
    import numpy as np
    from pycathy_wrapper_amd import routing as RT
    net = RT.synthetic_network(201, 201, dx=10.0, seed=4)
    rng = np.random.default_rng(21)
    shp = (members, net.ngrid)
    area = net.dx * net.dy
    st = dict(surface_water=rng.uniform(0.0, 2e-6, shp) * area, q_in_kk=rng.uniform(0.0, 0.05, shp),
              q_out_kk1=rng.uniform(0.0, 0.04, shp), q_out_kk2=rng.uniform(0.0, 0.02, shp),
              volume_kk=rng.uniform(0.0, 0.01, shp) * area, q_in_kkp1=np.zeros(shp), q_out_kkp1_1=np.zeros(shp),
              q_out_kkp1_2=np.zeros(shp), volume_kkp1=np.zeros(shp), h_water_kkp1=np.zeros(shp))
    g = RT.RouteGPU(net)
    log(f"[route] {net.ngrid} cells, {g.nlevel} levels, {members} members")
    for _ in range(warmup):
        g.step(15.0, st)
    t0 = time.perf_counter()
    for _ in range(steps):
        st["q_in_kkp1"][...] = 0.0
        g.step(15.0, st)
    dt = time.perf_counter() - t0
    g.close()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    orc = P.Oracle()
    one = {k: np.ascontiguousarray(v[0]) for k, v in st.items()}
    nrep, t1 = 0, time.perf_counter()
    while nrep < 3 or time.perf_counter() - t1 < 2.0:
        one["q_in_kkp1"][...] = 0.0
        orc.route(net, 15.0, one)
        nrep += 1
    cpu = (time.perf_counter() - t1) / nrep
    print(json.dumps(dict(metric="routing member-substeps/s (ROUTE + ALTEZZE, 201x201 cells)",
                          value=members * steps / dt, unit="member-substeps/s", n_gpus=1, steps=steps,
                          warmup=warmup, ms_per_step=1e3 * dt / steps, higher_is_better=True, dtype="f64",
                          data="synthetic", config=dict(workload="route_201x201", members=members,
                                                        levels=g.nlevel, host_buffers=True),
                          cpu_baseline=dict(value=1.0 / cpu, unit="member-substeps/s", cores=1, kind="port",
                                            sample=f"{nrep} substeps of member 0"))), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=200)
    ap.add_argument("--nstr", type=int, default=30)
    ap.add_argument("--ordering", choices=["column3", "natural"], default="column3")
    ap.add_argument("--cpu-steps", type=int, default=1, help="time steps of the CPU baseline (0 skips it)")
    ap.add_argument("--cpu-seconds", type=float, default=60.0, help="stop the CPU baseline after this long")
    ap.add_argument("--calib-mb", type=float, default=0.0,
                    help="after the timed steps, run the copy kernel over this many MB (PMC calibration)")
    ap.add_argument("--members", type=int, default=3, help="ensemble members")
    ap.add_argument("--group-rows", type=int, default=-1,
                    help="lockstep groups of members up to this many rows each, one stream per group (0: one "
                         "stream per member, no lockstep; -1: the group sizes, PCG batches and SpMV variant of the "
                         "autotuner's table pycathy_wrapper_amd/tuning.json)")
    ap.add_argument("--fp64-factor", type=int, default=1,
                    help="also run member 0's steps with an FP64 IC(0) factor (config.fp64_factor; 0 skips)")
    ap.add_argument("--stress", type=float, default=0.0,
                    help="instead: the stress ensemble (64 S + 48 M + 16 L members times this factor; dry start "
                         "psi = -5 m, 2 h of 2e-6 m/s rain; Newton for half the members, Picard for the rest)")
    ap.add_argument("--mixed", type=float, default=0.0,
                    help="instead: the mixed-size ensemble (32 S + 24 M + 8 L members times this factor, "
                         "41x41x15 / 101x101x20 / 201x201x30 nodes), all members concurrently")
    ap.add_argument("--learned-members", type=int, default=8,
                    help="small members whose first steps train the learned initial iterate (0 skips the leg)")
    ap.add_argument("--learned-train-steps", type=int, default=4000)
    ap.add_argument("--route-only", type=int, default=0, metavar="MEMBERS",
                    help="time the surface routing substep (ROUTE + ALTEZZE) of MEMBERS members on a "
                         "201x201-cell network instead, print its own JSON line and exit")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR_OR_CSV",
                    help="summarise rocprofv3 --pmc CSV output (FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--full-run", type=float, default=120.0,
                    help="wall seconds allowed for member 0's whole 2-day run, timed after the default line's "
                         "legs (0 skips it)")
    ap.add_argument("--cold-reps", type=int, default=50,
                    help="SpMV launches timed after a 1 GB cache flush each (0: skip)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the ensemble leg's timed steps (for profiling; prints no JSON line)")
    ap.add_argument("--trace-summary", nargs="+", metavar="CSV",
                    help="summarise rocprofv3 --kernel-trace CSV output (GPU time shared among concurrent "
                         "kernels, idle gaps) and exit")
    ap.add_argument("--pmc-all", action="store_true",
                    help="with --pmc-summary: per-launch HBM bytes of every kernel, not only the SpMV")
    args = ap.parse_args()
    if args.pmc_summary:
        return pmc_summary(args.pmc_summary, "k_spmv_dot2", args.calib_mb * 1e6, args.pmc_all)
    if args.trace_summary:
        return trace_summary(args.trace_summary)
    if args.gpus != 1:
        print(f"bench.py: one GPU only; --gpus {args.gpus} is not supported (members of an ensemble shard "
              f"across GPUs without communication; no multi-GPU run is part of this build)", file=sys.stderr)
        return 2

    from pycathy_wrapper_amd import gpu

    if args.route_only > 0:
        return route_bench(args.route_only, args.steps, args.warmup)
    if args.mixed > 0 or args.stress > 0:
        return mixed_ensemble(args)

    t0 = time.perf_counter()
    pr = build_problem(args.cells, args.nstr)
    m = pr.mesh
    log(f"[setup] mesh n={m.n} nt={m.nt} built in {time.perf_counter() - t0:.1f}s")
    name, cus = gpu.device_info()
    order = gpu.ORDER_COLUMN3 if args.ordering == "column3" else gpu.ORDER_NATURAL
    t0 = time.perf_counter()
    g = gpu.FlowGPU(pr, order)
    log(f"[setup] device {name} ({cus} CUs), flow created in {time.perf_counter() - t0:.1f}s")
    from pycathy_wrapper_amd import autotune
    tuned = autotune.entry_for(m.n) if args.group_rows < 0 else None
    if tuned is not None:
        autotune.apply(g, tuned)
        log(f"[setup] tuned launch shapes for {m.n} nodes: {dict((k, tuned[k]) for k in autotune.DEFAULT_ENTRY)}")
    lay = g.layout()
    log(f"[setup] full pattern nnz={lay['nnz']}, SELL-64 slots={lay['nnz_sell']} "
        f"({lay['nnz_sell'] / lay['nnz'] - 1:.1%} padding), longest row {lay['maxrow']}, coded columns "
        f"{lay['coded']} (check mismatches {lay['code_mismatch']})")

    for w in range(args.warmup):
        rc, s = g.advance()
        log(f"[warmup] step {w + 1} dt={s['DELTAT']:.4g} iter={s['ITER']:.0f} pcg={s['NITERT']:.0f}")
    g.init(pr.psi0)  # timed steps restart from the initial condition
    others = []
    for i in range(1, args.members):
        t0 = time.perf_counter()
        pi = build_problem(args.cells, args.nstr, seed=2 + i)
        gi = gpu.FlowGPU(pi, order)
        if tuned is not None:
            autotune.apply(gi, tuned)
        for w in range(args.warmup):
            gi.advance()
        gi.init(pi.psi0)
        others.append((pi, gi))
        log(f"[setup] member {i} (seed {2 + i}) built, warmed up and reset in {time.perf_counter() - t0:.1f}s")

    if args.ensemble_only:  # for profiling the ensemble: its timed leg alone, no JSON line
        ens = grouped_run([(pr, g)] + others, args.steps, args.group_rows)
        log(f"[ensemble-only] {sum(r['steps'] for r in ens['members'])} member-steps in {ens['wall']:.3f}s")
        return 0
    # member 0's K steps alone, timed without per-launch events; then the
    # same K steps again from the initial state with the SpMV's dispatch
    # timestamps recorded (the roofline), which the events slow slightly
    dts = []
    steps = picard = pcg = backs = 0
    sim1 = 0.0
    step_wall = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        rc, s = g.advance()
        step_wall.append(time.perf_counter() - ts)
        c = g.counters()
        steps += 1
        picard += c["picard"]   # every Picard iteration, back-stepped attempts included
        pcg += c["pcg"]
        backs += s["BACK"]
        sim1 = s["TIME"]
        dts.append(s["DELTAT"])
        if rc:
            break
    t1 = time.perf_counter()
    g.init(pr.psi0)
    g.profile(True)
    for k in range(steps):
        rc_p, _ = g.advance()
        if rc_p:
            break
    spmv_ms, spmv_n = g.profile(None)
    g.profile(False)
    wall = t1 - t0
    log(f"[timed] {steps} steps in {wall:.3f}s: {picard:.0f} Picard, {pcg:.0f} PCG iterations (back-stepped "
        f"attempts included), {backs:.0f} back-steps, sim 0 -> {sim1:.2f} s")
    spmv_bytes = g.spmv_bytes()
    avg_ms = spmv_ms / max(1, spmv_n)
    achieved = spmv_bytes / (avg_ms * 1e-3) / 1e9 if spmv_n else 0.0
    log(f"[roofline] k_spmv_dot2: {spmv_n} launches, {avg_ms * 1e3:.2f} us avg, {spmv_bytes / 1e6:.1f} MB "
        f"-> {achieved:.0f} GB/s")
    # the same SpMV with the caches cleared before each launch (1 GB copy:
    # A and p come from HBM, not the 256 MB last-level cache) and warm
    cold_ms, warm_ms = g.bench_spmv(args.cold_reps, 1024.0) if args.cold_reps > 0 else (0.0, 0.0)
    cold_gbs = spmv_bytes / (cold_ms * 1e-3) / 1e9 if cold_ms > 0 else 0.0
    if args.cold_reps > 0:
        log(f"[roofline] k_spmv_dot2 after a 1 GB cache flush: {cold_ms * 1e3:.2f} us avg over {args.cold_reps} "
            f"-> {cold_gbs:.0f} GB/s ({cold_gbs / HBM_PEAK_GBS:.3f} of peak); warm {warm_ms * 1e3:.2f} us")
    if args.calib_mb > 0:
        ms = gpu.calib_copy(int(args.calib_mb * 1e6), 5)
        log(f"[calib] copy of {args.calib_mb:.0f} MB: {ms * 1e3:.1f} us/launch")
    allm = [(pr, g)] + others
    ens = grouped_run(allm, steps, args.group_rows)
    for r in ens["members"]:
        r.pop("step_s")
    ens_steps = sum(r["steps"] for r in ens["members"])
    log(f"[grouped] {len(ens['members'])} members in lockstep groups {ens['groups']}: {ens_steps} member-steps "
        f"in {ens['wall']:.3f}s -> {ens_steps / ens['wall']:.2f} member-steps/s "
        f"(member 0 alone: {steps / wall:.2f} steps/s)")
    one = lockstep_run(allm, steps)
    one_steps = sum(r["steps"] for r in one["members"])
    one_rate = one_steps / one["wall"]
    log(f"[lockstep] all {len(allm)} members in one group: {one_steps} member-steps in {one['wall']:.3f}s -> "
        f"{one_rate:.2f} member-steps/s; {one['info']}")
    streams = ensemble_run([(pr, g)] + others, steps)
    for r in streams["members"]:
        r.pop("step_s")
    st_steps = sum(r["steps"] for r in streams["members"])
    streams_rate = st_steps / streams["wall"]
    log(f"[streams] {len(streams['members'])} members: {st_steps} member-steps in {streams['wall']:.3f}s -> "
        f"{streams_rate:.2f} member-steps/s")
    for _, gi in others:
        gi.close()
    psi32 = g.psi()  # member 0 after its K steps (FP32 factor), before the full run moves it on
    full = None
    if args.full_run > 0:
        full = full_run(pr, g, args.full_run)
        log(f"[full run] member 0 alone: {full['steps']} steps, sim 0 -> {full['sim_s']:.0f} s in "
            f"{full['wall_s']:.2f}s ({full['steps_per_s']:.2f} steps/s, {full['sim_s_per_wall_s']:.0f} sim-s per "
            f"wall-s), {full['picard']:.0f} Picard, {full['pcg']:.0f} PCG, {full['back_steps']:.0f} back-steps, "
            f"rc {full['rc']}; passes {full['marks_h']}")
    fp64 = None
    if args.fp64_factor and args.ordering == "column3":
        # the same steps of member 0 with its IC(0) factor in FP64 (the
        # column ordering's factor is FP32 otherwise)
        import numpy as np
        g.factor_fp64(True)
        g.init(pr.psi0)
        n64 = pic64 = pcg64 = 0
        t0 = time.perf_counter()
        for k in range(steps):
            rc, s = g.advance()
            c = g.counters()
            n64 += 1
            pic64 += c["picard"]
            pcg64 += c["pcg"]
            if rc:
                break
        w64 = time.perf_counter() - t0
        rel = float(np.max(np.abs(g.psi() - psi32)) / np.max(np.abs(psi32)))
        g.factor_fp64(False)
        fp64 = dict(steps=n64, steps_per_s=n64 / w64, picard=pic64, pcg=pcg64, psi_rel_diff_vs_fp32=rel)
        log(f"[fp64 factor] {n64} steps in {w64:.3f}s -> {n64 / w64:.2f} steps/s: {pic64:.0f} Picard, {pcg64:.0f} "
            f"PCG iterations; psi rel diff vs the FP32 factor {rel:.2e}")
    learned = None
    if args.learned_members > 0:
        learned = learned_leg(g, pr, args, dts, steps)
    cpu = None
    if args.cpu_steps > 0 and args.cpu_seconds > 0:
        port = cpu_baseline(pr, args.cpu_steps, args.cpu_seconds)
        cpu = reference_baseline(args, args.members, args.cpu_steps, args.cpu_seconds)
        if cpu is None:  # no reference build for this size: the port is the baseline
            cpu = port
        else:
            cpu["port"] = {k: port[k] for k in ("value", "steps", "seconds", "sample")}
        k = min(cpu["steps"], len(step_wall))
        gsec = sum(step_wall[:k])
        cpu["device_seconds_same_steps"] = gsec
        log(f"[cpu] steps 1..{k}: {cpu['kind']} {cpu['seconds']:.2f} s, device {gsec:.3f} s")
    traffic = read_traffic()
    out = {
        "metric": METRIC,
        "value": ens_steps / ens["wall"],
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * ens["wall"] / ens_steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"hillslope_large ensemble: {args.members} members (seeds 2..{args.members + 1}), each "
                        f"{m.nrow}x{m.ncol} cells, {m.nstr} graded layers, N={m.n} nodes, NT={m.nt} tets, steps "
                        f"1..{steps} of the 2-day run, plain Picard + IC(0)-PCG ({args.ordering} ordering, FP32 "
                        f"factor, FP64 state/vectors), members in lockstep groups of "
                        f"{ens['groups']} members ("
                        + ("autotuned" if args.group_rows < 0 else f"up to {args.group_rows} rows each")
                        + "; one launch per kernel for a group's members, per-member PCG convergence masks; "
                        "groups on streams of their own)",
            "lockstep_group_rows": args.group_rows,
            "lockstep_groups": ens["groups"],
            "tuned": None if tuned is None else {k: tuned[k] for k in autotune.DEFAULT_ENTRY},
            "one_lockstep_group": {"member_steps_per_s": one_rate, "wall_s": one["wall"], **one["info"]},
            "streams": {"member_steps_per_s": streams_rate, "wall_s": streams["wall"],
                        "note": "the same members on one HIP stream and host thread each"},
            "members": args.members,
            "nodes": m.n,
            "tets": m.nt,
            "member_steps": ens_steps,
            "picard_per_s": sum(r["picard"] for r in ens["members"]) / ens["wall"],
            "pcg_per_s": sum(r["pcg"] for r in ens["members"]) / ens["wall"],
            "ensemble": ens["members"],
            "single_member": {"steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps,
                              "picard_per_s": picard / wall, "pcg_per_s": pcg / wall,
                              "sim_s_per_wall_s": sim1 / wall, "picard_iterations": picard,
                              "pcg_iterations": pcg, "back_steps": backs},
            "learned_guess": False,
            "learned": learned,
            "full_run": full,
            "fp64_factor": fp64,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_spmv_dot2",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_us": avg_ms * 1e3,
            "launches": spmv_n,
            "bytes_per_launch": spmv_bytes,
            "cold": {"avg_us": cold_ms * 1e3, "achieved": cold_gbs, "frac": cold_gbs / HBM_PEAK_GBS,
                     "launches": args.cold_reps, "flush_mb": 1024, "warm_avg_us": warm_ms * 1e3},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
