"""A whole subsurface run on the device, writing CATHY's output files.

``run_flow`` is the loop a CATHY host runs around FLOW3D (cathy_main.f:
2882-3860) with the time stepping, back-steps and boundary updates inside
``cathy_gpu_flow3d_advance``: it writes an ``mbeconv`` row per accepted
step and ``psi`` / ``sw`` (text + full-precision sidecars) at the initial
time, at the requested steps or times, and at the last step
(``outputs.CathyOutputs``).
"""
from __future__ import annotations

from . import gpu as G
from .outputs import CathyOutputs


def run_flow(problem, outdir: str, ordering: int = G.ORDER_COLUMN3, output_steps=(), output_times=(),
             max_steps: int | None = None) -> dict:
    """Run ``problem`` to TMAX (or ``max_steps``); returns a summary with
    the final return code (1 = TMAX, 2 = NOBACK), the step count and the
    Picard / linear iteration totals."""
    steps = set(int(s) for s in output_steps)
    times = sorted(float(t) for t in output_times)
    g = G.FlowGPU(problem, ordering)
    picard = linear = 0.0
    k = 0
    rc = 0
    try:
        with CathyOutputs(outdir) as out:
            out.nodal(0, 0.0, g.psi(), g.sw())
            while True:
                rc, st = g.advance()
                c = g.counters()
                picard += c["picard"]
                linear += c["pcg"]
                k += 1
                out.step(st)
                due = int(st["NSTEP"]) in steps
                while times and st["TIME"] >= times[0] - 1e-9 * max(1.0, abs(times[0])):
                    times.pop(0)
                    due = True
                last = rc != 0 or (max_steps is not None and k >= max_steps)
                if due or last:
                    out.nodal(int(st["NSTEP"]), st["TIME"], g.psi(), g.sw())
                if last:
                    break
    finally:
        g.close()
    return dict(rc=rc, steps=k, picard=picard, linear=linear)
