"""The Fortran drop-in boundary on the device.

1. pycathy_wrapper_amd/fortran/symslv_check.f90, a Fortran program, calls
   cathy_gpu_symslv through the ISO_C_BINDING module (the binding a CATHY
   build uses) on a small system with Dirichlet rows.
2. The reference's own FLOW3D / PICARD, compiled from its sources with the
   GPU SYMSLV shim linked in its place (oracle/_ref/libcathy_ref_gpu.so),
   runs the hillslope_small day: every linear solve of the reference's
   Picard loop happens on the MI355X.  The run must reproduce the
   reference's own run (tests/golden/hillslope_small_ref.npz): same time
   steps, Picard and PCG counts, psi to 1e-6 and mass balance to 1e-8.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CHECK = os.path.join(ROOT, "pycathy_wrapper_amd", "_build", "fortran", "symslv_check")
REF_GPU = os.path.join(ROOT, "oracle", "_ref", "libcathy_ref_gpu.so")
GOLD = os.path.join(HERE, "golden", "hillslope_small_ref.npz")


def test_fortran_program_calls_symslv():
    if not os.path.exists(CHECK):
        pytest.skip("Fortran layer not built (make -C pycathy_wrapper_amd/fortran)")
    out = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, file=sys.stderr)
    assert out.returncode == 0
    assert "SYMSLV CHECK OK" in out.stdout


def test_reference_flow3d_with_gpu_symslv_matches_reference(tmp_path):
    if not os.path.exists(REF_GPU):
        pytest.skip("oracle/_ref/libcathy_ref_gpu.so not built (needs the reference sources at build time)")
    g = np.load(GOLD)
    saved = [int(v) for v in g["steps"]]
    out = tmp_path / "dropin.npz"
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "dropin_run.py"), str(out),
                        ",".join(map(str, saved))], capture_output=True, text=True, timeout=600)
    print(r.stderr[-2000:], file=sys.stderr)
    assert r.returncode == 0, r.stdout[-2000:]
    d = np.load(out)
    names = [str(v) for v in g["stat_names"]]
    assert [str(v) for v in d["stat_names"]] == names
    ref, got = g["stats"], d["stats"]
    assert got.shape == ref.shape and int(d["rc"]) == int(g["rc_last"])
    col = {nm: j for j, nm in enumerate(names)}
    for nm in ("DELTAT", "TIME", "ITER", "NITERT", "BACK"):
        assert np.array_equal(got[:, col[nm]], ref[:, col[nm]]), nm
    scale = 1e-6 * np.abs(ref[:, col["STORE1"]]).max()
    for nm in ("DSTORE", "VIN", "VOUT", "ERRAS", "CDSTOR", "CERRAS"):
        err = np.abs(got[:, col[nm]] - ref[:, col[nm]])
        assert np.all(err <= 1e-8 * np.maximum(np.abs(ref[:, col[nm]]), scale)), nm
    for i in range(len(saved)):
        rel = np.max(np.abs(d["psi"][i] - g["psi"][i])) / np.max(np.abs(g["psi"][i]))
        assert rel <= 1e-6, (saved[i], rel)


@pytest.mark.parametrize("case", ["reservoir", "masked"])
def test_reference_route_call_runs_on_device(case):
    """ROUTE replaced by pycathy_wrapper_amd/fortran/route_gpu.f in the
    reference's library (linked ahead of route.o): the caller writes the
    network and flows into COMMON /RIVERNETWORK/ and /SURFWATER/ and calls
    ROUTE with the reference's arguments, then the reference's own ALTEZZE.
    "reservoir": the network holds a reservoir (SERBATOIO on the device)
    beside the channel cells.  "masked": 30% of the grid lies outside the
    catchment, so NCELL < NROW*NCOL as in any DEM with ZONE = 0 cells.  Six
    chained substeps against the oracle (bit-identical to the reference's
    ROUTE and ALTEZZE), to 1e-12, the pool level too (tests/dropin_route_run.py,
    one process per network)."""
    if not os.path.exists(REF_GPU):
        pytest.skip("oracle/_ref/libcathy_ref_gpu.so not built (needs the reference sources)")
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "dropin_route_run.py"), case],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], file=sys.stderr)
    assert r.returncode == 0 and "ok:" in r.stdout, r.stdout[-2000:]
