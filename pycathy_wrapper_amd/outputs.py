"""CATHY's subsurface output files, written from the device state.

The text files keep the reference's formats exactly, so pyCATHY's readers
(``pyCATHY/importers/cathy_outputs.py:418-657``) parse them unchanged:

* ``psi`` / ``sw``: per output time a header ``(I7,1PE16.8,'     NSTEP   TIME')``
  and the N nodal values as ``5(1PE15.6)`` (detout.f:29-45, :126-128); the
  ``sw`` file opens with ``'  0   HSPSW'`` (cathy_main.f:2821, :4134);
* ``mbeconv``: the ``#INITIAL VOLUME`` header (cathy_main.f:2864, FORMAT
  1220 at :4021-4030) and one row per accepted step (:3269-3273, FORMAT 1240
  at :4043-4044).

Beside them, ``psi.f64`` / ``sw.f64`` hold the same fields at full
precision (the text keeps 7 digits, too few to check a 1e-6 relative
criterion): per record int32 NSTEP, float64 TIME, int32 N, then N float64,
little-endian.
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np


def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran ``1PEw.d``: one digit before the point, ``d`` after, a
    two-digit exponent as ``E+dd``; three-digit exponents drop the ``E``
    (``+ddd``), as Ew.d without an Ee field prints them."""
    x = float(x)
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("-Infinity" if x < 0 else "Infinity").rjust(w)
    s = f"{x:.{d}E}"
    mant, exp = s.split("E")
    e = int(exp)
    if abs(e) <= 99:
        out = f"{mant}E{e:+03d}"
    else:
        out = f"{mant}{e:+04d}"
    if len(out) > w:
        return "*" * w
    return out.rjust(w)


def nodal_block(nstep: int, time: float, values) -> str:
    """One output time of ``psi`` / ``sw`` (detout.f FORMATs 1000, 1020)."""
    v = np.asarray(values, dtype=np.float64)
    lines = [f"{nstep:7d}{fortran_e(time, 16, 8)}     NSTEP   TIME"]
    for i in range(0, len(v), 5):
        lines.append("".join(fortran_e(x, 15, 6) for x in v[i:i + 5]))
    return "\n".join(lines) + "\n"


SW_HEADER = "  0   HSPSW\n"

MBECONV_HEADER = (
    "#INITIAL VOLUME OF WATER IN THE SUBSURFACE = {store0}\n"
    "#NSTEP       DELTAT         TIME  NLIN   AVG.LIN       STORE1       STORE2       DSTORE"
    "   CUM.DSTORE       VIN  CUM.VIN       VOUT    CUM.VOUT   VIN+VOUT   CUM. VIN  M.BAL.ERR"
    "   REL. MBE   CUM. MBE      CUM.\n"
    "#                                 ITER.    ITER.                             (>0 x inc)"
    "                                                                      + VOUT  (Vi+Vo-Ds)"
    "        (%)             ABS(MBE)\n")


def mbeconv_header(store0: float) -> str:
    return MBECONV_HEADER.format(store0=fortran_e(store0, 13, 5))


def mbeconv_row(st: dict) -> str:
    """One accepted step (FORMAT 1240): NSTEP, DELTAT, TIME, ITER,
    NITERT/ITER, STORE1, STORE2, DSTORE, CUM.DSTORE, VIN, CUM.VIN, VOUT,
    CUM.VOUT, VIN+VOUT, CUM.(VIN+VOUT), M.BAL.ERR, REL.MBE, CUM.MBE,
    CUM.ABS(MBE).  ``st`` holds the cathy_gpu_flow3d_advance statistics."""
    it = int(st["ITER"])
    parts = [f"{int(st['NSTEP']):6d}", fortran_e(st["DELTAT"], 13, 6), fortran_e(st["TIME"], 13, 6), f"{it:6d}",
             fortran_e(float(st["NITERT"]) / float(it), 10, 3), fortran_e(st["STORE1"], 13, 6),
             fortran_e(st["STORE2"], 13, 6), fortran_e(st["DSTORE"], 13, 5), fortran_e(st["CDSTOR"], 13, 5),
             fortran_e(st["VIN"], 10, 3), fortran_e(st["CVIN"], 10, 3)]
    for k in ("VOUT", "CVOUT", "VINOUT", "VTOT", "ERRAS", "ERREL", "CERRAS"):
        parts.append(fortran_e(st[k], 11, 3))
    parts.append(fortran_e(st["CAERAS"], 10, 3))
    return "".join(parts) + "\n"


class CathyOutputs:
    """The ``psi``, ``sw`` and ``mbeconv`` files of one run in ``outdir``,
    with the full-precision sidecars ``psi.f64`` and ``sw.f64``."""

    def __init__(self, outdir: str):
        os.makedirs(outdir, exist_ok=True)
        self.outdir = outdir
        self._psi = open(os.path.join(outdir, "psi"), "w")
        self._sw = open(os.path.join(outdir, "sw"), "w")
        self._sw.write(SW_HEADER)
        self._mb = open(os.path.join(outdir, "mbeconv"), "w")
        self._psi_bin = open(os.path.join(outdir, "psi.f64"), "wb")
        self._sw_bin = open(os.path.join(outdir, "sw.f64"), "wb")
        self._mb_started = False

    def nodal(self, nstep: int, time: float, psi, sw) -> None:
        """PSI and SW at an output time (DETOUT)."""
        self._psi.write(nodal_block(nstep, time, psi))
        self._sw.write(nodal_block(nstep, time, sw))
        for f, v in ((self._psi_bin, psi), (self._sw_bin, sw)):
            v = np.ascontiguousarray(v, dtype="<f8")
            f.write(struct.pack("<idi", int(nstep), float(time), v.size))
            f.write(v.tobytes())

    def step(self, st: dict) -> None:
        """One mbeconv row; the header (initial volume = the first step's
        STORE1) is written before the first row."""
        if not self._mb_started:
            self._mb.write(mbeconv_header(st["STORE1"]))
            self._mb_started = True
        self._mb.write(mbeconv_row(st))

    def close(self) -> None:
        for f in (self._psi, self._sw, self._mb, self._psi_bin, self._sw_bin):
            f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_sidecar(path: str):
    """Records of a ``psi.f64`` / ``sw.f64`` sidecar: list of (nstep, time, values)."""
    out = []
    with open(path, "rb") as f:
        while True:
            head = f.read(16)
            if len(head) < 16:
                break
            nstep, time, n = struct.unpack("<idi", head)
            out.append((nstep, time, np.frombuffer(f.read(8 * n), dtype="<f8").copy()))
    return out


def read_nodal_text(path: str):
    """Blocks of a ``psi`` / ``sw`` text file: list of (nstep, time, values)."""
    blocks = []
    cur = None
    with open(path) as f:
        for line in f:
            if "NSTEP" in line and "TIME" in line:
                parts = line.split()
                cur = [int(parts[0]), float(parts[1]), []]
                blocks.append(cur)
            elif cur is not None and line.strip():
                cur[2].extend(float(v) for v in line.split())
    return [(a, b, np.array(c)) for a, b, c in blocks]
