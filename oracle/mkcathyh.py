"""A CATHY.H sized for one mesh, for a second build of the reference's flow
routines (TEST INFRASTRUCTURE: the reference solver that tests and the CPU
baseline run; nothing here is product code).

pyCATHY sizes the CATHY processor per project by rewriting its CATHY.H
(pyCATHY/cathy_tools.py:792-1030, update_cathyH): it keeps the header's
first 109 lines and writes the PARAMETER lines from the DEM's rows and
columns, the layer count, the zone count and a few parm values.  This script
applies those rules to the CATHY.H of the reference source tree and writes
the result, with symbolic links to the reference's .f sources and its other
.H files, into one directory: a Fortran INCLUDE resolves in the directory of
the file being compiled, so the unmodified sources compile there against
the sized header.  Also writes sizes.json (the array bounds a caller of the
built library needs).  Only build outputs go to --out (under oracle/_ref/).
"""
from __future__ import annotations

import argparse
import glob
import json
import os


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--src", required=True, help="the reference's CATHY source directory")
    ap.add_argument("--out", required=True)
    ap.add_argument("--rows", type=int, required=True, help="DEM rows (ROWMAX)")
    ap.add_argument("--cols", type=int, required=True, help="DEM columns (COLMAX)")
    ap.add_argument("--nstr", type=int, required=True, help="layers (MAXSTR)")
    ap.add_argument("--nzone", type=int, required=True, help="zones (MAXZON)")
    ap.add_argument("--n1", type=int, default=30, help="most element connections of a node (N1MAX)")
    ap.add_argument("--nprt", type=int, default=1)
    ap.add_argument("--nr", type=int, default=0)
    ap.add_argument("--numvp", type=int, default=1)
    # bounds of the listed boundary conditions, which update_cathyH leaves at
    # 1 and a user raises by hand for non-empty nansfdirbc / nansfneubc /
    # sfbc files (cathy_tools.py:2270-2273), and of the vegetation types
    ap.add_argument("--npmax", type=int, default=1)
    ap.add_argument("--nqmax", type=int, default=1)
    ap.add_argument("--nsfmax", type=int, default=1)
    ap.add_argument("--nnsfmx", type=int, default=1)
    ap.add_argument("--maxveg", type=int, default=1)
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    head = open(os.path.join(a.src, "CATHY.H")).readlines()[:109]
    # update_cathyH's values (cathy_tools.py:848-895) and derived ones (:897-901)
    nodmax = (a.rows + 1) * (a.cols + 1)
    maxcel = a.rows * a.cols
    ntrmax = 2 * maxcel
    maxtrm = a.rows * a.cols * a.nstr * 30
    nfacemax = nodmax * 5
    lines = [
        f"      PARAMETER (ROWMAX={a.rows},COLMAX={a.cols},DEMRES=1)\n",
        "      PARAMETER (MAXCEL=ROWMAX*COLMAX,MAXRES=1)\n",
        f"      PARAMETER (NODMAX={nodmax})\n",
        f"      PARAMETER (NTRMAX={ntrmax})\n",
        f"      PARAMETER (NP2MAX=1,MAXSTR={a.nstr})\n",
        f"      PARAMETER (NFACEMAX={nfacemax})\n",
        "      PARAMETER (NMAX=NODMAX*(MAXSTR + 1),NTEMAX=3*NTRMAX*MAXSTR)\n",
        f"      PARAMETER (NPMAX={a.npmax},NPMAX_TRA=1,NQMAX={a.nqmax},NSFMAX={a.nsfmax})\n",
        f"      PARAMETER (NNSFMX={a.nnsfmx},MAXDIR=NODMAX+NPMAX+NSFMAX*NNSFMX)\n",
        "      PARAMETER (MAXNUDN=1,MAXNUDT=1,MAXNUDC=1)\n",
        f"      PARAMETER (MAXZON={a.nzone},MAXTRM={maxtrm},MAXIT=30,MAXVEG={a.maxveg})\n",
        f"      PARAMETER (NRMAX={a.nr + 1},MAXPRT={a.nprt},MAXVP={a.numvp})\n",
        f"      PARAMETER (N1MAX={a.n1},NTPMAX=N1MAX*NMAX)\n",
        "      PARAMETER (MAXBOT=1,INTBOT=1,MAXQOUT=1)\n",
        "cxcx  PARAMETER (NIAUXMAX=NFACEMAX + MAXTRM + 1)\n",
        "cxcx  PARAMETER (NRAUXMAX=5*NFACEMAX + MAXTRM,NQMAX_TRA=NODMAX)\n",
        "      PARAMETER (NIAUXMAX=NMAX + MAXTRM + 1)\n",
        "      PARAMETER (NRAUXMAX=5*NMAX + MAXTRM,NQMAX_TRA=NODMAX)\n",
        f"      PARAMETER (MAXVTKPRT={a.nprt})\n",
        "      PARAMETER (MAXFCONTONODE=100,MAXLKP=3)\n",
    ]
    hdr = os.path.join(a.out, "CATHY.H")
    tmp = hdr + ".tmp"
    with open(tmp, "w") as f:
        f.writelines(head + lines)
    os.replace(tmp, hdr)
    for path in glob.glob(os.path.join(a.src, "*.f")) + glob.glob(os.path.join(a.src, "*.H")):
        name = os.path.basename(path)
        if name == "CATHY.H":
            continue
        link = os.path.join(a.out, name)
        if os.path.lexists(link):
            os.remove(link)
        os.symlink(path, link)
    sizes = dict(rowmax=a.rows, colmax=a.cols, maxcel=maxcel, nodmax=nodmax, maxstr=a.nstr,
                 nmax=nodmax * (a.nstr + 1), ntemax=3 * ntrmax * a.nstr, maxzon=a.nzone, maxtrm=maxtrm, n1max=a.n1,
                 npmax=a.npmax, nqmax=a.nqmax, nsfmax=a.nsfmax, nnsfmx=a.nnsfmx, maxveg=a.maxveg)
    with open(os.path.join(a.out, "sizes.json"), "w") as f:
        json.dump(sizes, f)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
