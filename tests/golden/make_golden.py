#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ (run in the development
container, where the reference tree is mounted at /root/reference).

1. meshes.npz -- mesh golden vectors held by the reference's stored runs:
   for each example project, the DEM raster (prepro/dem), zone raster
   (prepro/zone), the grid parameters (input/dem_parameters, the fields
   datin.f:185-290 reads) and the mesh CATHY wrote (output/grid3d, written by
   gen3d.f:91-101 before riord.f sorts the vertices: NNOD N NT, TETRA(5,NT),
   X Y Z per node at 1PE13.6), plus NTERM from output/risul when printed.
2. hillslope_small_ref.npz -- the hillslope_small run (21x21 nodes, 10
   layers, 1 day) stepped by the reference's own FLOW3D (compiled from the
   reference sources by oracle/Makefile into oracle/_ref/libcathy_ref.so and
   driven by oracle/refrun.py): per-step statistics of every step and psi/Sw
   at selected steps.
4. weil_outputs.npz -- text of the weil run's output files as the reference
   wrote them (examples/SSHydro/weill_exemple/output/): the first two output
   times of psi and sw (with sw's header line) and the whole mbeconv.
5. hillslope_bc_ref.npz -- configs.hillslope_bc (fixed heads, listed
   Neumann fluxes, root uptake) to its 3 h end by the reference's FLOW3D,
   BCONE/BCNXT/BCBAK on the written nansfdirbc/nansfneubc files, NEUMANN
   and ETRAN (oracle/_ref/bc): per-step statistics, psi/Sw/IFATM at
   selected steps.
6. hillslope_seepage_ref.npz -- configs.hillslope_seepage (two seepage
   faces, exit points in the convergence test) by the reference's FLOW3D
   with SFVONE/SFVREC on the written sfbc file, SFINIT, EXTALL/EXTCVG:
   per-step statistics, psi/Sw and the faces' SFEX at selected steps.
7. hillslope_small_as_written_ref.npz -- hillslope_small under its parm
   as written (DTMIN = 1 s, ITUNS 10/5/7): the reference's run to its
   NOBACK stop, every step's statistics and the final psi/Sw.
3. hillslope_small_newton_ref.npz -- the first 60 steps of the same run
   with the Newton linearisation (IOPT=2, ISOLV=0: ILU(0)-BiCGSTAB).  The
   reference's Newton mass-balance terms are stale (its Newton branch never
   sets FLOW3D's local SWNEW/SWTIMEP) and are stored as computed.

8. bc_noflow_project/, outlets_project/, weil_project/,
   spatial_atmbc_project/, soil3d_project/ -- example projects as pyCATHY
   lays them out (input/, prepro/: the files project.load_project reads),
   copied from examplesTmp/SSHydro/bc_noflow, examplesTmp/catchments/outlets,
   examples/SSHydro/weill_exemple, examples/SSHydro/atmbc_spatially_from_weill
   (per-node rainfall, HSPATM = 0), examples/SSHydro/soil3d_from_weill
   (400 soil zones) and examples/SSHydro/atmbc_spatially_from_weill_withnodata
   (a DEM with no-data cells, 326 surface nodes), with
   what the reference program wrote for them: output/mbeconv as written,
   output/psi_sw.npz (the psi and sw blocks of output/psi and output/sw,
   values as printed, with their NSTEP and TIME) and, for weil, the head of
   output/risul (parameter echo, soil table, initial heads).
9. bc_noflow_ref.npz, outlets_ref.npz, weil_ref.npz, spatial_atmbc_ref.npz,
   soil3d_ref.npz, nodata_ref.npz -- those projects run by the
   reference's FLOW3D with the coupled switching (SURF true: SWITCH in
   CONVER, ADRSTN at step starts and in BKSTEP), their own nansfneubc /
   sfbc / atmbc files read by the reference's BCONE / SFVONE / ATMONE
   (oracle/_ref/bc): per-step statistics for the steps the stored mbeconv
   holds, psi/Sw/IFATM/PONDNOD at selected steps.  weil ponds from its
   first step on (its surface starts saturated): its run takes the whole
   routing block of the time loop (oracle/refproject.py reads the project
   with the reference's DATIN, whose drainage network SURF_FLOWTRA routes)
   to TMAX = 7200 s.

Only data goes into the fixtures; no reference source text.
"""
from __future__ import annotations

import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"

MESH_CASES = [
    "examples/SSHydro/weill_exemple",
    "examples/SSHydro/atmbc_spatially_from_weill_withnodata",
    "examples/SSHydro/bc_noflow",
    "examples/SSHydro/meshing_from_weill",
    "examplesTmp/catchments/outlets",
    "examplesTmp/Luca/my_first_project4",
    "examples/DA/DA_with_non_uniform_ic",
    "pyCATHY/tests/data/simu_test_DA/tensiometers_DA",
]


def read_raster(path: str, header: bool = False):
    """GRASS-ASCII raster as rast_input_dem.f reads it (6 header lines)."""
    with open(path) as f:
        lines = f.read().splitlines()
    hdr = {}
    for ln in lines[:6]:
        k, v = ln.split(":", 1)
        hdr[k.strip()] = float(v)
    nrow, ncol = int(hdr["rows"]), int(hdr["cols"])
    vals = np.array(" ".join(lines[6:]).split(), dtype=np.float64)
    ras = vals[: nrow * ncol].reshape(nrow, ncol)
    return (ras, hdr) if header else ras


def read_dem_parameters(path: str) -> dict:
    toks = []
    for t in open(path).read().split():
        try:
            toks.append(float(t))
        except ValueError:
            break
    dx, dy, factor, dostep, nzone, nstr, n1, ivert, isp, base = toks[:10]
    nstr = int(nstr)
    return dict(dx=dx, dy=dy, factor=factor, dostep=int(dostep), nzone=int(nzone), nstr=nstr, n1=int(n1),
                ivert=int(ivert), isp=int(isp), base=base, zratio=np.array(toks[10:10 + nstr]))


def read_grid3d(path: str):
    with open(path) as f:
        nnod, n, nt = (int(v) for v in f.readline().split())
        tet = np.array([f.readline().split() for _ in range(nt)], dtype=np.int32)
        xyz = np.array([f.readline().split() for _ in range(n)], dtype=np.float64)
    return nnod, n, nt, tet, xyz


def read_nterm(path: str) -> int:
    if not os.path.exists(path):
        return -1
    with open(path, errors="replace") as f:
        for ln in f:
            m = re.search(r"NTERM \(# OF NONZERO TERMS IN SYS\. MAT\.\)\s*=\s*(\d+)", ln)
            if m:
                return int(m.group(1))
    return -1


def make_meshes(out: str) -> None:
    arrs = {}
    names = []
    for i, case in enumerate(MESH_CASES):
        d = os.path.join(REF, case)
        prm = read_dem_parameters(os.path.join(d, "input/dem_parameters"))
        dem, hdr = read_raster(os.path.join(d, "prepro/dem"), header=True)
        zp = os.path.join(d, "prepro/zone")
        zone = read_raster(zp).astype(np.int32) if os.path.exists(zp) else np.ones(dem.shape, np.int32)
        nnod, n, nt, tet, xyz = read_grid3d(os.path.join(d, "output/grid3d"))
        p = f"c{i}_"
        arrs[p + "dem"] = dem
        arrs[p + "zone"] = zone
        arrs[p + "zratio"] = prm.pop("zratio")
        arrs[p + "scalars"] = np.array([prm[k] for k in ("dx", "dy", "factor", "base")] + [hdr["west"], hdr["south"]])
        arrs[p + "ints"] = np.array([prm[k] for k in ("dostep", "nzone", "nstr", "n1", "ivert", "isp")] +
                                    [nnod, n, nt, read_nterm(os.path.join(d, "output/risul"))], dtype=np.int64)
        arrs[p + "tetra"] = tet
        arrs[p + "xyz"] = xyz
        names.append(case)
        print(f"{case}: nnod={nnod} n={n} nt={nt} nterm={arrs[p + 'ints'][-1]}")
    arrs["cases"] = np.array(names)
    np.savez_compressed(out, **arrs)


def make_hillslope_small(out: str, steps_saved=(1, 2, 5, 10, 25, 50, 100, 150)) -> None:
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from pycathy_wrapper_amd import configs as CF
    import pyoracle as P
    import refrun as R
    pr = CF.hillslope_small()
    rr = R.RefFlowRun(P.RefLib(), pr.mesh, pr.soil, pr.psi0, float(pr.atm_vals[0]), pr.params.as_dict())
    stats, psi, sw, saved = [], [], [], []
    keys = None
    k = 0
    while True:
        rc, s = rr.step()
        k += 1
        keys = keys or list(s)
        stats.append([s[nm] for nm in keys])
        if k in steps_saved or rc:
            saved.append(k)
            psi.append(rr.psi().copy())
            sw.append(rr.sw().copy())
        if rc:
            break
    np.savez_compressed(out, stat_names=np.array(keys), stats=np.array(stats), steps=np.array(saved),
                        psi=np.array(psi), sw=np.array(sw), rc_last=rc)
    print(f"hillslope_small: {k} steps, saved {saved}")


def make_hillslope_small_newton(out: str, nsteps: int = 60, steps_saved=(1, 5, 10, 20, 40, 60)) -> None:
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from pycathy_wrapper_amd import configs as CF
    import pyoracle as P
    import refrun as R
    pr = CF.hillslope_small(params=CF.hillslope_small_params(iopt=2))
    rr = R.RefFlowRun(P.RefLib(), pr.mesh, pr.soil, pr.psi0, float(pr.atm_vals[0]), pr.params.as_dict())
    stats, psi, sw, keys = [], [], [], None
    for k in range(1, nsteps + 1):
        rc, s = rr.step()
        keys = keys or list(s)
        stats.append([s[nm] for nm in keys])
        if k in steps_saved:
            psi.append(rr.psi().copy())
            sw.append(rr.sw().copy())
        if rc:
            break
    np.savez_compressed(out, stat_names=np.array(keys), stats=np.array(stats), steps=np.array(steps_saved),
                        psi=np.array(psi), sw=np.array(sw))
    print(f"hillslope_small newton: {len(stats)} steps")


def make_hillslope_bc(out: str, steps_saved=(1, 2, 10, 50, 100, 200, 300, 420, 440, 460, 476, 480)) -> None:
    """hillslope_bc (fixed heads, listed Neumann fluxes under free drainage,
    two-type root uptake; configs.hillslope_bc) stepped to its 3 h end by
    the reference's own FLOW3D with its own BCONE / BCNXT / BCBAK reading
    the nansfdirbc / nansfneubc files bcfiles.write_bc writes, NEUMANN and
    ETRAN, built with BC-sized bounds (oracle/_ref/bc)."""
    import tempfile
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pycathy_wrapper_amd import bcfiles as B
    from pycathy_wrapper_amd import configs as CF
    import pyoracle as P
    import refrun as R
    pr = CF.hillslope_bc()
    m = pr.mesh
    with tempfile.TemporaryDirectory() as d:
        paths = {k: os.path.join(d, k) for k in ("dir", "neu", "atmbc")}
        B.write_bc(paths["dir"], pr.bc_dir.file_records(m.nnod, m.nstr))
        B.write_bc(paths["neu"], pr.bc_neu.file_records(m.nnod, m.nstr))
        R.write_atmbc(paths["atmbc"], pr.atm_times, pr.atm_vals, 1)
        rr = R.RefFlowRun(P.RefLib(P.REF_BC_SO), m, pr.soil, pr.psi0, 0.0, pr.params.as_dict(),
                          atmbc=paths["atmbc"], dirbc=paths["dir"], neubc=paths["neu"], veg=pr.veg)
        stats, psi, sw, ifatm, saved, keys = [], [], [], [], [], None
        k = 0
        while True:
            rc, s = rr.step()
            k += 1
            keys = keys or list(s)
            stats.append([s[nm] for nm in keys])
            if k in steps_saved or rc:
                saved.append(k)
                psi.append(rr.psi().copy())
                sw.append(rr.sw().copy())
                ifatm.append(rr.a["IFATM"][:m.nnod].copy())
            if rc:
                break
        rr.close()
    np.savez_compressed(out, stat_names=np.array(keys), stats=np.array(stats), steps=np.array(saved),
                        psi=np.array(psi), sw=np.array(sw), ifatm=np.array(ifatm), rc_last=rc)
    print(f"hillslope_bc: {k} steps, saved {saved}")


def make_hillslope_seepage(out: str, steps_saved=(1, 2, 5, 10, 20, 40, 100, 200, 250, 280, 300)) -> None:
    """hillslope_seepage (two seepage faces, ISFCVG = 1; configs.
    hillslope_seepage) to its 2 h end by the reference's FLOW3D with its own
    SFVONE / SFVREC reading the sfbc file bcfiles.write_sfbc writes, SFINIT,
    EXTALL / EXTCVG and BKSTEP (oracle/_ref/bc).  The file repeats the
    faces in a second record at t = 1e30: SFVONE reads the next record's
    header at once, so the reference needs one beyond the run."""
    import tempfile
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pycathy_wrapper_amd import bcfiles as B
    from pycathy_wrapper_amd import configs as CF
    import pyoracle as P
    import refrun as R
    pr = CF.hillslope_seepage()
    m = pr.mesh
    faces = pr.seepage["faces"]
    with tempfile.TemporaryDirectory() as d:
        sf = os.path.join(d, "sfbc")
        B.write_sfbc(sf, [B.SFRecord(0.0, faces), B.SFRecord(1e30, faces)])
        lib = P.RefLib(P.REF_BC_SO)
        rr = R.RefFlowRun(lib, m, pr.soil, pr.psi0, float(pr.atm_vals[0]), pr.params.as_dict(), sfbc=sf,
                          isfcvg=pr.seepage["isfcvg"])
        fm = lib.nsfmax
        pos = [(j * fm + i) for i, f in enumerate(faces) for j in range(len(f))]
        stats, psi, sw, sfex, saved, keys = [], [], [], [], [], None
        k = 0
        while True:
            rc, s = rr.step()
            k += 1
            keys = keys or list(s)
            stats.append([s[nm] for nm in keys])
            if k in steps_saved or rc:
                saved.append(k)
                psi.append(rr.psi().copy())
                sw.append(rr.sw().copy())
                sfex.append(rr.a["SFEX"][pos].copy())
            if rc:
                break
        rr.close()
    np.savez_compressed(out, stat_names=np.array(keys), stats=np.array(stats), steps=np.array(saved),
                        psi=np.array(psi), sw=np.array(sw), sfex=np.array(sfex), rc_last=rc)
    print(f"hillslope_seepage: {k} steps, saved {saved}")


def make_hillslope_small_as_written(out: str) -> None:
    """hillslope_small with the workload's parm as written (DTMIN = 1 s,
    ITUNS/ITUNS1/ITUNS2 = 10/5/7): the reference's FLOW3D stops with NOBACK
    (DELTAT at DTMIN and ITUNS reached); every step and the final state."""
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pycathy_wrapper_amd import configs as CF
    import pyoracle as P
    import refrun as R
    pr = CF.hillslope_small(params=CF.Params(toluns=1e-6))
    rr = R.RefFlowRun(P.RefLib(), pr.mesh, pr.soil, pr.psi0, float(pr.atm_vals[0]), pr.params.as_dict())
    stats, keys = [], None
    while True:
        rc, s = rr.step()
        keys = keys or list(s)
        stats.append([s[nm] for nm in keys])
        if rc:
            break
    np.savez_compressed(out, stat_names=np.array(keys), stats=np.array(stats), psi=rr.psi(), sw=rr.sw(),
                        rc_last=rc)
    print(f"hillslope_small as written: {len(stats)} steps, rc {rc}")


def make_outputs(out: str, nblocks: int = 2) -> None:
    d = os.path.join(REF, "examples/SSHydro/weill_exemple/output")
    texts = {}
    for name in ("psi", "sw"):
        lines = open(os.path.join(d, name)).read().splitlines(keepends=True)
        heads = [i for i, ln in enumerate(lines) if "NSTEP" in ln and "TIME" in ln]
        end = heads[nblocks] if len(heads) > nblocks else len(lines)
        texts[name] = "".join(lines[:end])
    texts["mbeconv"] = open(os.path.join(d, "mbeconv")).read()
    np.savez_compressed(out, **{k: np.array(v) for k, v in texts.items()})
    print("weil outputs:", {k: len(v) for k, v in texts.items()})


PROJECTS = {
    "bc_noflow_project": "examplesTmp/SSHydro/bc_noflow",
    "outlets_project": "examplesTmp/catchments/outlets",
    "weil_project": "examples/SSHydro/weill_exemple",
    "spatial_atmbc_project": "examples/SSHydro/atmbc_spatially_from_weill",
    "soil3d_project": "examples/SSHydro/soil3d_from_weill",
    "nodata_project": "examples/SSHydro/atmbc_spatially_from_weill_withnodata",
    "mesh10x20_project": "examples/SSHydro/meshing_from_weill",
    "camporese_project": "examples/SSHydro/meshing_from_Camporese",
}
ROUTED = ("weil_project", "spatial_atmbc_project", "soil3d_project", "nodata_project", "mesh10x20_project",
          "camporese_project")
# projects whose stored output belongs to other inputs (meshing_from_weill's
# mbeconv starts from 165 m^3 of water, its inputs hold 1,100): no output kept
STALE_OUTPUT = ("mesh10x20_project",)
PROJECT_FILES = ("input/parm", "input/soil", "input/ic", "input/atmbc", "input/dem_parameters",
                 "input/nansfdirbc", "input/nansfneubc", "input/sfbc", "input/root_map", "prepro/dem",
                 "prepro/zone")
# the drainage network and reservoir files DATIN reads for ISIMGR = 2
# (the projects whose surface ponds)
ROUTING_FILES = ("input/posizione_serb", "input/livelli_iniz_s", "prepro/lakes_map", "prepro/qoi_a",
                 "prepro/dtm_w_1", "prepro/dtm_w_2", "prepro/dtm_p_outflow_1", "prepro/dtm_p_outflow_2",
                 "prepro/dtm_local_slope_1", "prepro/dtm_local_slope_2", "prepro/dtm_epl_1", "prepro/dtm_epl_2",
                 "prepro/dtm_kSs1_sf_1", "prepro/dtm_kSs1_sf_2", "prepro/dtm_Ws1_sf_1", "prepro/dtm_Ws1_sf_2",
                 "prepro/dtm_b1_sf", "prepro/dtm_y1_sf", "prepro/dtm_nrc")


def read_blocks(path: str):
    """The blocks of a psi / sw output file (detout.f): a header line
    'NSTEP TIME' then N values; returns [(nstep, time, values)]."""
    lines = open(path).read().splitlines()
    heads = [i for i, ln in enumerate(lines) if "NSTEP" in ln and "TIME" in ln]
    out = []
    for j, h in enumerate(heads):
        f = lines[h].split()
        end = heads[j + 1] if j + 1 < len(heads) else len(lines)
        vals = np.array(" ".join(lines[h + 1:end]).split(), dtype=np.float64)
        out.append((int(f[0]), float(f[1]), vals))
    return out


def copy_projects() -> None:
    """Item 8: the project directories and what the reference wrote."""
    import shutil
    for name, src in PROJECTS.items():
        dst = os.path.join(HERE, name)
        for rel in PROJECT_FILES + (ROUTING_FILES if name in ROUTED else ()):
            s = os.path.join(REF, src, rel)
            if os.path.exists(s):
                os.makedirs(os.path.dirname(os.path.join(dst, rel)), exist_ok=True)
                shutil.copyfile(s, os.path.join(dst, rel))
        if name in STALE_OUTPUT:
            continue
        os.makedirs(os.path.join(dst, "output"), exist_ok=True)
        if name == "weil_project":
            lines = open(os.path.join(REF, src, "output/risul")).read().splitlines(keepends=True)
            end = next(i for i, ln in enumerate(lines) if "TIME STEP:" in ln)
            with open(os.path.join(dst, "output/risul_head"), "w") as f:
                f.writelines(lines[:end])
            continue
        shutil.copyfile(os.path.join(REF, src, "output/mbeconv"), os.path.join(dst, "output/mbeconv"))
        arrs = {}
        for kind in ("psi", "sw"):
            blocks = read_blocks(os.path.join(REF, src, "output", kind))
            keep = {}
            for nstep, t, v in blocks:   # first block of each NSTEP (the last is written twice)
                if len(v) == len(blocks[0][2]):   # a block cut short (a run stopped while writing) is left out
                    keep.setdefault(nstep, (t, v))
            steps = sorted(keep)
            arrs[kind + "_nstep"] = np.array(steps)
            arrs[kind + "_time"] = np.array([keep[k][0] for k in steps])
            arrs[kind] = np.array([keep[k][1] for k in steps])
        np.savez_compressed(os.path.join(dst, "output/psi_sw.npz"), **arrs)
        print(f"{name}: psi blocks at steps {list(arrs['psi_nstep'])}")


def make_ponded_project() -> None:
    """weil_project started ponded: its input/ic set to INDP 2, IPOND 1 with
    a 2 cm ponding head everywhere (datin.f:377-403, icvhe.f: hydrostatic
    with the water table 2 cm above the surface), all else unchanged."""
    src, dst = os.path.join(HERE, "weil_project"), os.path.join(HERE, "ponded_project")
    if os.path.exists(dst):
        shutil.rmtree(dst)
    shutil.copytree(src, dst, ignore=shutil.ignore_patterns("output"))
    with open(os.path.join(dst, "input", "ic"), "w") as fh:
        fh.write("2 1 INDP IPOND\n0.02   PONDNOD\n")


def make_project_run(name: str, out: str, steps_saved=(1, 2, 5, 10, 50, 100, 200, 300), nsteps=10 ** 9) -> None:
    """Item 9: the project run by the reference's FLOW3D with the coupled
    switching, for as many steps as the stored mbeconv holds."""
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pycathy_wrapper_amd import project as PJ
    import pyoracle as P
    import refrun as R
    import refproject as RP
    root = os.path.join(HERE, name)
    pr = PJ.load_project(root)
    m = pr.mesh
    mbe = os.path.join(root, "output/mbeconv")
    nmb = sum(1 for ln in open(mbe) if not ln.startswith("#")) if os.path.exists(mbe) else nsteps
    io = lambda k: os.path.join(root, "input", k)
    ref = P.RefLib(P.REF_BC_SO)
    # the routing block where the project holds its drainage network
    routing = RP.RefProject(ref, root) if os.path.exists(os.path.join(root, "prepro", "qoi_a")) else None
    pond0 = pr.routing.get("pond0") if pr.routing is not None else None
    rr = R.RefFlowRun(ref, m, pr.soil, pr.psi0, 0.0, pr.params.as_dict(), atmbc=io("atmbc"),
                      dirbc=io("nansfdirbc"), neubc=io("nansfneubc"), veg=pr.veg, sfbc=io("sfbc"), surf=pr.surf,
                      routing=routing, pond0=pond0)
    stats, psi, sw, ifatm, pond, saved, keys = [], [], [], [], [], [], None
    k = 0
    while True:
        rc, s = rr.step()
        k += 1
        keys = keys or list(s)
        stats.append([s[nm] for nm in keys])
        last = rc or k == nmb
        if k in steps_saved or last:
            saved.append(k)
            psi.append(rr.psi().copy())
            sw.append(rr.sw().copy())
            ifatm.append(rr.a["IFATM"][:m.nnod].copy())
            pond.append(rr.a["PONDNOD"][:m.nnod].copy())
        if last:
            break
    rr.close()
    if routing is not None:
        routing.close()
    np.savez_compressed(out, stat_names=np.array(keys), stats=np.array(stats), steps=np.array(saved),
                        psi=np.array(psi), sw=np.array(sw), ifatm=np.array(ifatm), pondnod=np.array(pond),
                        rc_last=rc, routed=rr.routed if routing is not None else 0)
    print(f"{name}: {k} steps, rc {rc}, saved {saved}")


if __name__ == "__main__":
    make_meshes(os.path.join(HERE, "meshes.npz"))
    make_hillslope_small(os.path.join(HERE, "hillslope_small_ref.npz"))
    make_hillslope_small_newton(os.path.join(HERE, "hillslope_small_newton_ref.npz"))
    make_outputs(os.path.join(HERE, "weil_outputs.npz"))
    make_hillslope_bc(os.path.join(HERE, "hillslope_bc_ref.npz"))
    make_hillslope_seepage(os.path.join(HERE, "hillslope_seepage_ref.npz"))
    make_hillslope_small_as_written(os.path.join(HERE, "hillslope_small_as_written_ref.npz"))
    copy_projects()
    make_project_run("bc_noflow_project", os.path.join(HERE, "bc_noflow_ref.npz"))
    make_project_run("outlets_project", os.path.join(HERE, "outlets_ref.npz"),
                     steps_saved=(1, 2, 5, 10, 50, 100, 200, 300, 400, 500, 600, 700))
    make_project_run("weil_project", os.path.join(HERE, "weil_ref.npz"),
                     steps_saved=(1, 2, 10, 50, 100, 179, 235))
    make_project_run("mesh10x20_project", os.path.join(HERE, "mesh10x20_ref.npz"),
                     steps_saved=(1, 2, 10, 50, 100, 195))
    make_project_run("camporese_project", os.path.join(HERE, "camporese_ref.npz"),
                     steps_saved=(1, 2, 10, 50, 100, 200, 235))
    make_ponded_project()
    make_project_run("ponded_project", os.path.join(HERE, "ponded_ref.npz"), steps_saved=(1, 2, 5, 10, 20, 40, 60),
                     nsteps=60)
    make_project_run("spatial_atmbc_project", os.path.join(HERE, "spatial_atmbc_ref.npz"),
                     steps_saved=(1, 2, 10, 100, 500, 1027))
    make_project_run("soil3d_project", os.path.join(HERE, "soil3d_ref.npz"), steps_saved=(1, 2, 10, 20, 40))
    make_project_run("nodata_project", os.path.join(HERE, "nodata_ref.npz"), steps_saved=(1, 2, 10, 50, 100, 180))
