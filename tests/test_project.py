"""pyCATHY project directories read as the reference reads them
(pycathy_wrapper_amd/project.py), and the coupled atmospheric switching of
ISIMGR = 2 runs, pinned on what the reference program wrote for the same
projects (tests/golden/*_project/output: data files the example projects
hold).

1. weil_exemple: every parameter of input/parm and input/soil that DATIN
   echoes to output/risul (datin.f:489-620), the soil table row by row,
   N / NT / NTERM of the mesh built from prepro/dem and input/dem_parameters,
   and the initial heads INITAL printed (ICVHE, 4 significant digits) equal
   the reader's values.
2. bc_noflow (20x20 cells, one inactive, 15 layers, 440 listed Neumann
   nodes) and catchments/outlets (a 256-node catchment with inactive cells,
   15 layers): the meshes have the reference's node / tetrahedron counts
   (output/grid3d headers: 440 7040 35910 and 256 4096 20250).
3. The reference's FLOW3D run on those projects with the coupled switching
   (SWITCH in CONVER, ADRSTN at step starts and back-steps; oracle/refrun.py
   with surf=True, golden bc_noflow_ref.npz / outlets_ref.npz), and on
   atmbc_spatially_from_weill (per-node rainfall, 1027 steps) and
   soil3d_from_weill (400 soil zones, the 40 steps its output holds) and
   atmbc_spatially_from_weill_withnodata (a DEM with no-data cells, 180
   steps, ponded at 15 of them) with the routing block too
   (spatial_atmbc_ref.npz, soil3d_ref.npz, nodata_ref.npz), equals the
   reference program's own stored mbeconv at every stored step -- DELTAT,
   TIME, the Picard count NLIN, STORE1/2, DSTORE, VIN, VOUT, the mass-balance
   error and their cumulative sums -- to the digits printed, and psi / Sw
   equal output/psi and output/sw where those hold a block.  So the reader,
   the mesh and the coupled switching reproduce the reference program end
   to end on its own example inputs, with no ponding (OVL. FLUX is zero
   throughout both runs' hgatmsf).  Where oracle/_ref/bc is built, the
   bc_noflow golden is re-derived and must be identical.  weil_exemple
   ponds from its first step on: with the routing block of the time loop
   (the reference's SURF_FLOWTRA and PONDUPD on the project as its DATIN
   reads it) its 235 steps equal the stored mbeconv, psi and sw.
4. Refusals: a coupled project that starts ponded (IPOND != 0; inital.f:201
   sets PONDING, so step 1 would route); ISIMGR = 3; NLRELX != 0.
"""
from __future__ import annotations

import os
import re
import shutil
import sys

import numpy as np
import pytest

from pycathy_wrapper_amd import project as PJ

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)
REF_BC_SO = os.path.join(ROOT, "oracle", "_ref", "bc", "libcathy_ref_bc.so")

# mbeconv columns (cathy_main.f:4043): the golden statistic each one holds
MBE_COLS = {1: "DELTAT", 2: "TIME", 3: "ITER", 5: "STORE1", 6: "STORE2", 7: "DSTORE", 8: "CDSTOR", 9: "VIN",
            10: "CVIN", 11: "VOUT", 12: "CVOUT", 13: "VINOUT", 14: "VTOT", 15: "ERRAS", 17: "CERRAS", 18: "CAERAS"}


def _echo(path: str) -> dict:
    """NAME -> printed value for the 'NAME (description) = value' lines of
    risul (a description may run over several lines before its '=')."""
    out, name = {}, None
    for ln in open(path).read().splitlines():
        m = re.match(r"^\s{5}([A-Z][A-Z0-9_]*)\s*\(", ln)
        if m:
            name = m.group(1)
        if "=" in ln and name:
            out[name] = ln.split("=", 1)[1].strip()
            name = None
    return out


def _printed_close(ours: float, token: str, slack: float = 1e-9) -> bool:
    """``ours`` rounds to the printed ``token`` (E format): within half a
    unit of its last printed digit, plus a relative slack."""
    mant, exp = token.upper().split("E")
    digits = len(mant.replace("-", "").replace(".", ""))
    unit = 10.0 ** (int(exp) - (digits - 1))
    return abs(ours - float(token)) <= 0.5 * unit * (1 + 1e-6) + slack * abs(float(token))


def test_weil_parameters_match_reference_echo():
    root = os.path.join(GOLD, "weil_project")
    echo = _echo(os.path.join(root, "output", "risul_head"))
    parm = PJ.read_parm(os.path.join(root, "input", "parm"))
    for key in ("ISIMGR", "IPRT1", "IPRT", "NPRT", "NUMVP", "NR", "KSLOPE", "ISFONE", "ISFCVG", "LUMP", "IOPT",
                "NLRELX", "L2NORM", "ITUNS", "ITUNS1", "ITUNS2", "ITMXCG"):
        assert parm[key.lower()] == int(echo[key]), key
    # PONDH_MIN is echoed as read (datin.f sets 1e10 only for ISIMGR <= 1; weil has ISIMGR 2)
    for key in ("PONDH_MIN", "TOLKSL", "TETAF", "TOLUNS", "TOLSWI", "ERNLMX", "TOLCG", "DELTAT", "DTMIN", "DTMAX",
                "TMAX", "DTMAGA", "DTMAGM", "DTREDS", "DTREDM"):
        assert _printed_close(parm[key.lower()], echo[key], 0.0), (key, parm[key.lower()], echo[key])
    dp = PJ.read_dem_parameters(os.path.join(root, "input", "dem_parameters"))
    soil = PJ.read_soil(os.path.join(root, "input", "soil"), dp["nstr"], dp["nzone"])
    assert _printed_close(soil["pmin"], echo["PMIN"], 0.0)
    assert soil["ipeat"] == int(echo["IPEAT"])
    # the soil table (layer, zone, PERMX PERMY PERMZ ELSTOR POROS VGN VGRMC VGPSAT)
    lines = open(os.path.join(root, "output", "risul_head")).read().splitlines()
    h = next(i for i, ln in enumerate(lines) if "X-PERM" in ln and "VGPSAT" in ln)
    keys = ("permx", "permy", "permz", "elstor", "poros", "vgn", "vgrmc", "vgpsat")
    rows = 0
    for ln in lines[h + 1:]:
        f = ln.split()
        if len(f) != 10:
            break
        lay, zone = int(f[0]) - 1, int(f[1]) - 1
        for k, tok in zip(keys, f[2:]):
            assert _printed_close(soil["table"][k][lay, zone], tok, 0.0), (lay, zone, k)
        rows += 1
    assert rows == dp["nstr"] * dp["nzone"]
    pr = PJ.load_project(root)
    m = pr.mesh
    from pycathy_wrapper_amd import mesh as M
    assert m.n == int(echo["N"]) and m.nt == int(echo["NT"])
    assert len(M.sym_pattern(m.tetra, m.n)[1]) == int(echo["NTERM"])
    # INITIAL PRESSURE HEAD: N pairs (node, value) at 1PE10.3
    h = next(i for i, ln in enumerate(lines) if "INITIAL PRESSURE HEAD" in ln)
    toks = " ".join(lines[h + 1:h + 1 + (m.n + 3) // 4]).split()
    nodes, vals = np.array(toks[0::2], dtype=int), toks[1::2]
    assert np.array_equal(nodes, np.arange(1, m.n + 1))
    assert all(_printed_close(pr.psi0[i], v, 0.0) for i, v in enumerate(vals))
    # hydrostatic from the surface: every node saturated, so the initial
    # water volume is the pore volume (risul: INITIAL VOLUME OF WATER)
    assert np.all(pr.psi0 >= 0.0)
    pores = np.zeros(m.n)
    np.add.at(pores, m.tetra[:, :4].ravel(), np.repeat(m.volu / 4.0, 4) *
              np.repeat(pr.soil["poros"][m.layer_of_element(), m.tetra[:, 4] - 1], 4))
    vol = [ln for ln in lines if "INITIAL VOLUME OF WATER" in ln][0].split("=")[1].strip()
    assert _printed_close(float(pores.sum()), vol, 0.0)
    assert pr.surf and pr.params.pondh_min == 0.0


@pytest.mark.parametrize("name,sizes", [("bc_noflow_project", (440, 7040, 35910)),
                                        ("outlets_project", (256, 4096, 20250))])
def test_project_meshes_match_reference_grid3d(name, sizes):
    pr = PJ.load_project(os.path.join(GOLD, name))
    assert (pr.mesh.nnod, pr.mesh.n, pr.mesh.nt) == sizes
    assert pr.surf and pr.params.pondh_min == 0.0 and pr.bc_neu is not None and pr.veg is not None
    assert pr.bc_dir is None and pr.seepage is None


def _mbeconv(name):
    rows = []
    for ln in open(os.path.join(GOLD, name, "output", "mbeconv")):
        if not ln.startswith("#") and ln.strip():
            rows.append(ln.split())
    return rows


@pytest.mark.parametrize("name", ["bc_noflow", "outlets", "spatial_atmbc", "soil3d", "nodata", "camporese"])
def test_coupled_runs_equal_reference_program_outputs(name):
    gold = np.load(os.path.join(GOLD, f"{name}_ref.npz"))
    col = {nm: j for j, nm in enumerate(str(v) for v in gold["stat_names"])}
    stats = gold["stats"]
    rows = _mbeconv(f"{name}_project")
    assert len(stats) == len(rows)
    for k, row in enumerate(rows):
        assert int(row[0]) == k + 1
        for c, key in MBE_COLS.items():
            v = stats[k, col[key]]
            if key == "ITER":
                assert int(v) == int(row[c]), (k + 1, key)
            else:
                assert _printed_close(v, row[c]), (k + 1, key, v, row[c])
        # bc_noflow and outlets never pond (their routing never runs);
        # spatial_atmbc and soil3d pond from their first step on and route
        # every later one; nodata ponds now and then
        if name in ("bc_noflow", "outlets", "spatial_atmbc", "soil3d", "camporese"):
            assert stats[k, col["PONDING"]] == (0 if name in ("bc_noflow", "outlets") else 1)
    if name in ("spatial_atmbc", "soil3d", "camporese"):
        assert int(gold["routed"]) == len(stats) - 1
    if name == "nodata":
        npond = int(stats[:-1, col["PONDING"]].sum())
        assert 0 < npond < len(stats) - 1 and int(gold["routed"]) >= npond
    out = np.load(os.path.join(GOLD, f"{name}_project", "output", "psi_sw.npz"))
    steps = list(gold["steps"])
    hit = 0
    for j, ns in enumerate(out["psi_nstep"]):
        if ns in steps:
            i = steps.index(ns)
            assert out["psi_time"][j] == pytest.approx(stats[ns - 1, col["TIME"]], rel=1e-8)
            for kind in ("psi", "sw"):   # 1PE15.6: 7 significant digits
                b = out[kind][j]
                assert np.all(np.abs(gold[kind][i] - b) <= 5.0000001e-7 * np.abs(b) + 1e-300), (ns, kind)
            hit += 1
    assert hit >= (0 if name == "soil3d" else 1)   # soil3d's stored run printed no block after step 0


def _blocks(text: str):
    """(NSTEP, values) of each block of a psi / sw output text."""
    lines = text.splitlines()
    heads = [i for i, ln in enumerate(lines) if "NSTEP" in ln and "TIME" in ln]
    out = []
    for j, h in enumerate(heads):
        end = heads[j + 1] if j + 1 < len(heads) else len(lines)
        out.append((int(lines[h].split()[0]), np.array(" ".join(lines[h + 1:end]).split(), dtype=np.float64)))
    return out


def test_weil_coupled_run_equals_reference_program():
    """weil_exemple starts saturated to the surface (ICVHE), so ATMONE makes
    every surface node IFATM 2 under PONDH_MIN = 0, the first step's SWITCH
    ponds them and every later step starts with the routing block: the
    overland flow routed through the project's drainage network
    (SURF_FLOWTRA) and the ponded heads set by PONDUPD.  The reference's
    routines run that loop here on the project as DATIN reads it
    (weil_ref.npz: 235 steps to TMAX = 7200 s, 234 of them routed), and it
    equals the reference program's stored mbeconv at every step and its
    psi / sw at step 179, to the digits printed."""
    gold = np.load(os.path.join(GOLD, "weil_ref.npz"))
    col = {nm: j for j, nm in enumerate(str(v) for v in gold["stat_names"])}
    stats = gold["stats"]
    assert int(gold["rc_last"]) == 1 and int(gold["routed"]) == len(stats) - 1
    out = np.load(os.path.join(GOLD, "weil_outputs.npz"))
    rows = [ln.split() for ln in str(out["mbeconv"]).splitlines() if not ln.startswith("#") and ln.strip()]
    assert len(rows) == len(stats) == 235
    for k, row in enumerate(rows):
        for c, key in MBE_COLS.items():
            v = stats[k, col[key]]
            if key == "ITER":
                assert int(v) == int(row[c]), (k + 1, key)
            else:
                assert _printed_close(v, row[c]), (k + 1, key, v, row[c])
    assert np.all(stats[:, col["PONDING"]] == 1)
    steps = list(gold["steps"])
    for kind in ("psi", "sw"):
        blocks = dict(_blocks(str(out[kind])))
        b = blocks[179]
        a = gold[kind][steps.index(179)]
        assert np.all(np.abs(a - b) <= 5.0000001e-7 * np.abs(b) + 1e-300), kind
    assert np.any(gold["pondnod"][-1] > 0.0)


@pytest.mark.parametrize("name", ["weil_project", "mesh10x20_project", "camporese_project"])
def test_reader_matches_reference_datin(name):
    """project.load_project on weil_exemple, meshing_from_weill (a 10 x 20
    DEM) and meshing_from_Camporese against the reference's own
    DATIN reading the same files (oracle/refproject.py): every parm and
    soil value DATIN returns, the surface mesh (TRIANG, X, Y, Z), the
    vegetation types, and the drainage network and cell numbering of the
    routing block (INDEX, INDEX_WITH_LAKES, CELL, QOI_SN, the DTM grids as
    COMMON /RIVERNETWORK/ holds them, DELTA_X / DELTA_Y)."""
    if not os.path.exists(REF_BC_SO):
        pytest.skip("oracle/_ref/bc not built (make -C oracle ref-bc)")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    import refproject as RP
    root = os.path.join(GOLD, name)
    rp = RP.RefProject(P.RefLib(P.REF_BC_SO), root)
    try:
        pr = PJ.load_project(root)
        m, prm, parm = pr.mesh, pr.params, pr.meta["parm"]
        for key in ("DELTAT", "DTMIN", "DTMAX", "TMAX", "DTMAGA", "DTMAGM", "DTREDS", "DTREDM", "TOLUNS",
                    "TOLSWI", "ERNLMX", "TOLCG", "TETAF", "PONDH_MIN"):
            assert getattr(prm, key.lower()) == rp.scalar(key), key
        for key in ("ITUNS", "ITUNS1", "ITUNS2", "ITMXCG", "LUMP", "IOPT", "L2NORM"):
            assert getattr(prm, key.lower()) == rp.scalar(key), key
        for key in ("ISIMGR", "KSLOPE", "ISFCVG", "NLRELX", "IPRT1", "NPRT"):
            assert parm[key.lower()] == rp.scalar(key), key
        assert (rp.scalar("NNOD"), rp.scalar("NTRI"), rp.scalar("NSTR"), rp.scalar("N"), rp.scalar("NT")) == \
            (m.nnod, m.ntri, m.nstr, m.n, m.nt)
        ms = rp.ref.maxstr
        for key in ("PERMX", "PERMY", "PERMZ", "ELSTOR", "POROS", "VGNCELL", "VGRMCCELL", "VGPSATCELL"):
            tab = rp.a[key][:ms * rp.scalar("NZONE")].reshape(-1, ms).T[:m.nstr]
            mine = pr.soil[{"VGNCELL": "vgn", "VGRMCCELL": "vgrmc", "VGPSATCELL": "vgpsat"}.get(key, key.lower())]
            assert np.array_equal(mine, tab), key
        tri = rp.a["TRIANG"][:4 * m.ntri].reshape(m.ntri, 4)
        assert np.array_equal(np.sort(tri[:, :3], axis=1) - 1, m.triang[:, :3]) and np.array_equal(tri[:, 3],
                                                                                                    m.triang[:, 3])
        for key, v in (("X", m.x), ("Y", m.y), ("Z", m.z)):
            assert np.array_equal(rp.a[key][:m.nnod], v[:m.nnod]), key
        assert np.array_equal(rp.a["VEG_TYPE"][:m.nnod], pr.veg["veg_type"])
        r = pr.routing
        assert r is not None and (r["nrow"], r["ncol"]) == (rp.scalar("NROW"), rp.scalar("NCOL"))
        assert np.array_equal(r["index"], rp.grid("INDEX")) and np.array_equal(r["index_wl"],
                                                                                rp.grid("INDEX_WITH_LAKES"))
        nc = rp.scalar("NCELL_COARSE")
        assert np.array_equal(r["cell"] + 1, rp.a["CELL"][:5 * nc].reshape(nc, 5)[:, :4])
        rn = rp.rivernetwork()
        net = r["net"]
        assert (net.dx, net.dy) == (rn["DELTA_X"], rn["DELTA_Y"]) and r["ncell"] == rp.scalar("NCELL")
        assert np.array_equal(net.qoi, rn["QOI_SN"])
        names = dict(w1="DTM_W_1", w2="DTM_W_2", slope1="DTM_LOCAL_SLOPE_1", slope2="DTM_LOCAL_SLOPE_2",
                     epl1="DTM_EPL_1", epl2="DTM_EPL_2", kss1="DTM_KSS1_SF_1", kss2="DTM_KSS1_SF_2",
                     ws1="DTM_WS1_SF_1", ws2="DTM_WS1_SF_2", b1="DTM_B1_SF", y1="DTM_Y1_SF", nrc="DTM_NRC",
                     pout1="DTM_P_OUTFLOW_1", pout2="DTM_P_OUTFLOW_2")
        for k, name in names.items():   # Network holds them as stored: row J from the south
            assert np.array_equal(getattr(net, k), rn[name][::-1]), k
    finally:
        rp.close()


def test_bc_noflow_golden_rederived():
    if not os.path.exists(REF_BC_SO):
        pytest.skip("oracle/_ref/bc not built (make -C oracle ref-bc)")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    import refrun as R
    root = os.path.join(GOLD, "bc_noflow_project")
    pr = PJ.load_project(root)
    io = lambda k: os.path.join(root, "input", k)
    gold = np.load(os.path.join(GOLD, "bc_noflow_ref.npz"))
    rr = R.RefFlowRun(P.RefLib(P.REF_BC_SO), pr.mesh, pr.soil, pr.psi0, 0.0, pr.params.as_dict(),
                      atmbc=io("atmbc"), dirbc=io("nansfdirbc"), neubc=io("nansfneubc"), veg=pr.veg,
                      sfbc=io("sfbc"), surf=True)
    try:
        for k in range(len(gold["stats"])):
            rc, s = rr.step()
            assert np.array_equal(np.array(list(s.values()), dtype=np.float64), gold["stats"][k]), k + 1
        assert rc == int(gold["rc_last"])
        assert np.array_equal(rr.psi(), gold["psi"][-1])
    finally:
        rr.close()


def test_project_refusals(tmp_path):
    src = os.path.join(GOLD, "bc_noflow_project")
    root = tmp_path / "prj"
    shutil.copytree(src, root)
    ic = root / "input" / "ic"
    ic.write_text("0\t1\tINDP IPOND\n-1.5e+01\n0.0\n")
    with pytest.raises(ValueError, match="IPOND"):
        PJ.load_project(str(root))
    pr = PJ.load_project(str(root), subsurface_only=True)   # runs as ISIMGR = 1
    assert not pr.surf and pr.params.pondh_min == 1e10
    ic.write_text("0\t0\tINDP IPOND\n-1.5e+01\n")
    parm = (root / "input" / "parm").read_text().splitlines()
    parm[1] = "3 0.0 0\tISIMGR PONDH_MIN VELREC"
    (root / "input" / "parm").write_text("\n".join(parm) + "\n")
    with pytest.raises(ValueError, match="ISIMGR = 3"):
        PJ.load_project(str(root))
    parm[1] = "2 0.0 0\tISIMGR PONDH_MIN VELREC"
    parm[7] = "1 0.8\tNLRELX OMEGA"
    (root / "input" / "parm").write_text("\n".join(parm) + "\n")
    with pytest.raises(ValueError, match="NLRELX"):
        PJ.load_project(str(root))


def test_ponded_start_reader_and_reference():
    """The project started ponded (INDP 2, IPOND 1, 2 cm): the reader keeps
    PONDNOD as the reference's DATIN reads it and forms the initial head as
    ICVHE does (water table 2 cm above the surface); the reference harness,
    started with PONDING set, re-derives the first steps of its golden run
    (tests/golden/ponded_ref.npz)."""
    if not os.path.exists(REF_BC_SO):
        pytest.skip("oracle/_ref/bc not built (make -C oracle ref-bc)")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    import refproject as RP
    import refrun as R
    root = os.path.join(GOLD, "ponded_project")
    pr = PJ.load_project(root)
    m = pr.mesh
    rp = RP.RefProject(P.RefLib(P.REF_BC_SO), root)
    try:
        assert rp.scalar("IPOND") == 1 and rp.scalar("INDP") == 2
        assert np.array_equal(pr.routing["pond0"], rp.a["PONDNOD"][:m.nnod])
        assert np.allclose(pr.psi0[:m.nnod], 0.02, rtol=0, atol=1e-12)
        assert np.allclose(pr.psi0, np.tile(m.z[:m.nnod] + 0.02, m.nstr + 1) - m.z, rtol=0, atol=1e-12)
        io = lambda k: os.path.join(root, "input", k)
        gold = np.load(os.path.join(GOLD, "ponded_ref.npz"))
        names = [str(v) for v in gold["stat_names"]]
        rr = R.RefFlowRun(rp.ref, m, pr.soil, pr.psi0, 0.0, pr.params.as_dict(), atmbc=io("atmbc"),
                          dirbc=io("nansfdirbc"), neubc=io("nansfneubc"), veg=pr.veg, sfbc=io("sfbc"), surf=True,
                          routing=rp, pond0=pr.routing["pond0"])
        for k in range(3):
            rc, st = rr.step()
            assert [st[nm] for nm in names] == list(gold["stats"][k]), k
        assert rr.routed >= 3
        rr.close()
    finally:
        rp.close()


def test_raster_rows_read_as_the_reference_reads_them(tmp_path):
    """RAST_INPUT_* read one row per list-directed READ: a row starts on a
    new line, may continue over the next ones, and drops what its last line
    holds beyond the row; a file that ends early is an error (the reference
    reads past its end)."""
    hdr = "north: 0\nsouth: 0\neast: 0\nwest: 0\nrows: 2\ncols: 3\n"
    p = tmp_path / "r"
    p.write_text(hdr + "1 2 3 99 98\n4\n5 6 97\n")
    r, h = PJ.read_raster(str(p))
    assert r.tolist() == [[1, 2, 3], [4, 5, 6]] and h["rows"] == 2
    p.write_text(hdr + "1 2 3 4 5 6\n")
    with pytest.raises(ValueError, match="ends at row 2"):
        PJ.read_raster(str(p))
