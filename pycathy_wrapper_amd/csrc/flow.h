// flow.h -- one mesh's subsurface flow run on the device (Flow): the
// device-resident state, FLOW3D's Picard loop (flow3d.f / picard.f) with the
// IC(0)-PCG of solscal-extended.f, and the per-time-step host logic of
// cathy_main.f, split into the phases a lockstep ensemble interleaves.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "cathy_gpu.h"
#include "flow_kernels.h"
#include "guess.h"
#include "route_link.h"

namespace cg {

// ------------------------------------------------------------- schedule --
struct Schedule {
  std::vector<ColDesc> cols;     // column ordering: colour classes
  std::vector<ListLevel> lv;     // natural ordering: DAG levels
  DBuf<int> rows;                // row lists of the natural levels
};

// Level schedule of the natural elimination order on a full CSR (sorted
// columns): level(r) = 1 + max level of its lower neighbours.
inline void natural_levels(int n, const std::vector<int> &rowptr, const std::vector<int> &col,
                           Schedule &sch, hipStream_t st) {
  std::vector<int> lev(n, 0);
  int nlev = 0;
  for (int r = 0; r < n; ++r) {
    int l = 0;
    for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
      int j = col[p];
      if (j >= r) break;
      l = std::max(l, lev[j] + 1);
    }
    lev[r] = l;
    nlev = std::max(nlev, l + 1);
  }
  std::vector<int> cnt(nlev + 1, 0), rows(n);
  for (int r = 0; r < n; ++r) cnt[lev[r] + 1]++;
  for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int r = 0; r < n; ++r) rows[fill[lev[r]]++] = r;
  sch.lv.clear();
  for (int l = 0; l < nlev; ++l) sch.lv.push_back(ListLevel{cnt[l], cnt[l + 1] - cnt[l]});
  sch.rows.upload(rows, st);
}

// --------------------------------------------------------------- params --
struct Params {
  int iopt = 1, l2norm = 0, ituns = 10, ituns1 = 5, ituns2 = 7, itmxcg = 500, ieto = 0;
  int free_drainage = 1, hspatm = 1;
  double tetaf = 1.0, toluns = 1e-6, tolswi = 1e30, ernlmx = 1e30, tolcg = 1e-10;
  double deltat = 10.0, dtmin = 1.0, dtmax = 3600.0, tmax = 86400.0;
  double dtmaga = 0.0, dtmagm = 1.1, dtreds = 0.0, dtredm = 0.5;
  double pmin = -5.0, pondh_min = 0.0, scf = 0.0;
};

// ------------------------------------------------ listed-node BC series --
// A non-atmospheric Dirichlet or Neumann series (input/nansfdirbc,
// input/nansfneubc): records of (time, node rows, values) in file order and
// the three-record window BCONE / BCNXT / BCBAK keep over them (bcone.f,
// bcnxt.f, bcbak.f).  The window only moves forward; the active record
// (ANBC, ABCNOD, BC) is the window's second record while the third lies
// ahead, else the third; BCBAK (a back-step to a time at or before the
// second record's) takes the first.
struct BcSeries {
  std::vector<double> t;      // record times
  std::vector<int> off;       // record k: entries off[k] .. off[k+1]
  std::vector<int> row;       // device rows
  std::vector<int> node;      // CATHY ids (0-based)
  std::vector<double> val;
  int next = 0, hti = 0, act = -1;
  double tim[3] = {0, 0, 0};
  int rec[3] = {-1, -1, -1};  // record held in each window slot (-1: none)
  bool set() const { return !t.empty(); }
  int count(int k) const { return k < 0 ? 0 : off[k + 1] - off[k]; }
  int max_count() const {
    int m = 0;
    for (size_t k = 0; k + 1 < off.size(); ++k) m = std::max(m, off[k + 1] - off[k]);
    return m;
  }
  // bcnxt.f:30-41: shift the window until TIME <= its third time
  void advance_window(double time) {
    while (!(time <= tim[2])) {
      tim[0] = tim[1]; tim[1] = tim[2];
      rec[0] = rec[1]; rec[1] = rec[2];
      if (next >= (int)t.size()) { hti = 1; break; }
      tim[2] = t[next];
      rec[2] = next++;
    }
    act = tim[2] > tim[1] ? rec[1] : rec[2];  // bcnxt.f:42-54
  }
  void bcone(double time) {  // bcone.f:20-60
    next = 0; hti = 0;
    tim[0] = tim[1] = 0.0;
    rec[0] = rec[1] = -1;
    tim[2] = t[0];
    rec[2] = 0;
    next = 1;
    advance_window(time);
  }
  void bcnxt(double time) {
    if (hti) return;
    advance_window(time);
  }
  void bcbak() {  // bcbak.f:14-21
    if (tim[0] >= tim[1]) return;
    act = rec[0];
  }
  // bkstep.f:73-88: forward again past the window's second time, else back
  void back(double time) {
    if (time > tim[1]) bcnxt(time);
    else bcbak();
  }
};

// ------------------------------------------------------------------ Flow --
struct Flow {
  int n = 0, nt = 0, ntri = 0, nnod = 0, nstr = 0, nnz = 0, ordering = 0, lump = 1;
  // the 3-colour column ordering; its rows in tiles of row_tile() chains
  // per colour (0: one tile per colour, plane by plane)
  bool column_order() const {
    return ordering == CATHY_GPU_ORDER_COLUMN3 || ordering == CATHY_GPU_ORDER_COLUMN3C ||
           ordering == CATHY_GPU_ORDER_COLUMN3T;
  }
  int row_tile() const {
    return ordering == CATHY_GPU_ORDER_COLUMN3C ? 1 : (ordering == CATHY_GPU_ORDER_COLUMN3T ? kChains : 0);
  }
  bool f32 = true;  // IC(0) factor precision
  // PCG SpMV reads the lower triangle through the transpose (column
  // ordering with valid chain tables): one member alone runs faster, since
  // the smaller HBM stream leaves the last-level cache to the sweeps; several
  // members sharing the GPU run slower, so it is off unless asked for.
  bool spmv_transposed = false;
  hipStream_t st = nullptr;
  std::vector<int> perm, iperm;  // CATHY id -> device id, device -> CATHY

  // mesh / pattern (device numbering)
  DBuf<int> so, scol, slen, sdiag, stpos;  // SELL-64 pattern
  int nnz_sell = 0, maxrow = 0;
  DBuf<int> plan_off, plan_ie;  // lumped assembly plan
  DBuf<uint32_t> plan_io;
  DBuf<double4> plan_s4;
  DBuf<double> plan_w2, plan_gw;
  bool use_plan = false;
  DBuf<int> pair_ptr;  // IC(0) update pattern per SELL slot
  DBuf<int2> pair_rc;
  DBuf<uint32_t> ftab;  // IC(0) chain tables (column ordering), see FacTab
  int ftab_nch = 0;
  bool ftab_ok = false;
  DBuf<double> dbeta;
  DBuf<int> inc_ptr, inc, refid, row_surf, surf_row, bot_row;
  DBuf<uint32_t> incoff;
  DBuf<int4> enode;
  DBuf<double> S, gz, volu, m4d, m4full, lmass, arenod;
  DBuf<double> psat, vgn, vgm, vgn1, vgnr, vgmr, vgpsn, pnot, rmcp, snodi, pnodi, volnod, kznod;
  Schedule sched;

  // state
  DBuf<double> pnew, pold, ptnew, ptold, ptimep, pdiff, sw, ckrw, ckrwp, et1, et2, swnew, swtimep;
  DBuf<double> ckrwe, et1e, A, tnoti, xt5, bcflux, qnode, qbot;
  DBuf<uint8_t> isdir;
  DBuf<int> ifatm, ifatmp;
  DBuf<double> atmpot, atmact, atmold;
  // listed-node boundary conditions: fixed heads (CONTP/PRESC, QPNEW and
  // QPOLD by list position), listed Neumann fluxes (CONTQ/Q beyond the base
  // nodes of free drainage), root uptake (ETRAN's QTRANIE) -- all optional
  BcSeries bc_dir, bc_neu;
  DBuf<uint8_t> bcfix;                       // per row: 2 fixed head, 4 seepage exit
  DBuf<int> dir_rows, neu_rows, atm_off;     // current lists (rows), IFATM = -1 surface ids
  DBuf<double> dir_vals, neu_vals, qpnew, qpold;
  std::vector<int> h_dir_rows, h_neu_rows;   // rows marked / fluxes added by the last apply
  int np_cur = 0, nx_cur = 0, n_atm_off = 0;
  double nx_in = 0.0, nx_out = 0.0;          // sums of the listed Neumann fluxes (fluxmb.f:39-46)
  bool veg_on = false;
  // coupled surface (ISIMGR = 2): SWITCH in CONVER and ADRSTN at step
  // starts and back-steps; pond is SWITCH's PONDING flag on the device
  bool surf = false;
  DBuf<int> pond;
  // The routing block of a coupled run (cathy_main.f:2993-3058): the
  // overland flux OVFLNOD to cells (NOD_CELL), to routing order
  // (TRANSFER_F3D_SURF), ROUTE + ALTEZZE substeps on the linked routing
  // handle, the depths back to cells (TRANSFER_SURF_F3D) and nodes
  // (CELL_NOD) as PONDNOD, then PONDUPD; with the back-step and step-end
  // copies of the routing state (cathy_main.f:3145-3157, 3770-3782,
  // bkstep.f:148-157).  DOSTEP = 1, no lakes or reservoirs.
  struct Routing {
    RouteLink link;
    int nrow = 0, ncol = 0, ngrid = 0, ncoarse = 0, npair = 0;
    double dx = 0.0, dy = 0.0;
    DBuf<int> corner, target, pair_dst, pair_src, csrc, nptr, nlist, err;
    DBuf<double> ovflnod, ovflp, pondnod, ovflcel, pondcel, coarse;
    // [10][ngrid]: SURFACE_WATER_SN, Q_IN_KK_SN, Q_OUT_KK_SN_1/2, VOLUME_KK_SN,
    // Q_IN_KKP1_SN, Q_OUT_KKP1_SN_1/2, VOLUME_KKP1_SN, H_WATER_KKP1_SN
    DBuf<double> state;
    DBuf<double> sav, pst;  // [4][ngrid]: the _SAV and _P copies of the KK arrays
    DBuf<double> cu_ak;     // CU_MAX, AK_MAX
    double ak_max = 0.0, ak_max_sav = 0.0, ak_max_p = 0.0;
    bool ak_fresh = false;  // AK_MAX on the device is newer (routed this step)
    int pond_host = 0, pondp = 0;
    int routed = 0;         // routing blocks run (all attempts)
  };
  std::unique_ptr<Routing> rt;
  DBuf<int> vegc, permd;
  DBuf<double> zc, vegtab, qtr;
  // seepage faces (sfbc): nodes face by face, top to bottom; SFEX / SFEXIT
  // / SFEXP and SFQ / SFQP by node position (cathy_main.f:1085-1100)
  std::vector<int> sf_node, sf_off;
  DBuf<int> sf_rows, sfex, sfexit, sfexp;
  DBuf<double> sfq, sfqp;
  int nsf_nodes = 0, isfcvg = 0, ksf = 0;
  double sfflw_cur = 0.0, sfflwp = 0.0;
  bool lists_on() const { return bc_dir.set() || bc_neu.set() || nsf_nodes > 0; }
  DBuf<float> F32;
  DBuf<double> F64, dinv;
  DBuf<double> r, p, Bv, z;
  DBuf<double> part;  // partial sums, kSlots slots of kMaxRedBlocks
  DBuf<int> ipart;
  DBuf<PcgScal> scal;
  DBuf<int> flags;  // [0] PCG done, [1] IC(0) breakdown count
  DBuf<double> res;
  std::vector<double> h_res;
  // what the host polls, in pinned memory: the solver's done flag and
  // scalars, MASBAL's reduced values, single sums
  struct HostBox {
    int done;
    PcgScal pcg;
    BicgScal bicg;
    double res[24];
    double v;
  };
  Pinned<HostBox> hbox;

  // host time / balance state (cathy_main.f)
  Params prm;
  std::vector<double> atm_t, atm_v;  // series
  int atm_w = 0;                     // values per record of the series (1 or NNOD, from HSPATM when set)
  int atm_next = 0, htiatm = 0;
  double atmtim[3] = {0, 0, 0};
  std::vector<double> atminp;  // 3 x (1 or nnod) rates
  double time = 0, timep = 0, deltat = 0;
  int dtgmin = 1, nstep = 1, iter = 1, nitert = 0, itlin = 0, itrtot = 0, klsfai = 0;
  double store1 = 0, store2 = 0, dstore = 0, cdstor = 0, cvin = 0, cvout = 0, vtot = 0, cerras = 0,
         caeras = 0, vin = 0, vout = 0, erras = 0, errel = 0;
  double adinp = 0, adoutp = 0, ndinp = 0, ndoutp = 0, aninp = 0, anoutp = 0, nninp = 0, nnoutp = 0;
  double cur[8] = {0};  // adin adout ndin ndout anin anout nnin nnout of the last MASBAL
  double pinf = 0, pl2 = 0, finf = 0, fl2 = 0, pikmax = 0;
  int ikmax = 0, lsfail = 0, noback = 0, bckstp = 0, finished = 0, initialised = 0;
  // counters
  double c_picard = 0, c_pcg = 0, c_fall = 0, c_solve_ms = 0, c_dev_ms = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // SpMV launch timing (HIP events on this stream around k_spmv_dot2)
  bool prof = false;
  std::vector<hipEvent_t> pev;
  int pev_used = 0;
  double spmv_ms = 0.0;
  long spmv_launches = 0;

  ~Flow() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : pev) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }

  hipEvent_t next_event() {
    if (pev_used == (int)pev.size()) {
      hipEvent_t e;
      CG_HIP(hipEventCreate(&e));
      pev.push_back(e);
    }
    return pev[pev_used++];
  }
  // after a stream sync: fold recorded (start, stop) pairs into spmv_ms
  void collect_events() {
    for (int i = 0; i + 1 < pev_used; i += 2) {
      float ms = 0.f;
      CG_HIP(hipEventElapsedTime(&ms, pev[i], pev[i + 1]));
      spmv_ms += ms;
      spmv_launches += 1;
    }
    pev_used = 0;
  }

  SoilSoA soil() const {
    return SoilSoA{psat.p, vgn.p, vgm.p, vgn1.p, vgnr.p, vgmr.p, vgpsn.p, pnot.p, rmcp.p, snodi.p, pnodi.p};
  }
  // partial-sum slots: 0 XLUNG, 1 (p,r), 2 (p,Ap), 3 r_free^2, 4 (Ap,z),
  // 5-10 norms/DSTORE, 11-16 boundary fluxes, 17 storage
  // 0-4 and 18 also serve the BiCGSTAB of the Newton path (no PCG there)
  static constexpr int kSlots = 19;
  double *P(int slot) { return part.p + (size_t)slot * kMaxRedBlocks; }
  int red_grid(size_t work) const { return grid_for(work, kBlock, kMaxRedBlocks); }
  Sell sell() const {
    Sell s{};
    s.so = so.p;
    s.col = scol.p;
    s.len = slen.p;
    s.diag = sdiag.p;
    s.code = ccode.p;
    s.clen = cclen.p;
    s.cdiag = ccdiag.p;
    s.coded = coded;
    s.nchain = nnod;
    s.width = cwidth;
    s.nplane = nstr + 1;
    s.ncolr = cncolr;
    s.ctw = row_tile();
    for (int c = 0; c < 4; ++c) {
      s.cbase[c] = ccbase[c];
      s.cnc[c] = ccnc[c];
      s.cgoff[c] = ccgoff[c];
    }
    return s;
  }

  // Coded column source (see Sell in flow_kernels.h), derived from the
  // explicit pattern and required to reproduce it row by row; if any row
  // disagrees the explicit column ids stay in use.
  DBuf<uint32_t> ccode;
  DBuf<uint8_t> cclen, ccdiag;
  int coded = 0, cwidth = 0, cncolr = 0;
  int ccbase[4] = {0, 0, 0, 0}, ccnc[4] = {1, 1, 1, 1}, ccgoff[4] = {0, 0, 0, 0};
  void build_codes() {
    coded = 0;
    if (!column_order()) return;
    const int ncl = (int)sched.cols.size();
    if (ncl < 1 || ncl > 4 || nnod >= (1 << 27) || maxrow > 255) return;
    for (int c = 0; c < ncl; ++c) {
      ccbase[c] = sched.cols[c].base;
      ccnc[c] = sched.cols[c].nc;
      ccgoff[c] = sched.cols[c].g0;
    }
    const int nch = nnod, W = maxrow, np = nstr + 1;
    const uint32_t unset = 0xffffffffu;
    std::vector<uint32_t> code((size_t)3 * W * nch, unset);
    std::vector<int> clen((size_t)3 * nch, -1), cdiag((size_t)3 * nch, -1);
    auto decode = [&](int r, int &c, int &k, int &t) {
      c = 0;
      for (int i = 1; i < ncl; ++i)
        if (r >= ccbase[i]) c = i;
      const int off = r - ccbase[c];
      if (row_tile()) {
        tile_decode(off, ccnc[c], row_tile(), np, k, t);
      } else {
        k = off / ccnc[c];
        t = off - k * ccnc[c];
      }
    };
    for (int r = 0; r < n; ++r) {
      int c, k, t;
      decode(r, c, k, t);
      const size_t g = (size_t)ccgoff[c] + t;
      const int pt = k == 0 ? 0 : (k == np - 1 ? 2 : 1);
      const int len = h_rowptr[r + 1] - h_rowptr[r];
      int dg = -1;
      for (int j = 0; j < len; ++j)
        if (h_col[h_rowptr[r] + j] == r) dg = j;
      int &L = clen[pt * nch + g], &D = cdiag[pt * nch + g];
      if (L < 0) {
        L = len;
        D = dg;
      } else if (L != len || D != dg) {
        return;
      }
      for (int j = 0; j < W; ++j) {
        uint32_t w;
        if (j < len) {
          int c2, k2, t2;
          decode(h_col[h_rowptr[r] + j], c2, k2, t2);
          const int dk = k2 - k;
          if (dk < -1 || dk > 1) return;
          w = (uint32_t)t2 | ((uint32_t)(dk + 1) << 27) | ((uint32_t)c2 << 29);
        } else {
          w = (uint32_t)t | (1u << 27) | ((uint32_t)c << 29);  // padding: the row itself
        }
        uint32_t &slot = code[((size_t)pt * W + j) * nch + g];
        if (slot == unset) slot = w;
        else if (slot != w) return;
      }
    }
    std::vector<uint8_t> l8(clen.size()), d8(cdiag.size());
    for (size_t i = 0; i < clen.size(); ++i) {
      l8[i] = (uint8_t)std::max(0, clen[i]);
      d8[i] = (uint8_t)std::max(0, cdiag[i]);
    }
    for (uint32_t &w : code)
      if (w == unset) w = 1u << 27;  // plane types without rows (never read)
    ccode.upload(code, st);
    cclen.upload(l8, st);
    ccdiag.upload(d8, st);
    CG_HIP(hipStreamSynchronize(st));
    cwidth = W;
    cncolr = ncl;
    coded = 1;
    // the device decode must reproduce every explicit column id
    DBuf<int> bad(1);
    CG_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_check_codes, dim3(grid_for(n)), dim3(kBlock), 0, st, n, sell(), bad.p);
    int hb = 0;
    bad.download(&hb, 1, st);
    CG_HIP(hipStreamSynchronize(st));
    if (hb != 0) coded = 0;
    code_mismatch = hb;
  }
  int code_mismatch = 0;
  PairList pairs() const { return PairList{pair_ptr.p, pair_rc.p}; }

  // ------------------------------------------------------------ set-up --
  void setup(int n_, int nt_, int ntri_, int nnod_, int nstr_, int nzone, int maxstr, const int *tetra,
             const double *bi, const double *ci, const double *di, const double *volu_, const int *ivol,
             const double *volnod_, const double *arenod_, const double *permx, const double *permy,
             const double *permz, const double *snodi_, const double *pnodi_, const double *vgn_,
             const double *vgrmc_, const double *vgpsat_, const double *kznod_, int lump_, int ordering_) {
    n = n_; nt = nt_; ntri = ntri_; nnod = nnod_; nstr = nstr_; lump = lump_; ordering = ordering_;
    nzone_ = nzone;
    maxstr_ = maxstr;
    CG_REQUIRE(n > 0 && nt > 0 && nnod > 0 && nstr > 0, "empty mesh");
    CG_REQUIRE((long)nnod * (nstr + 1) == n, "N must equal NNOD*(NSTR+1) (gen3d.f numbering)");
    CG_REQUIRE((long)3 * ntri * nstr == nt, "NT must equal 3*NTRI*NSTR (gen3d.f)");
    CG_REQUIRE(nzone >= 1 && maxstr >= nstr, "NZONE/MAXSTR");
    CG_REQUIRE(ordering == CATHY_GPU_ORDER_NATURAL || column_order(), "ORDERING");
    CG_REQUIRE((long)n <= (long)kMaxRedBlocks * kBlock, "N above the per-launch row limit");
    CG_REQUIRE(!column_order() || (nstr + 1) <= 64, "column ordering needs NSTR < 64");
    f32 = column_order();
    CG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    CG_HIP(hipEventCreate(&ev0));
    CG_HIP(hipEventCreate(&ev1));
    for (long e = 0; e < (long)nt * 5; e += 5) {
      for (int k = 0; k < 4; ++k)
        CG_REQUIRE(tetra[e + k] >= 1 && tetra[e + k] <= n, "TETRA node id out of range");
      CG_REQUIRE(tetra[e + 4] >= 1 && tetra[e + 4] <= nzone, "TETRA zone out of range");
    }
    build_ordering(tetra);
    build_pattern(tetra);
    build_elements(tetra, bi, ci, di, volu_, ivol, permx, permy, permz, maxstr, pnodi_);
    build_nodes(snodi_, pnodi_, vgn_, vgrmc_, vgpsat_, volnod_, kznod_, arenod_);
    alloc_state();
    CG_HIP(hipStreamSynchronize(st));
  }

  // SYMSLV set-up: the pattern of a 1-based upper CSR with the diagonal
  // first in each row (strpic.f layout), natural order, FP64 factor.
  // csr_map[q] = COEF1 index of SELL slot q (-1 = padding).
  DBuf<int> csr_map;
  int nterm = 0;
  void setup_csr(int n_, int nterm_, const int *topol, const int *ja) {
    n = n_;
    nterm = nterm_;
    nt = 0;
    nnod = 0;
    ordering = CATHY_GPU_ORDER_NATURAL;
    f32 = false;
    CG_REQUIRE(n > 0 && nterm > 0, "empty system");
    CG_REQUIRE((long)n <= (long)kMaxRedBlocks * kBlock, "N above the per-launch row limit");
    CG_REQUIRE(topol[0] == 1 && topol[n] - 1 == nterm, "TOPOL(1) must be 1 and TOPOL(N+1)-1 = NTERM");
    for (int r = 0; r < n; ++r) {
      CG_REQUIRE(topol[r + 1] > topol[r], "TOPOL not increasing");
      CG_REQUIRE(ja[topol[r] - 1] == r + 1, "JA: the diagonal must come first in each row");
      for (int q = topol[r]; q < topol[r + 1] - 1; ++q)
        CG_REQUIRE(ja[q] > r + 1 && ja[q] <= n, "JA: upper-triangle column out of range");
    }
    CG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    CG_HIP(hipEventCreate(&ev0));
    CG_HIP(hipEventCreate(&ev1));
    perm.resize(n);
    iperm.resize(n);
    for (int i = 0; i < n; ++i) perm[i] = iperm[i] = i;
    build_pattern(nullptr, topol, ja);
    std::vector<int> hmap(nnz_sell, -1);
    auto find = [&](int r, int c) {  // COEF1 index of (r, c), r <= c
      for (int q = topol[r] - 1; q < topol[r + 1] - 1; ++q)
        if (ja[q] - 1 == c) return q;
      return -1;
    };
    for (int r = 0; r < n; ++r)
      for (int k = h_rowptr[r]; k < h_rowptr[r + 1]; ++k) {
        const int c = h_col[k];
        const int q = c >= r ? find(r, c) : find(c, r);
        CG_REQUIRE(q >= 0, "pattern entry missing from the upper CSR");
        hmap[h_so[r >> 6] + 64 * (k - h_rowptr[r]) + (r & 63)] = q;
      }
    csr_map.upload(hmap, st);
    alloc_solver();
    coef.alloc(nterm);
    CG_HIP(hipStreamSynchronize(st));
  }
  DBuf<double> coef;
  std::vector<uint8_t> h_isdir;

  // SYMSLV (solscal-extended.f:4669-4699): INCLDP + PRODDP + GRADDP +
  // RESSYM on COEF1 x = TNOTI.  Rows NODDIR(1:NUMDIR) carry the
  // 1e-9*RMAX penalty diagonal in the reference, which decouples them to
  // within ~1e-90 relative; here they are masked out of the factor and
  // the products, and return x = 0.
  int symslv(int numdir, const int *noddir, int itmxcg, double tolcg, double rmin, const double *tnoti_h,
             const double *coef1, double *pnew_h, int *niter_out, double *err_out, double *res_out) {
    h_isdir.assign(n, 0);
    for (int k = 0; k < numdir; ++k) {
      CG_REQUIRE(noddir[k] >= 1 && noddir[k] <= n, "NODDIR entry out of range");
      h_isdir[noddir[k] - 1] = 1;
    }
    isdir.upload(h_isdir.data(), n, st);
    coef.upload(coef1, nterm, st);
    tnoti.upload(tnoti_h, n, st);
    hipLaunchKernelGGL(k_gather_coef, dim3(grid_for(nnz_sell)), dim3(kBlock), 0, st, nnz_sell, csr_map.p, coef.p,
                       A.p);
    double err = 0.0;
    const int niter = pcg(tolcg, itmxcg, &err, last_pcg > 0 ? last_pcg : 8);
    last_pcg = niter;
    // RESSYM: sqrt(sum_free (b - A x)^2 / XLUNG), or / N when XLUNG <= RMIN
    const int G = std::min(grid_for(n, kBlock), 1024);
    hipLaunchKernelGGL(k_residual0, dim3(G), dim3(kBlock), 0, st, n, sell(), A.p, pdiff.p, tnoti.p, isdir.p,
                       r.p, P(0));
    hipLaunchKernelGGL(k_sumsq_free, dim3(G), dim3(kBlock), 0, st, n, r.p, isdir.p, P(1));
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, st, G, P(0), res.p);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, st, G, P(1), res.p + 1);
    double hr[2];
    res.download(hr, 2, st);
    pdiff.download(pnew_h, n, st);
    CG_HIP(hipStreamSynchronize(st));
    *niter_out = niter;
    *err_out = err;
    *res_out = hr[0] <= rmin ? std::sqrt(hr[1] / n) : std::sqrt(hr[1] / hr[0]);
    return niter >= itmxcg ? CATHY_GPU_LSFAIL : 0;
  }

  // Surface 3-colouring by triangle propagation (a triangulated disc with
  // even interior degrees is 3-colourable and the colouring is forced);
  // greedy fallback otherwise.
  std::vector<int> surface_colours(const int *tetra, int &ncol) {
    std::vector<std::array<int, 3>> tris;
    for (int e = 0; e < 3 * ntri; ++e) {  // layer-1 tets with 3 surface nodes
      int v[4], m = 0;
      for (int k = 0; k < 4; ++k)
        if (tetra[5 * e + k] - 1 < nnod) v[m++] = tetra[5 * e + k] - 1;
      if (m == 3) tris.push_back({v[0], v[1], v[2]});
    }
    std::vector<std::vector<int>> adj(nnod);
    std::map<std::pair<int, int>, std::vector<int>> edge_tris;
    for (int t = 0; t < (int)tris.size(); ++t)
      for (int a = 0; a < 3; ++a) {
        int u = tris[t][a], w = tris[t][(a + 1) % 3];
        adj[u].push_back(w);
        adj[w].push_back(u);
        edge_tris[{std::min(u, w), std::max(u, w)}].push_back(t);
      }
    std::vector<int> c(nnod, -1);
    bool ok = true;
    std::vector<char> seen(tris.size(), 0);
    for (int t0 = 0; t0 < (int)tris.size() && ok; ++t0) {
      if (seen[t0]) continue;
      std::vector<int> stack{t0};
      seen[t0] = 1;
      // seed the first triangle of a component
      int used = 0;
      for (int a = 0; a < 3; ++a)
        if (c[tris[t0][a]] >= 0) used |= 1 << c[tris[t0][a]];
      for (int a = 0; a < 3; ++a)
        if (c[tris[t0][a]] < 0) {
          int k = 0;
          while (used & (1 << k)) ++k;
          c[tris[t0][a]] = k;
          used |= 1 << k;
        }
      while (!stack.empty() && ok) {
        int t = stack.back();
        stack.pop_back();
        int cols = 0;
        for (int a = 0; a < 3; ++a) cols |= 1 << c[tris[t][a]];
        if (cols != 7) { ok = false; break; }
        for (int a = 0; a < 3; ++a) {
          int u = tris[t][a], w = tris[t][(a + 1) % 3];
          for (int t2 : edge_tris[{std::min(u, w), std::max(u, w)}]) {
            if (seen[t2]) continue;
            seen[t2] = 1;
            for (int b = 0; b < 3; ++b) {
              int x = tris[t2][b];
              if (x == u || x == w) continue;
              int need = 7 & ~((1 << c[u]) | (1 << c[w]));
              int want = need == 1 ? 0 : need == 2 ? 1 : 2;
              if (c[x] < 0) c[x] = want;
              else if (c[x] != want) ok = false;
            }
            stack.push_back(t2);
          }
        }
      }
    }
    for (int s = 0; s < nnod && ok; ++s)
      for (int w : adj[s])
        if (c[s] == c[w] || c[s] < 0) ok = false;
    if (!ok) {  // greedy colouring (any graph)
      std::fill(c.begin(), c.end(), -1);
      for (int s = 0; s < nnod; ++s) {
        uint64_t used = 0;
        for (int w : adj[s])
          if (c[w] >= 0 && c[w] < 64) used |= 1ull << c[w];
        int k = 0;
        while (used & (1ull << k)) ++k;
        c[s] = k;
      }
    }
    ncol = 0;
    for (int s = 0; s < nnod; ++s) ncol = std::max(ncol, c[s] + 1);
    return c;
  }

  void build_ordering(const int *tetra) {
    perm.assign(n, 0);
    iperm.assign(n, 0);
    sched.lv.clear();
    sched.cols.clear();
    if (ordering == CATHY_GPU_ORDER_NATURAL) {
      for (int i = 0; i < n; ++i) perm[i] = iperm[i] = i;
      return;
    }
    int ncolr = 0;
    std::vector<int> c = surface_colours(tetra, ncolr);
    std::vector<int> cnt(ncolr, 0), rank(nnod);
    for (int s = 0; s < nnod; ++s) rank[s] = cnt[c[s]]++;
    std::vector<long> base(ncolr + 1, 0);
    for (int k = 0; k < ncolr; ++k) base[k + 1] = base[k] + (long)cnt[k] * (nstr + 1);
    // a colour's rows plane by plane (level-major), or in tiles of
    // row_tile() chains, each tile plane by plane
    std::vector<ColDesc> cd(ncolr);
    for (int k = 0; k < ncolr; ++k)
      cd[k] = ColDesc{(int)base[k], std::max(cnt[k], 1), nstr + 1, 0, 0, 0, row_tile() ? row_tile() : std::max(cnt[k], 1)};
    for (int s = 0; s < nnod; ++s)
      for (int k = 0; k <= nstr; ++k) {
        const long ref = (long)k * nnod + s;
        const long dev = col_row(cd[c[s]], rank[s], k);
        perm[ref] = (int)dev;
        iperm[dev] = (int)ref;
      }
    int g0 = 0;
    for (int k = 0; k < ncolr; ++k)
      if (cnt[k]) {
        ColDesc C = cd[k];
        C.g0 = g0;
        sched.cols.push_back(C);
        g0 += cnt[k];
      }
  }

  // IC(0) chain tables (FacTab): the bytes of each (class, chain) come from
  // the class's first row; every row is then checked against its own
  // pattern, and any mismatch leaves the pair-list factor in use.
  void build_factor_tables(const std::vector<int> &hrp, const std::vector<int> &hcol, const std::vector<int> &hsd,
                           const std::vector<int> &htpos) {
    ftab_ok = false;
    if (!column_order() || sched.cols.empty()) return;
    int nch = 0;
    for (const ColDesc &C : sched.cols) nch = std::max(nch, C.g0 + C.nc);
    constexpr int NB = kFacWords * 4;
    auto pos_in = [&](int row, int col) -> int {
      const int *b = &hcol[hrp[row]], *e = &hcol[hrp[row + 1]];
      const int *f = std::lower_bound(b, e, col);
      return (f != e && *f == col) ? (int)(f - b) : -1;
    };
    // the bytes row r (plane k of a colour whose planes lie sk rows apart) needs
    auto row_bytes = [&](int r, int k, int sk, uint8_t *out) -> bool {
      std::fill(out, out + NB, (uint8_t)0xFF);
      const int jd = hsd[r], jp = k > 0 ? jd - 1 : jd;
      if (jd > 16) return false;
      for (int j = 1; j < jp; ++j) {
        const int c = hcol[hrp[r] + j];
        for (int a = 0; a < j; ++a) {
          const int b = pos_in(c, hcol[hrp[r] + a]);
          if (b >= 0) out[j * (j - 1) / 2 + a] = (uint8_t)b;
        }
      }
      if (k > 0) {
        const int pred = r - sk;
        if (hcol[hrp[r] + jd - 1] != pred) return false;
        for (int a = 0; a < jp; ++a) {
          const int b = pos_in(pred, hcol[hrp[r] + a]);
          if (b >= 0) out[120 + a] = (uint8_t)b;
        }
      }
      for (int j = 0; j < jd; ++j) out[136 + j] = (uint8_t)(htpos[hrp[r] + j] - hrp[hcol[hrp[r] + j]]);
      return true;
    };
    std::vector<uint8_t> tb((size_t)5 * nch * NB, 0xFF), tmp(NB);
    for (const ColDesc &C : sched.cols) {
      if (C.len * kChains > 256) return;  // the table kernel holds one row per thread
      for (int cls = 0; cls < 5; ++cls) {
        int krep = -1;
        for (int k = 0; k < C.len && krep < 0; ++k)
          if (fac_class(k, C.len) == cls) krep = k;
        if (krep < 0) continue;
        for (int t = 0; t < C.nc; ++t)
          if (!row_bytes(col_row(C, t, krep), krep, col_w(C, t), &tb[((size_t)cls * nch + C.g0 + t) * NB])) return;
      }
    }
    for (const ColDesc &C : sched.cols)
      for (int k = 0; k < C.len; ++k) {
        const int cls = fac_class(k, C.len);
        for (int t = 0; t < C.nc; ++t) {
          if (!row_bytes(col_row(C, t, k), k, col_w(C, t), tmp.data())) return;
          if (std::memcmp(tmp.data(), &tb[((size_t)cls * nch + C.g0 + t) * NB], NB) != 0) return;
        }
      }
    std::vector<uint32_t> w((size_t)5 * kFacWords * nch);
    for (int cls = 0; cls < 5; ++cls)
      for (int g = 0; g < nch; ++g) {
        const uint8_t *b = &tb[((size_t)cls * nch + g) * NB];
        for (int q = 0; q < kFacWords; ++q)
          w[((size_t)cls * kFacWords + q) * nch + g] = (uint32_t)b[4 * q] | ((uint32_t)b[4 * q + 1] << 8) |
                                                        ((uint32_t)b[4 * q + 2] << 16) | ((uint32_t)b[4 * q + 3] << 24);
      }
    ftab.upload(w, st);
    ftab_nch = nch;
    ftab_ok = true;
  }

  // Full pattern in device numbering, from the tetrahedra (tetra != null)
  // or from a 1-based upper CSR (topol, ja) in natural order.
  void build_pattern(const int *tetra, const int *topol = nullptr, const int *ja = nullptr) {
    const int W = kMaxRow;
    std::vector<int> slot((size_t)n * W), cnt(n, 0);
    auto ins = [&](int r, int c) {
      int *s = &slot[(size_t)r * W];
      int m = cnt[r], pos = 0;
      while (pos < m && s[pos] < c) ++pos;
      if (pos < m && s[pos] == c) return;
      CG_REQUIRE(m < W, "row longer than kMaxRow in the full pattern");
      std::memmove(s + pos + 1, s + pos, sizeof(int) * (m - pos));
      s[pos] = c;
      cnt[r] = m + 1;
    };
    if (tetra) {
      for (int e = 0; e < nt; ++e) {
        int v[4];
        for (int k = 0; k < 4; ++k) v[k] = perm[tetra[5 * e + k] - 1];
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) ins(v[a], v[b]);
      }
    } else {
      for (int r = 0; r < n; ++r) {
        ins(r, r);
        for (int q = topol[r] - 1; q < topol[r + 1] - 1; ++q) {
          const int c = ja[q] - 1;
          ins(r, c);
          ins(c, r);
        }
      }
    }
    std::vector<int> hrp(n + 1, 0);
    for (int r = 0; r < n; ++r) hrp[r + 1] = hrp[r] + cnt[r];
    nnz = hrp[n];
    std::vector<int> hcol(nnz), hdiag(n), htpos(nnz);
    for (int r = 0; r < n; ++r)
      for (int k = 0; k < cnt[r]; ++k) {
        hcol[hrp[r] + k] = slot[(size_t)r * W + k];
        if (slot[(size_t)r * W + k] == r) hdiag[r] = hrp[r] + k;
      }
    for (int r = 0; r < n; ++r)
      for (int p = hrp[r]; p < hrp[r + 1]; ++p) {
        int j = hcol[p];
        const int *b = &hcol[hrp[j]], *e = &hcol[hrp[j + 1]];
        const int *f = std::lower_bound(b, e, r);
        CG_REQUIRE(f != e && *f == r, "pattern not symmetric");
        htpos[p] = hrp[j] + (int)(f - b);
      }
    // SELL-64: slice s holds rows 64s..64s+63, entry j of row r at
    // so[s] + 64 j + (r & 63); padding = own row id, zero value.
    const int ns = (n + 63) / 64;
    std::vector<int> hso(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
      int w = 0;
      for (int r = 64 * s; r < std::min(n, 64 * s + 64); ++r) w = std::max(w, cnt[r]);
      CG_REQUIRE((long)hso[s] + 64L * w < (1L << 31), "matrix too large for 32-bit slots");
      hso[s + 1] = hso[s] + 64 * w;
    }
    nnz_sell = hso[ns];
    auto spos = [&](int r, int j) { return hso[r >> 6] + 64 * j + (r & 63); };
    std::vector<int> hsc(nnz_sell), hst(nnz_sell, -1), hlen(n), hsd(n);
    for (int s = 0; s < ns; ++s)  // padding slots point at their own row
      for (int q = hso[s]; q < hso[s + 1]; ++q) hsc[q] = std::min(n - 1, 64 * s + ((q - hso[s]) & 63));
    for (int r = 0; r < n; ++r) {
      hlen[r] = hrp[r + 1] - hrp[r];
      hsd[r] = hdiag[r] - hrp[r];
    }
    // gather counts of the column sweeps per colour: slots before the
    // diagonal less the chain predecessor, after it less the successor
    for (ColDesc &C : sched.cols) {
      C.nf = C.nb = 0;
      for (int k = 0; k < C.len; ++k)
        for (int t = 0; t < C.nc; ++t) {
          const int r = col_row(C, t, k);
          C.nf = std::max(C.nf, hsd[r] - (k > 0 ? 1 : 0));
          C.nb = std::max(C.nb, hlen[r] - hsd[r] - 1 - (k < C.len - 1 ? 1 : 0));
        }
    }
    for (int r = 0; r < n; ++r) {
      for (int p = hrp[r]; p < hrp[r + 1]; ++p) {
        const int j = p - hrp[r];
        hsc[spos(r, j)] = hcol[p];
        const int c = hcol[p];
        hst[spos(r, j)] = spos(c, htpos[p] - hrp[c]);
      }
    }
    // IC(0) update pattern: for each lower slot (r,j), c = col(r,j), the
    // slot pairs (a in row r, b in row c) with col(r,a) == col(c,b) < c
    maxrow = 0;
    for (int r = 0; r < n; ++r) maxrow = std::max(maxrow, hlen[r]);
    std::vector<int> hpp((size_t)nnz_sell + 1, 0);
    std::vector<int2> hpr;
    hpr.reserve((size_t)n * 16);
    {
      std::vector<std::vector<int2>> tmp;  // per slot, filled in slot order below
      std::vector<int> cntq(nnz_sell, 0);
      std::vector<std::pair<int, int2>> all;
      all.reserve((size_t)n * 16);
      for (int r = 0; r < n; ++r)
        for (int j = 0; j < hsd[r]; ++j) {
          const int c = hcol[hrp[r] + j];
          int a = 0, b = 0;
          while (a < j && b < hsd[c]) {
            const int ca = hcol[hrp[r] + a], cb = hcol[hrp[c] + b];
            if (ca < cb) ++a;
            else if (ca > cb) ++b;
            else {
              all.push_back({spos(r, j), make_int2(spos(r, a), spos(c, b))});
              ++a;
              ++b;
            }
          }
        }
      std::stable_sort(all.begin(), all.end(),
                       [](const std::pair<int, int2> &u, const std::pair<int, int2> &v) { return u.first < v.first; });
      for (const auto &e : all) hpp[e.first + 1]++;
      for (int q = 0; q < nnz_sell; ++q) hpp[q + 1] += hpp[q];
      hpr.resize(all.size());
      for (size_t m = 0; m < all.size(); ++m) hpr[m] = all[m].second;
    }
    pair_ptr.upload(hpp, st);
    if (hpr.empty()) hpr.push_back(make_int2(0, 0));
    pair_rc.upload(hpr, st);
    build_factor_tables(hrp, hcol, hsd, htpos);
    so.upload(hso, st);
    scol.upload(hsc, st);
    slen.upload(hlen, st);
    sdiag.upload(hsd, st);
    stpos.upload(hst, st);
    if (ordering == CATHY_GPU_ORDER_NATURAL) natural_levels(n, hrp, hcol, sched, st);
    // every level row must have its same-column neighbours where the column
    // kernels expect them: predecessor = last lower slot, successor = first
    // upper slot
    if (column_order())
      for (const ColDesc &C : sched.cols)
        for (int k = 0; k < C.len; ++k)
          for (int t = 0; t < C.nc; ++t) {
            const int r = col_row(C, t, k), w = col_w(C, t);
            if (k > 0) CG_REQUIRE(hsd[r] >= 1 && hsc[spos(r, hsd[r] - 1)] == r - w, "column predecessor slot");
            if (k + 1 < C.len)
              CG_REQUIRE(hsd[r] + 1 < hlen[r] && hsc[spos(r, hsd[r] + 1)] == r + w, "column successor slot");
            for (int j = 0; j < hlen[r]; ++j) {  // no coupling inside a colour class but the chain
              const int c = hsc[spos(r, j)];
              if (c >= C.base && c < C.base + C.nc * C.len && c != r)
                CG_REQUIRE(c == r - w || c == r + w, "colour class not independent");
            }
          }
    h_rowptr = std::move(hrp);
    h_col = std::move(hcol);
    h_so = std::move(hso);
    if (tetra) build_codes();
  }
  std::vector<int> h_rowptr, h_col, h_so;

  void build_elements(const int *tetra, const double *bi, const double *ci, const double *di,
                      const double *volu_, const int *ivol, const double *permx, const double *permy,
                      const double *permz, int maxstr, const double *pnodi_) {
    std::vector<int4> hen(nt);
    std::vector<double> hS((size_t)nt * 10), hgz((size_t)nt * 4), hvol(nt);
    std::vector<int> icnt(n + 1, 0);
    std::vector<double> pel(nt);
    for (int e = 0; e < nt; ++e) {
      int v[4];
      for (int k = 0; k < 4; ++k) v[k] = perm[tetra[5 * e + k] - 1];
      hen[e] = make_int4(v[0], v[1], v[2], v[3]);
      const int istr = e / (3 * ntri);
      const int mt = tetra[5 * e + 4] - 1;
      const size_t ix = (size_t)mt * maxstr + istr;  // PERMX(MAXSTR,NZONE)
      const double vr = 1.0 / volu_[e];
      const double k1 = permx[ix] * vr, k2 = permy[ix] * vr, k3 = permz[ix] * vr;
      int q = 0;
      for (int a = 0; a < 4; ++a) {
        const double pb = k1 * bi[4 * e + a], pc = k2 * ci[4 * e + a], pd = k3 * di[4 * e + a];
        for (int b = a; b < 4; ++b)
          hS[(size_t)e * 10 + q++] = pb * bi[4 * e + b] + pc * ci[4 * e + b] + pd * di[4 * e + b];
        hgz[(size_t)e * 4 + a] = permz[ix] * di[4 * e + a] * (double)ivol[e];
      }
      hvol[e] = volu_[e];
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s = s + pnodi_[tetra[5 * e + k] - 1];
      pel[e] = s * 0.25;  // picuns.f:48 NODELT(PNODI)
      for (int k = 0; k < 4; ++k) icnt[v[k] + 1]++;
    }
    for (int r = 0; r < n; ++r) icnt[r + 1] += icnt[r];
    std::vector<int> hinc((size_t)4 * nt), fill(icnt.begin(), icnt.end() - 1);
    std::vector<uint32_t> hoff((size_t)4 * nt);
    for (int e = 0; e < nt; ++e) {  // ascending e per row
      const int v[4] = {hen[e].x, hen[e].y, hen[e].z, hen[e].w};
      for (int k = 0; k < 4; ++k) {
        const int r = v[k];
        const int i = fill[r]++;
        hinc[i] = e * 4 + k;
        uint32_t pk = 0;
        for (int l = 0; l < 4; ++l) {
          const int *b = &h_col[h_rowptr[r]], *en = &h_col[h_rowptr[r + 1]];
          const int *f = std::lower_bound(b, en, v[l]);
          pk |= (uint32_t)(f - b) << (8 * l);
        }
        hoff[i] = pk;
      }
    }
    // COEF4 (porosity mass, constant): lumped diagonal or full (locmas.f)
    double lm[16];
    for (int k = 0; k < 4; ++k)
      for (int l = 0; l < 4; ++l)
        lm[k * 4 + l] = lump ? (k == l ? 0.25 : 0.0) : (k == l ? 0.1 : 0.05);
    std::vector<double> hm4d(n, 0.0), hm4f;
    if (!lump) hm4f.assign(nnz_sell, 0.0);
    auto spos = [&](int r, int j) { return h_so[r >> 6] + 64 * j + (r & 63); };
    for (int r = 0; r < n; ++r)
      for (int i = icnt[r]; i < icnt[r + 1]; ++i) {
        const int e = hinc[i] >> 2, k = hinc[i] & 3;
        const double ve = volu_[e] * pel[e];
        if (lump) {
          hm4d[r] = hm4d[r] + ve * lm[k * 4 + k];
        } else {
          for (int l = 0; l < 4; ++l) {
            const int o = (hoff[i] >> (8 * l)) & 0xff;
            hm4f[spos(r, o)] = hm4f[spos(r, o)] + ve * lm[k * 4 + l];
          }
        }
      }
    use_plan = lump && maxrow <= 31;
    if (use_plan) build_plan(hS, hgz, hvol, hinc, hoff, icnt);
    enode.upload(hen, st);
    S.upload(hS, st);
    gz.upload(hgz, st);
    volu.upload(hvol, st);
    inc_ptr.upload(icnt, st);
    inc.upload(hinc, st);
    incoff.upload(hoff, st);
    m4d.upload(hm4d, st);
    if (!lump) m4full.upload(hm4f, st);
    else m4full.alloc(1);
    lmass.upload(std::vector<double>(lm, lm + 16), st);
    CG_HIP(hipStreamSynchronize(st));
  }

  // Lumped assembly plan (see AsmPlan): per row, one record per incident
  // element; SELL weights of the storage (V/16 per node of each incident
  // element) and gravity (gz/4).
  void build_plan(const std::vector<double> &hS, const std::vector<double> &hgz, const std::vector<double> &hvol,
                  const std::vector<int> &hinc, const std::vector<uint32_t> &hoff, const std::vector<int> &icnt) {
    const int ns = (n + 63) / 64;
    auto spos = [&](int r, int j) { return h_so[r >> 6] + 64 * j + (r & 63); };
    auto up = [](int k, int l) {
      int a = k < l ? k : l, b = k < l ? l : k;
      return a * 4 - (a * (a - 1)) / 2 + (b - a);
    };
    std::vector<int> hoffp(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
      int w = 0;
      for (int r = 64 * s; r < std::min(n, 64 * s + 64); ++r) w = std::max(w, icnt[r + 1] - icnt[r]);
      CG_REQUIRE((long)hoffp[s] + 64L * w < (1L << 31), "assembly plan too large");
      hoffp[s + 1] = hoffp[s] + 64 * w;
    }
    const size_t tot = hoffp[ns];
    std::vector<int> hie(tot, -1);
    std::vector<uint32_t> hio(tot, 0xFFFFFFFFu);  // padding: no slot byte is 0xFF
    std::vector<double4> hs4(tot, make_double4(0, 0, 0, 0));
    std::vector<double> hw2(nnz_sell, 0.0), hgw(nnz_sell, 0.0);
    for (int r = 0; r < n; ++r) {
      const size_t b0 = hoffp[r >> 6] + (r & 63);
      for (int i = icnt[r]; i < icnt[r + 1]; ++i) {
        const int e = hinc[i] >> 2, k = hinc[i] & 3;
        const size_t q = b0 + 64 * (size_t)(i - icnt[r]);
        hie[q] = e;
        hio[q] = hoff[i];
        const double *se = &hS[(size_t)e * 10];
        hs4[q] = make_double4(se[up(k, 0)], se[up(k, 1)], se[up(k, 2)], se[up(k, 3)]);
        for (int l = 0; l < 4; ++l) {
          const int j = (hoff[i] >> (8 * l)) & 0xff;
          hw2[spos(r, j)] += hvol[e] * 0.25 * 0.25;
          hgw[spos(r, j)] += hgz[(size_t)e * 4 + k] * 0.25;
        }
      }
    }
    plan_off.upload(hoffp, st);
    plan_ie.upload(hie, st);
    plan_io.upload(hio, st);
    plan_s4.upload(hs4, st);
    plan_w2.upload(hw2, st);
    plan_gw.upload(hgw, st);
    CG_HIP(hipStreamSynchronize(st));
  }

  void build_nodes(const double *snodi_, const double *pnodi_, const double *vgn_, const double *vgrmc_,
                   const double *vgpsat_, const double *volnod_, const double *kznod_,
                   const double *arenod_) {
    std::vector<double> a_psat(n), a_vgn(n), a_vgm(n), a_vgn1(n), a_vgnr(n), a_vgmr(n), a_vgpsn(n),
        a_pnot(n), a_rmcp(n), a_snodi(n), a_pnodi(n), a_volnod(n), a_kz(n);
    std::vector<int> hrefid(n), hrow_surf(n, -1), hsurf_row(nnod), hbot_row(nnod);
    for (int ref = 0; ref < n; ++ref) {
      const int d = perm[ref];
      const double g = vgn_[ref];
      CG_REQUIRE(g > 1.0 && pnodi_[ref] > 0.0, "vG n must exceed 1 and porosity be positive");
      const double m = (g - 1.0) / g;  // chparm.f:24-31
      a_vgn[d] = g;
      a_vgm[d] = m;
      a_vgn1[d] = g - 1.0;
      a_vgnr[d] = 1.0 / g;
      a_vgmr[d] = 1.0 / m;
      a_vgpsn[d] = std::pow(std::fabs(vgpsat_[ref]), g);
      a_psat[d] = vgpsat_[ref];
      a_pnot[d] = (pnodi_[ref] - vgrmc_[ref]) / pnodi_[ref];
      a_rmcp[d] = vgrmc_[ref] / pnodi_[ref];
      a_snodi[d] = snodi_[ref];
      a_pnodi[d] = pnodi_[ref];
      a_volnod[d] = volnod_[ref];
      a_kz[d] = kznod_ ? kznod_[ref] : 0.0;
      hrefid[d] = ref;
      if (ref < nnod) hrow_surf[d] = ref;
    }
    for (int s = 0; s < nnod; ++s) {
      hsurf_row[s] = perm[s];
      hbot_row[s] = perm[(long)nstr * nnod + s];
    }
    psat.upload(a_psat, st); vgn.upload(a_vgn, st); vgm.upload(a_vgm, st); vgn1.upload(a_vgn1, st);
    vgnr.upload(a_vgnr, st); vgmr.upload(a_vgmr, st); vgpsn.upload(a_vgpsn, st); pnot.upload(a_pnot, st);
    rmcp.upload(a_rmcp, st); snodi.upload(a_snodi, st); pnodi.upload(a_pnodi, st);
    volnod.upload(a_volnod, st); kznod.upload(a_kz, st);
    refid.upload(hrefid, st); row_surf.upload(hrow_surf, st); surf_row.upload(hsurf_row, st);
    bot_row.upload(hbot_row, st);
    arenod.upload(std::vector<double>(arenod_, arenod_ + nnod), st);
    CG_HIP(hipStreamSynchronize(st));
  }

  void alloc_state() {
    for (DBuf<double> *b : {&pnew, &pold, &ptnew, &ptold, &ptimep, &sw, &ckrw, &ckrwp, &et1, &et2, &swnew,
                            &swtimep, &xt5, &bcflux, &qnode})
      b->alloc(n);
    ckrwe.alloc(nt);
    et1e.alloc(nt);
    ifatm.alloc(nnod); ifatmp.alloc(nnod);
    pond.alloc(1);
    hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, pond.p, 0);
    atmpot.alloc(nnod); atmact.alloc(nnod); atmold.alloc(nnod); qbot.alloc(nnod);
    alloc_solver();
    CG_HIP(hipMemsetAsync(qnode.p, 0, sizeof(double) * n, st));
  }
  // what SYMSLV needs: matrix, factor, PCG vectors and reduction scratch
  void alloc_solver() {
    for (DBuf<double> *b : {&pdiff, &tnoti, &dinv, &r, &p, &Bv, &z}) b->alloc(n);
    A.alloc(nnz_sell);
    CG_HIP(hipMemsetAsync(A.p, 0, sizeof(double) * nnz_sell, st));
    isdir.alloc(n);
    CG_HIP(hipMemsetAsync(isdir.p, 0, n, st));
    if (f32) {
      F32.alloc(nnz_sell);
      CG_HIP(hipMemsetAsync(F32.p, 0, sizeof(float) * nnz_sell, st));
    } else {
      F64.alloc(nnz_sell);
      CG_HIP(hipMemsetAsync(F64.p, 0, sizeof(double) * nnz_sell, st));
    }
    part.alloc((size_t)kSlots * kMaxRedBlocks);
    ipart.alloc(kMaxRedBlocks);
    scal.alloc(1);
    dbeta.alloc(1);
    flags.alloc(2);
    res.alloc(24);
    h_res.assign(24, 0.0);
    if (!hbox.p) hbox.alloc(1);
    CG_HIP(hipMemsetAsync(flags.p, 0, sizeof(int) * 2, st));
  }

  // IC(0) factor precision of the column ordering (FP32 by default; the
  // natural ordering keeps the reference's FP64).  Reallocates the factor.
  void set_factor_fp64(bool on) {
    CG_REQUIRE(on || column_order(), "the natural ordering keeps the FP64 factor");
    if (on == !f32) return;
    CG_HIP(hipStreamSynchronize(st));
    if (on) {
      F64.alloc(nnz_sell);
      CG_HIP(hipMemsetAsync(F64.p, 0, sizeof(double) * nnz_sell, st));
      F32.release();
    } else {
      F32.alloc(nnz_sell);
      CG_HIP(hipMemsetAsync(F32.p, 0, sizeof(float) * nnz_sell, st));
      F64.release();
    }
    f32 = !on;
    CG_HIP(hipStreamSynchronize(st));
  }

  // ------------------------------------------------------- kernels glue --
  size_t col_smem(const ColDesc &C) const { return sizeof(double) * 2 * (size_t)C.len * kChains; }

  // Column sweeps with the gather window sized for the colour.  LAST
  // fuses the colour's backward sweep (it needs C.nb == 0).
  template <typename FT, bool LAST = false>
  void launch_fwd_col3(const ColDesc &C, const Sell &sl, const FT *F, const double *di, const double *tin,
                       double *out, const int *done) {
    const dim3 g(grid_for(C.nc, kChains)), b(256);
    const size_t sm = (LAST ? 2 : 1) * col_smem(C);
    switch (sweep_window(C.nf)) {
      case 0: hipLaunchKernelGGL((k_fwd_col3<FT, 0, LAST>), g, b, sm, st, C, sl, F, di, tin, out, done); break;
      case 4: hipLaunchKernelGGL((k_fwd_col3<FT, 4, LAST>), g, b, sm, st, C, sl, F, di, tin, out, done); break;
      case 8: hipLaunchKernelGGL((k_fwd_col3<FT, 8, LAST>), g, b, sm, st, C, sl, F, di, tin, out, done); break;
      case 12: hipLaunchKernelGGL((k_fwd_col3<FT, 12, LAST>), g, b, sm, st, C, sl, F, di, tin, out, done); break;
      default: hipLaunchKernelGGL((k_fwd_col3<FT, 16, LAST>), g, b, sm, st, C, sl, F, di, tin, out, done); break;
    }
  }
  template <typename FT>
  void launch_bwd_col3(const ColDesc &C, const Sell &sl, const FT *F, const double *di, const double *xin,
                       double *out, const int *done) {
    const dim3 g(grid_for(C.nc, kChains)), b(256);
    const size_t sm = col_smem(C);
    switch (sweep_window(C.nb)) {
      case 0: hipLaunchKernelGGL((k_bwd_col3<FT, 0>), g, b, sm, st, C, sl, F, di, xin, out, done); break;
      case 4: hipLaunchKernelGGL((k_bwd_col3<FT, 4>), g, b, sm, st, C, sl, F, di, xin, out, done); break;
      case 8: hipLaunchKernelGGL((k_bwd_col3<FT, 8>), g, b, sm, st, C, sl, F, di, xin, out, done); break;
      case 12: hipLaunchKernelGGL((k_bwd_col3<FT, 12>), g, b, sm, st, C, sl, F, di, xin, out, done); break;
      default: hipLaunchKernelGGL((k_bwd_col3<FT, 16>), g, b, sm, st, C, sl, F, di, xin, out, done); break;
    }
  }
  size_t fac_smem(const ColDesc &C) const {
    return (sizeof(double) * 2 + sizeof(int) * 3) * (size_t)C.len * kChains;
  }

  template <typename FT>
  void precond_apply_t(const FT *F, const double *tin, double *out, const int *done) {
    // z = (L L^T)^-1 t   (proddp.f)
    const Sell sl = sell();
    if (column_order()) {
      // the last colour couples to no later colour: its two sweeps share
      // one launch
      const size_t nc = sched.cols.size();
      const bool fuse = nc > 0 && sched.cols[nc - 1].nb == 0;
      for (size_t c = 0; c < nc; ++c) {
        if (fuse && c + 1 == nc) launch_fwd_col3<FT, true>(sched.cols[c], sl, F, dinv.p, tin, out, done);
        else launch_fwd_col3<FT>(sched.cols[c], sl, F, dinv.p, tin, out, done);
      }
      for (size_t c = fuse ? nc - 1 : nc; c-- > 0;)
        launch_bwd_col3<FT>(sched.cols[c], sl, F, dinv.p, out, out, done);
    } else {
      for (const ListLevel &L : sched.lv)
        hipLaunchKernelGGL(k_fwd_list<FT>, dim3(grid_for(L.cnt, 64)), dim3(64), 0, st, L, sched.rows.p, sl, F,
                           dinv.p, tin, out, done);
      for (size_t l = sched.lv.size(); l-- > 0;) {
        const ListLevel &L = sched.lv[l];
        hipLaunchKernelGGL(k_bwd_list<FT>, dim3(grid_for(L.cnt, 64)), dim3(64), 0, st, L, sched.rows.p, sl, F,
                           dinv.p, out, out, done);
      }
    }
  }
  void precond_apply(const double *tin, double *out, const int *done) {
    if (f32) precond_apply_t<float>(F32.p, tin, out, done);
    else precond_apply_t<double>(F64.p, tin, out, done);
  }

  template <typename FT>
  void factor_t(FT *F) {
    const Sell sl = sell();
    if (column_order()) {
      const FacTab ft{ftab.p, ftab_nch};
      for (const ColDesc &C : sched.cols) {
        const dim3 g(grid_for(C.nc, kChains)), b(256);
        const size_t sm = fac_smem(C);
        const int w = sweep_window(C.nf);
        if (ftab_ok && w <= 12) {
          switch (w) {
            case 0: hipLaunchKernelGGL((k_ic_factor_col3t<FT, 0>), g, b, sm, st, C, sl, ft, A.p, isdir.p, F, dinv.p, flags.p + 1); break;
            case 4: hipLaunchKernelGGL((k_ic_factor_col3t<FT, 4>), g, b, sm, st, C, sl, ft, A.p, isdir.p, F, dinv.p, flags.p + 1); break;
            case 8: hipLaunchKernelGGL((k_ic_factor_col3t<FT, 8>), g, b, sm, st, C, sl, ft, A.p, isdir.p, F, dinv.p, flags.p + 1); break;
            default: hipLaunchKernelGGL((k_ic_factor_col3t<FT, 12>), g, b, sm, st, C, sl, ft, A.p, isdir.p, F, dinv.p, flags.p + 1); break;
          }
        } else {
          hipLaunchKernelGGL(k_ic_factor_col3<FT>, g, b, sm, st, C, sl, pairs(), stpos.p, A.p, isdir.p, F, dinv.p,
                             flags.p + 1);
        }
      }
    } else {
      for (const ListLevel &L : sched.lv)
        hipLaunchKernelGGL(k_ic_factor_list<FT>, dim3(grid_for(L.cnt, 64)), dim3(64), 0, st, L, sched.rows.p,
                           sl, pairs(), stpos.p, A.p, isdir.p, F, dinv.p, flags.p + 1);
    }
  }
  void factor() {
    if (f32) factor_t<float>(F32.p);
    else factor_t<double>(F64.p);
  }

  // k_spmv_dot2 on the current matrix and search direction, timed by the
  // dispatch packet's events: out[0] the mean over nrep launches each after
  // a copy of flush_bytes through scratch buffers (which evicts the L2s and
  // the last-level cache, so A and p come from HBM), out[1] the mean of a
  // second launch right after (warm caches).  Touches only the SpMV's
  // output and partials, which every PCG iteration recomputes.
  void bench_spmv(int nrep, size_t flush_bytes, double *out) {
    CG_REQUIRE(initialised && nrep >= 1, "bench_spmv needs an initialised run and nrep >= 1");
    const size_t fw = std::max<size_t>(1, flush_bytes / sizeof(double));
    DBuf<double> fa(fw), fb(fw);
    CG_HIP(hipMemsetAsync(fa.p, 0, fw * sizeof(double), st));
    int *done = flags.p;
    CG_HIP(hipMemsetAsync(done, 0, sizeof(int), st));
    const int Gs = grid_for(n, kBlock);
    const FacTab fts{(spmv_transposed && ftab_ok && sell().coded && maxrow <= kUnroll) ? ftab.p : nullptr, ftab_nch};
    std::vector<hipEvent_t> ev(4 * (size_t)nrep);
    for (hipEvent_t &e : ev) CG_HIP(hipEventCreate(&e));
    for (int i = 0; i < nrep; ++i) {
      hipLaunchKernelGGL(k_copy, dim3((unsigned)((fw + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, fw, fa.p, fb.p);
      for (int h = 0; h < 2; ++h)
        hipExtLaunchKernelGGL(k_spmv_dot2, dim3(Gs), dim3(kBlock), 0, st, ev[4 * i + 2 * h], ev[4 * i + 2 * h + 1], 0,
                              n, sell(), fts, A.p, p.p, r.p, Bv.p, P(1), P(2), done);
    }
    CG_HIP(hipStreamSynchronize(st));
    double cold = 0.0, warm = 0.0;
    for (int i = 0; i < nrep; ++i) {
      float a = 0.f, b = 0.f;
      CG_HIP(hipEventElapsedTime(&a, ev[4 * i], ev[4 * i + 1]));
      CG_HIP(hipEventElapsedTime(&b, ev[4 * i + 2], ev[4 * i + 3]));
      cold += a;
      warm += b;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    out[0] = cold / nrep;
    out[1] = warm / nrep;
  }

  // SYMSLV on the current A / TNOTI: x = pdiff.  Returns NITER; err out.
  int pcg(double tol, int imax, double *err_out, int expect) {
    const int G = std::min(grid_for(n, kBlock), 1024);
    factor();
    int *done = flags.p;
    precond_apply(tnoti.p, pdiff.p, nullptr);  // x0 = M^-1 b (symslv :4686)
    hipLaunchKernelGGL(k_residual0, dim3(G), dim3(kBlock), 0, st, n, sell(), A.p, pdiff.p, tnoti.p, isdir.p,
                       r.p, P(0));
    hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(kBlock), 0, st, G, P(0), scal.p, done, tol, imax);
    precond_apply(r.p, p.p, nullptr);  // p = M^-1 r
    const int Gs = grid_for(n, kBlock);  // SpMV: one thread per row
    const int Gr = std::min(grid_for(n, kBlock), kVecGroups);
    // the SpMV reads the lower triangle through the transpose where the chain
    // tables hold every row's transposed slots (column ordering, coded rows)
    const FacTab fts{(spmv_transposed && ftab_ok && sell().coded && maxrow <= kUnroll) ? ftab.p : nullptr, ftab_nch};
    int batch = std::max(1, std::min(expect, imax));
    int launched = 0;
    int h_done = 0;
    PcgScal hs{};
    while (true) {
      for (int b = 0; b < batch; ++b) {
        if (prof) {
          // the dispatch packet's own start/end timestamps (kernel time only)
          hipEvent_t e0 = next_event(), e1 = next_event();
          hipExtLaunchKernelGGL(k_spmv_dot2, dim3(Gs), dim3(kBlock), 0, st, e0, e1, 0, n, sell(), fts, A.p, p.p,
                                r.p, Bv.p, P(1), P(2), done);
        } else {
          hipLaunchKernelGGL(k_spmv_dot2, dim3(Gs), dim3(kBlock), 0, st, n, sell(), fts, A.p, p.p, r.p, Bv.p,
                             P(1), P(2), done);
        }
        hipLaunchKernelGGL(k_pcg_alpha, dim3(Gr), dim3(kBlock), 0, st, n, Gs, P(1), P(2), Bv.p, p.p, isdir.p,
                           r.p, pdiff.p, P(3), scal.p, done);
        precond_apply(r.p, z.p, done);
        hipLaunchKernelGGL(k_dot, dim3(Gr), dim3(kBlock), 0, st, n, Bv.p, z.p, P(4), done);
        hipLaunchKernelGGL(k_pcg_beta, dim3(Gr), dim3(kBlock), 0, st, n, Gr, P(4), P(3), z.p, p.p, scal.p, done);
      }
      launched += batch;
      CG_HIP(hipMemcpyAsync(&hbox.p->done, done, sizeof(int), hipMemcpyDeviceToHost, st));
      CG_HIP(hipMemcpyAsync(&hbox.p->pcg, scal.p, sizeof(PcgScal), hipMemcpyDeviceToHost, st));
      CG_HIP(hipStreamSynchronize(st));
      h_done = hbox.p->done;
      hs = hbox.p->pcg;
      if (prof) {
        // launches after convergence exit at once; count only the live ones
        const int executed = h_done ? hs.niter : hs.niter - 1;  // bodies run so far
        const int live = std::max(0, std::min(batch, executed - (launched - batch)));
        pev_used = std::min(pev_used, 2 * live);
        collect_events();
      }
      if (h_done) break;
      CG_REQUIRE(launched <= imax + 2, "PCG did not stop");
      // later batches are short (two by default): counts fall over a step's
      // Picard iterations, so larger follow-up batches mostly add launches
      // that exit at once
      batch = std::max(1, std::min(pcg_followup, imax - hs.niter + 1));
    }
    *err_out = hs.err;
    return hs.niter;
  }

  // ---------------------------------------------- Newton (IOPT = 2) --
  DBuf<double> etai_n, dckrw, detai, ones, b_rt, b_s, b_t, b_v, b_p, b_y, b_z1, b_z2;
  DBuf<BicgScal> bscal;
  bool newton_ready = false;
  void alloc_newton() {
    if (newton_ready) return;
    CG_REQUIRE(use_plan, "the Newton path needs the lumped assembly (LUMP=1)");
    for (DBuf<double> *b : {&etai_n, &dckrw, &detai, &ones, &b_rt, &b_s, &b_t, &b_v, &b_p, &b_y, &b_z1, &b_z2})
      b->alloc(n);
    bscal.alloc(1);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, st, (size_t)n, 1.0, ones.p);
    CG_HIP(hipStreamSynchronize(st));
    newton_ready = true;
  }
  // y = L^-1 t (forward sweep, L non-unit) and x = U^-1 y (unit U: the
  // backward sweep with inverse diagonal 1).
  template <typename FT>
  void ilu_fwd_t(const FT *F, const double *tin, double *out, const int *done) {
    const Sell sl = sell();
    if (column_order()) {
      for (const ColDesc &C : sched.cols) launch_fwd_col3<FT>(C, sl, F, dinv.p, tin, out, done);
    } else {
      for (const ListLevel &L : sched.lv)
        hipLaunchKernelGGL(k_fwd_list<FT>, dim3(grid_for(L.cnt, 64)), dim3(64), 0, st, L, sched.rows.p, sl, F,
                           dinv.p, tin, out, done);
    }
  }
  template <typename FT>
  void ilu_bwd_t(const FT *F, const double *xin, double *out, const int *done) {
    const Sell sl = sell();
    if (column_order()) {
      for (size_t c = sched.cols.size(); c-- > 0;) launch_bwd_col3<FT>(sched.cols[c], sl, F, ones.p, xin, out, done);
    } else {
      for (size_t l = sched.lv.size(); l-- > 0;) {
        const ListLevel &L = sched.lv[l];
        hipLaunchKernelGGL(k_bwd_list<FT>, dim3(grid_for(L.cnt, 64)), dim3(64), 0, st, L, sched.rows.p, sl, F,
                           ones.p, xin, out, done);
      }
    }
  }
  void ilu_fwd(const double *tin, double *out, const int *done) {
    if (f32) ilu_fwd_t<float>(F32.p, tin, out, done);
    else ilu_fwd_t<double>(F64.p, tin, out, done);
  }
  void ilu_bwd(const double *xin, double *out, const int *done) {
    if (f32) ilu_bwd_t<float>(F32.p, xin, out, done);
    else ilu_bwd_t<double>(F64.p, xin, out, done);
  }
  template <typename FT>
  void ilu_factor_t(FT *F) {
    const Sell sl = sell();
    if (column_order()) {
      for (const ColDesc &C : sched.cols)
        hipLaunchKernelGGL(k_ilu_factor_col3<FT>, dim3(grid_for(C.nc, kChains)), dim3(256), ilu_smem(C), st, C,
                           sl, pairs(), stpos.p, A.p, isdir.p, F, dinv.p, flags.p + 1);
    } else {
      for (const ListLevel &L : sched.lv)
        hipLaunchKernelGGL(k_ilu_factor_list<FT>, dim3(grid_for(L.cnt, 64)), dim3(64), 0, st, L, sched.rows.p,
                           sl, pairs(), stpos.p, A.p, isdir.p, F, dinv.p, flags.p + 1);
    }
  }
  size_t ilu_smem(const ColDesc &C) const { return (sizeof(double) * 3 + sizeof(int) * 3) * (size_t)C.len * kChains; }
  // y = A x, one launch
  void spmv(const double *x, double *y, const int *done) {
    hipLaunchKernelGGL(k_spmv, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n, sell(), A.p, x, y, done);
  }
  // z = L^-1 A U^-1 h (prodbh)
  void prodbh(const double *h, double *z, const int *done) {
    ilu_bwd(h, b_z1.p, done);
    spmv(b_z1.p, b_z2.p, done);
    ilu_fwd(b_z2.p, z, done);
  }

  // NSYSLV with ISOLV=0 on the current A / TNOTI: x = pdiff.  Returns NITER.
  int bicgstab(double tol, int imax, double *err_out, int expect) {
    alloc_newton();
    if (f32) ilu_factor_t<float>(F32.p);
    else ilu_factor_t<double>(F64.p);
    int *done = flags.p;
    const int G = std::min(grid_for(n, kBlock), 1024);
    const int Gr = std::min(grid_for(n, kBlock), 256);
    // x0 = (LU)^-1 b (prod1); r = L^-1 (b - A x0); rt = r; y = U x0
    ilu_fwd(tnoti.p, b_z2.p, nullptr);
    ilu_bwd(b_z2.p, pdiff.p, nullptr);
    hipLaunchKernelGGL(k_residual0, dim3(G), dim3(kBlock), 0, st, n, sell(), A.p, pdiff.p, tnoti.p, isdir.p,
                       b_z1.p, P(0));
    ilu_fwd(b_z1.p, r.p, nullptr);
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, st, (size_t)n, r.p, b_rt.p);
    if (f32) hipLaunchKernelGGL(k_mul_upper<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, sell(), F32.p,
                                pdiff.p, b_y.p);
    else hipLaunchKernelGGL(k_mul_upper<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, sell(), F64.p,
                            pdiff.p, b_y.p);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, st, (size_t)n, 0.0, b_p.p);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, st, (size_t)n, 0.0, b_v.p);
    hipLaunchKernelGGL(k_dot, dim3(Gr), dim3(kBlock), 0, st, n, b_rt.p, r.p, P(1), nullptr);
    hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(kBlock), 0, st, G, P(0), Gr, P(1), bscal.p, done, tol, imax);
    int batch = std::max(1, std::min(expect, imax + 1));
    int launched = 0, h_done = 0;
    BicgScal hs{};
    while (true) {
      for (int b = 0; b < batch; ++b) {
        hipLaunchKernelGGL(k_bicg_p, dim3(Gr), dim3(kBlock), 0, st, n, r.p, b_v.p, b_p.p, bscal.p, done);
        prodbh(b_p.p, b_v.p, done);
        hipLaunchKernelGGL(k_dot, dim3(Gr), dim3(kBlock), 0, st, n, b_rt.p, b_v.p, P(2), done);
        hipLaunchKernelGGL(k_bicg_s, dim3(Gr), dim3(kBlock), 0, st, n, Gr, P(2), r.p, b_v.p, b_s.p, bscal.p, done);
        prodbh(b_s.p, b_t.p, done);
        hipLaunchKernelGGL(k_bicg_st, dim3(Gr), dim3(kBlock), 0, st, n, b_s.p, b_t.p, P(3), P(4), done);
        hipLaunchKernelGGL(k_bicg_update, dim3(Gr), dim3(kBlock), 0, st, n, Gr, P(3), P(4), b_p.p, b_s.p, b_t.p,
                           b_rt.p, isdir.p, b_y.p, r.p, P(18), P(1), bscal.p, done);
        hipLaunchKernelGGL(k_bicg_check, dim3(1), dim3(kBlock), 0, st, n, Gr, P(18), P(1), bscal.p, done);
      }
      launched += batch;
      CG_HIP(hipMemcpyAsync(&hbox.p->done, done, sizeof(int), hipMemcpyDeviceToHost, st));
      CG_HIP(hipMemcpyAsync(&hbox.p->bicg, bscal.p, sizeof(BicgScal), hipMemcpyDeviceToHost, st));
      CG_HIP(hipStreamSynchronize(st));
      h_done = hbox.p->done;
      hs = hbox.p->bicg;
      if (h_done) break;
      CG_REQUIRE(launched <= imax + 3, "BiCGSTAB did not stop");
      batch = std::max(2, std::min(8, imax + 1 - hs.niter + 1));
    }
    // x = U^-1 y
    ilu_bwd(b_y.p, pdiff.p, nullptr);
    *err_out = hs.err;
    return hs.niter;
  }

  // One Newton iteration (newton.f).
  void newton_iteration(double &err) {
    alloc_newton();
    const int gn = grid_for(n), ge = grid_for(nt);
    hipLaunchKernelGGL(k_constitutive_newton, dim3(gn), dim3(kBlock), 0, st, n, ptnew.p, soil(), sw.p, ckrw.p,
                       etai_n.p, dckrw.p, detai.p);
    hipLaunchKernelGGL(k_elem_avg, dim3(ge), dim3(kBlock), 0, st, nt, enode.p, ckrw.p, etai_n.p, ckrwe.p, et1e.p);
    hipLaunchKernelGGL(k_bc_rows, dim3(gn), dim3(kBlock), 0, st, n, row_surf.p, qnode.p, ifatm.p, atmact.p,
                       atmold.p, prm.tetaf, lists_on() ? bcfix.p : nullptr, isdir.p, bcflux.p);
    AsmArgs a{};
    a.n = n; a.sl = sell(); a.ckrwe = ckrwe.p; a.ptnew = ptnew.p; a.pnew = pnew.p; a.ptimep = ptimep.p;
    a.isdir = isdir.p; a.bcflux = bcflux.p; a.tetaf = prm.tetaf; a.rdt = 1.0 / deltat; a.A = A.p;
    a.tnoti = tnoti.p; a.xt5 = xt5.p;
    const AsmPlan pl{plan_off.p, plan_ie.p, plan_io.p, plan_s4.p, plan_w2.p, plan_gw.p};
    const NewtonArgs nw{etai_n.p, ckrw.p, dckrw.p, detai.p, gz.p};
    const int ga = grid_for(n, kBlock);
    if (maxrow <= 16) hipLaunchKernelGGL((k_assemble_newton<16>), dim3(ga), dim3(kBlock), 0, st, a, pl, nw);
    else hipLaunchKernelGGL((k_assemble_newton<32>), dim3(ga), dim3(kBlock), 0, st, a, pl, nw);
    const int expect = last_pcg > 0 ? last_pcg : 8;
    int niter = bicgstab(prm.tolcg, prm.itmxcg, &err, expect);
    last_pcg = niter;
    itlin += niter;
    nitert += niter;
    c_pcg += niter;
    if (niter >= prm.itmxcg) {
      lsfail = 1;
      klsfai += 1;
    } else {
      lsfail = 0;
    }
    hipLaunchKernelGGL(k_update_pnew, dim3(gn), dim3(kBlock), 0, st, n, pdiff.p, row_surf.p, ifatm.p, pold.p,
                       pnew.p);
  }

  // ------------------------------------------------ Picard iteration --
  // arguments of the assembly kernels for the current step size
  AsmArgs asm_args() const {
    AsmArgs a{};
    a.n = n; a.sl = sell(); a.inc_ptr = inc_ptr.p;
    a.inc = inc.p; a.incoff = incoff.p; a.S = S.p; a.gz = gz.p; a.volu = volu.p; a.ckrwe = ckrwe.p;
    a.et1e = et1e.p; a.m4d = m4d.p; a.m4full = m4full.p; a.lmass = lmass.p; a.ptnew = ptnew.p;
    a.pnew = pnew.p; a.ptimep = ptimep.p; a.swnew = swnew.p; a.swtimep = swtimep.p; a.et2 = et2.p;
    a.isdir = isdir.p; a.bcflux = bcflux.p; a.tetaf = prm.tetaf; a.rdt = 1.0 / deltat; a.A = A.p;
    a.tnoti = tnoti.p; a.xt5 = xt5.p;
    a.qtr = veg_on ? qtr.p : nullptr;
    return a;
  }
  AsmPlan asm_plan() const { return AsmPlan{plan_off.p, plan_ie.p, plan_io.p, plan_s4.p, plan_w2.p, plan_gw.p}; }
  void picard_iteration(double &err) {
    const int gn = grid_for(n), ge = grid_for(nt);
    hipLaunchKernelGGL(k_constitutive, dim3(gn), dim3(kBlock), 0, st, n, ptnew.p, pnew.p, soil(), sw.p,
                       ckrw.p, et1.p, et2.p, swnew.p);
    // the element means feed only the assembly; the plan kernel forms kr_e
    // from the row's own nodal kr when every row fits its 16-slot window
    if (!(use_plan && maxrow <= kUnroll))
      hipLaunchKernelGGL(k_elem_avg, dim3(ge), dim3(kBlock), 0, st, nt, enode.p, ckrw.p, et1.p, ckrwe.p,
                         et1e.p);
    hipLaunchKernelGGL(k_bc_rows, dim3(gn), dim3(kBlock), 0, st, n, row_surf.p, qnode.p, ifatm.p, atmact.p,
                       atmold.p, prm.tetaf, lists_on() ? bcfix.p : nullptr, isdir.p, bcflux.p);
    const AsmArgs a = asm_args();
    const int ga = grid_for(n, kBlock);
    if (use_plan) {
      const AsmPlan pl{plan_off.p, plan_ie.p, plan_io.p, plan_s4.p, plan_w2.p, plan_gw.p};
      if (maxrow <= 16) hipLaunchKernelGGL((k_assemble_plan<16>), dim3(ga), dim3(kBlock), 0, st, a, pl, ckrw.p, et1.p);
      else hipLaunchKernelGGL((k_assemble_plan<32>), dim3(ga), dim3(kBlock), 0, st, a, pl, ckrw.p, et1.p);
    } else if (maxrow <= 16) {
      if (lump) hipLaunchKernelGGL((k_assemble<16, true>), dim3(ga), dim3(kBlock), 0, st, a);
      else hipLaunchKernelGGL((k_assemble<16, false>), dim3(ga), dim3(kBlock), 0, st, a);
    } else {
      if (lump) hipLaunchKernelGGL((k_assemble<32, true>), dim3(ga), dim3(kBlock), 0, st, a);
      else hipLaunchKernelGGL((k_assemble<32, false>), dim3(ga), dim3(kBlock), 0, st, a);
    }
    CG_HIP(hipEventRecord(ev0, st));
    int niter = pcg(prm.tolcg, prm.itmxcg, &err, pcg_expect());
    CG_HIP(hipEventRecord(ev1, st));
    after_solve(niter);
    hipLaunchKernelGGL(k_update_pnew, dim3(gn), dim3(kBlock), 0, st, n, pdiff.p, row_surf.p, ifatm.p, pold.p,
                       pnew.p);
    shlpic_lists();
  }
  int last_pcg = 0;
  // PCG launch batches: the first covers pcg_first_pct % of the last solve's
  // iterations (8 before any solve), later ones pcg_followup iterations each
  int pcg_first_pct = 100, pcg_followup = 2;
  int pcg_expect() const {
    const int last = last_pcg > 0 ? last_pcg : 8;
    return std::max(1, (last * pcg_first_pct + 99) / 100);
  }

  // MASBAL (masbal.f:37-107) + NORMS (norms.f); result in h_res.
  void masbal_norms() {
    masbal_launch();
    const int nr = nsf_nodes > 0 ? 24 : 16;
    CG_HIP(hipMemcpyAsync(hbox.p->res, res.p, sizeof(double) * nr, hipMemcpyDeviceToHost, st));
    CG_HIP(hipStreamSynchronize(st));
    std::copy(hbox.p->res, hbox.p->res + nr, h_res.begin());
    masbal_host(h_res.data());
  }
  void masbal_launch() {
    const int gs = grid_for(nnod);
    if (prm.iopt == 2) {  // Sw at PNEW and PTIMEP for STORMB (the Picard branch has them from CHPIC0)
      hipLaunchKernelGGL(k_sw_only, dim3(grid_for(n)), dim3(kBlock), 0, st, n, pnew.p, soil(), swnew.p);
      hipLaunchKernelGGL(k_sw_only, dim3(grid_for(n)), dim3(kBlock), 0, st, n, ptimep.p, soil(), swtimep.p);
    }
    hipLaunchKernelGGL(k_bkpic, dim3(gs), dim3(kBlock), 0, st, nnod, surf_row.p, sell(), A.p, pdiff.p, xt5.p,
                       ifatm.p, atmold.p, prm.tetaf, atmact.p);
    // at most 1024 groups: their partials are summed by one group below
    const int G = std::min(red_grid(n), 1024), Gf = red_grid(nnod);
    NormPart np{P(5), P(6), P(7), P(8), P(9), P(10), ipart.p};
    hipLaunchKernelGGL(k_norms, dim3(G), dim3(kBlock), 0, st, n, refid.p, pnew.p, pold.p, tnoti.p, swnew.p,
                       swtimep.p, ptimep.p, volnod.p, snodi.p, pnodi.p, np);
    hipLaunchKernelGGL(k_flux, dim3(Gf), dim3(kBlock), 0, st, nnod, ifatm.p, atmact.p, qbot.p,
                       prm.free_drainage ? nnod : 0, P(11));
    hipLaunchKernelGGL(k_norms_final, dim3(1), dim3(kBlock), 0, st, G, Gf, np, P(11), res.p);
    if (np_cur > 0) {  // bkpic.f:38-41 at the fixed-head nodes, fluxmb.f:30-37 -> res[13], res[14]
      hipLaunchKernelGGL(k_bkpic_rows, dim3(grid_for(np_cur)), dim3(kBlock), 0, st, np_cur, dir_rows.p,
                         (const int *)nullptr, sell(), A.p, pdiff.p, xt5.p, prm.tetaf, qpold.p, qpnew.p);
      hipLaunchKernelGGL(k_list_flux, dim3(1), dim3(kBlock), 0, st, np_cur, qpnew.p, (const int *)nullptr,
                         res.p + 13);
    }
    if (nsf_nodes > 0) {  // bkpic.f:42-49 SFQ, fluxmb.f:53-62 SFFLW -> res[16..17], then CONVER's
      // EXTALL / EXTCVG (after NORMS) -> res[18] = KSF
      hipLaunchKernelGGL(k_bkpic_rows, dim3(grid_for(nsf_nodes)), dim3(kBlock), 0, st, nsf_nodes, sf_rows.p,
                         sfex.p, sell(), A.p, pdiff.p, xt5.p, prm.tetaf, sfqp.p, sfq.p);
      hipLaunchKernelGGL(k_list_flux, dim3(1), dim3(kBlock), 0, st, nsf_nodes, sfq.p, sfex.p, res.p + 16);
      hipLaunchKernelGGL(k_extall, dim3(1), dim3(kBlock), 0, st, nsf_nodes, sf_rows.p, sfex.p, sfexit.p, sfq.p,
                         pnew.p, bcfix.p, res.p + 18);
    }
  }
  // the host part of MASBAL on the reduced values R[16] (k_norms_final)
  void masbal_host(const double *R) {
    dstore = R[0];
    pl2 = R[1]; fl2 = R[2]; finf = R[3]; pinf = R[4]; pikmax = R[5]; ikmax = (int)R[6];
    const double adin = R[7], adout = R[8], anin = R[9], anout = R[10];
    const double nnin = R[11] + nx_in, nnout = R[12] + nx_out;  // free drainage, then the listed fluxes
    const double ndin = np_cur > 0 ? R[13] : 0.0, ndout = np_cur > 0 ? R[14] : 0.0;
    const double deltam = 0.5 * deltat;
    const double vadin = (adin + adinp) * deltam, vadout = (adout + adoutp) * deltam;
    const double vndin = (ndin + ndinp) * deltam, vndout = (ndout + ndoutp) * deltam;
    const double vanin = (anin + aninp) * deltam, vanout = (anout + anoutp) * deltam;
    const double vnnin = (nnin + nninp) * deltam, vnnout = (nnout + nnoutp) * deltam;
    vin = vadin + vndin + vanin + vnnin;
    vout = vadout + vndout + vanout + vnnout;
    if (nsf_nodes > 0) {  // masbal.f:69,73: VSFFLW = (SFFLW + SFFLWP) DELTAT/2
      sfflw_cur = R[16] + R[17];
      vout = vout + (sfflw_cur + sfflwp) * deltam;
      ksf = (int)R[18];
    }
    erras = vin + vout - dstore;
    errel = (vin + vout) != 0.0 ? 100.0 * erras / (vin + vout) : 0.0;
    cur[0] = adin; cur[1] = adout; cur[2] = ndin; cur[3] = ndout;
    cur[4] = anin; cur[5] = anout; cur[6] = nnin; cur[7] = nnout;
  }

  // ---------------------------------------- listed-node boundary series --
  // KIND 0: fixed heads (nansfdirbc, PRESC in m), 1: Neumann fluxes beyond
  // the base nodes of free drainage (nansfneubc, Q in m^3/s).  nodes are
  // 1-based CATHY ids; records as bcfiles.read_bc expands them.
  void set_bc(int kind, int nrec, const double *times, const int *counts, const int *nodes, const double *vals) {
    CG_REQUIRE(kind == 0 || kind == 1, "KIND is 0 (Dirichlet) or 1 (Neumann)");
    CG_REQUIRE(prm.iopt == 1, "listed boundary conditions need the Picard scheme (IOPT=1)");
    CG_REQUIRE(nrec >= 1 && times && counts, "a series needs at least one record");
    CG_REQUIRE(times[0] <= 0.0, "the first record of a series starts at time 0");
    BcSeries b;
    b.off.push_back(0);
    std::vector<char> seen(kind == 0 ? n : 0, 0);
    int e = 0;
    for (int k = 0; k < nrec; ++k) {
      CG_REQUIRE(counts[k] >= 0, "negative record length");
      CG_REQUIRE(k == 0 || times[k] >= times[k - 1], "record times must not decrease");
      CG_REQUIRE(counts[k] == 0 || (nodes && vals), "null node or value list");
      for (int i = 0; i < counts[k]; ++i, ++e) {
        const int nd = nodes[e] - 1;
        CG_REQUIRE(nd >= 0 && nd < n, "boundary node id out of 1..N");
        if (kind == 0) {
          CG_REQUIRE(!seen[nd], "a fixed-head node is listed twice in one record");
          seen[nd] = 1;
        }
        b.node.push_back(nd);
        b.row.push_back(perm[nd]);
        b.val.push_back(vals[e]);
      }
      if (kind == 0)
        for (int i = b.off.back(); i < (int)b.node.size(); ++i) seen[b.node[i]] = 0;
      b.off.push_back((int)b.node.size());
      b.t.push_back(times[k]);
    }
    (kind == 0 ? bc_dir : bc_neu) = std::move(b);
    initialised = 0;  // the run starts over at the next flow3d_init
  }
  // vegetation of the surface nodes (1-based types) and its Feddes table;
  // z(N) in CATHY numbering.  Root zones must end above the base.
  void set_vegetation(int nveg, const int *veg_type, const double *const tab[6], const double *z) {
    CG_REQUIRE(prm.iopt == 1, "root uptake needs the Picard scheme (IOPT=1)");
    CG_REQUIRE(nveg >= 1 && veg_type && z, "null argument");
    for (int i = 0; i < 6; ++i) CG_REQUIRE(tab[i], "null vegetation table");
    std::vector<int> hv(nnod);
    for (int s = 0; s < nnod; ++s) {
      const int v = veg_type[s] - 1;
      CG_REQUIRE(v >= 0 && v < nveg, "vegetation type out of 1..NVEG");
      const double zr = tab[3][v];
      CG_REQUIRE(zr > 0.0, "ZROOT must be > 0");
      CG_REQUIRE(z[s] - z[(long)nstr * nnod + s] > zr, "a root zone reaches the base of its column (etran.f:67-72)");
      for (int l = 0; l < nstr; ++l)
        CG_REQUIRE(z[(long)l * nnod + s] > z[(long)(l + 1) * nnod + s], "z must decrease down each column");
      hv[s] = v;
    }
    std::vector<double> ht(6 * (size_t)nveg);
    for (int i = 0; i < 6; ++i)
      for (int v = 0; v < nveg; ++v) ht[(size_t)i * nveg + v] = tab[i][v];
    vegc.upload(hv, st);
    vegtab.upload(ht, st);
    zc.upload(std::vector<double>(z, z + n), st);
    permd.upload(perm, st);
    qtr.alloc(n);
    CG_HIP(hipMemsetAsync(qtr.p, 0, sizeof(double) * n, st));
    CG_HIP(hipStreamSynchronize(st));
    n_veg = nveg;
    veg_on = true;
    initialised = 0;  // the run starts over at the next flow3d_init
  }
  // time-invariant seepage faces: NSF faces of COUNTS(i) 1-based nodes each,
  // top to bottom; ISFCVG = 1 makes the exit points part of the Picard
  // convergence test (inital.f:218-224)
  void set_seepage(int nsf, const int *counts, const int *nodes, int isfcvg_) {
    CG_REQUIRE(prm.iopt == 1, "seepage faces need the Picard scheme (IOPT=1)");
    CG_REQUIRE(nsf >= 0 && (nsf == 0 || (counts && nodes)), "null argument");
    CG_REQUIRE(isfcvg_ == 0 || isfcvg_ == 1, "ISFCVG is 0 or 1");
    std::vector<int> nd, off{0};
    std::vector<char> seen(n, 0);
    int e = 0;
    for (int f = 0; f < nsf; ++f) {
      CG_REQUIRE(counts[f] >= 1, "a seepage face needs at least one node");
      for (int j = 0; j < counts[f]; ++j, ++e) {
        const int v = nodes[e] - 1;
        CG_REQUIRE(v >= 0 && v < n, "seepage node id out of 1..N");
        CG_REQUIRE(!seen[v], "a node lies on two seepage faces or twice on one");
        seen[v] = 1;
        nd.push_back(v);
      }
      off.push_back((int)nd.size());
    }
    sf_node = std::move(nd);
    sf_off = std::move(off);
    nsf_nodes = (int)sf_node.size();
    isfcvg = isfcvg_;
    initialised = 0;  // the run starts over at the next flow3d_init
  }
  int n_veg = 0;
  VegTab veg_tab() const {
    const double *b = vegtab.p;
    return VegTab{b, b + n_veg, b + 2 * n_veg, b + 3 * n_veg, b + 4 * n_veg, b + 5 * n_veg};
  }
  // etran.f, once per time step before SWITCH_OLD (cathy_main.f:2929)
  void etran() {
    if (!veg_on) return;
    hipLaunchKernelGGL(k_etran, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, nstr, permd.p, zc.p, pnew.p,
                       atmpot.p, vegc.p, veg_tab(), prm.scf, qtr.p);
  }
  // device buffers of the lists, sized for the longest record
  std::vector<int> pos_tmp;
  void alloc_lists() {
    if (!lists_on()) return;
    bcfix.alloc(n);
    CG_HIP(hipMemsetAsync(bcfix.p, 0, n, st));
    CG_HIP(hipMemsetAsync(qnode.p, 0, sizeof(double) * n, st));
    const int md = std::max(1, bc_dir.max_count()), mq = std::max(1, bc_neu.max_count());
    dir_rows.alloc(md); dir_vals.alloc(md); qpnew.alloc(md); qpold.alloc(md);
    CG_HIP(hipMemsetAsync(qpnew.p, 0, sizeof(double) * md, st));
    CG_HIP(hipMemsetAsync(qpold.p, 0, sizeof(double) * md, st));
    neu_rows.alloc(mq); neu_vals.alloc(mq);
    atm_off.alloc(std::max(1, std::min(nnod, md + mq)));
    pos_tmp.assign(n, -1);
    np_cur = nx_cur = n_atm_off = 0;
    if (nsf_nodes > 0) {
      std::vector<int> rows(nsf_nodes);
      for (int i = 0; i < nsf_nodes; ++i) rows[i] = perm[sf_node[i]];
      sf_rows.upload(rows, st);
      for (DBuf<int> *b : {&sfex, &sfexit, &sfexp}) b->alloc(nsf_nodes);
      sfq.alloc(nsf_nodes);
      sfqp.alloc(nsf_nodes);
      CG_HIP(hipMemsetAsync(sfq.p, 0, sizeof(double) * nsf_nodes, st));
      CG_HIP(hipMemsetAsync(sfqp.p, 0, sizeof(double) * nsf_nodes, st));
    }
    sfflw_cur = sfflwp = 0.0;
    ksf = 0;
  }
  // sfinit.f:22-45: a seepage node at or above 1e-8 m of head starts as an
  // exit node with PTIMEP = PNEW = 0 (hp: PNEW = PTIMEP by CATHY id order of
  // device rows)
  void sfinit(std::vector<double> &hp) {
    if (nsf_nodes == 0) return;
    std::vector<int> e(nsf_nodes, 0);
    for (int i = 0; i < nsf_nodes; ++i) {
      const int r = perm[sf_node[i]];
      if (hp[r] >= 1.0e-8) {
        e[i] = 1;
        hp[r] = 0.0;
      }
    }
    sfex.upload(e, st);
    sfexit.upload(e, st);
    sfexp.upload(e, st);
    hipLaunchKernelGGL(k_sync_exit, dim3(grid_for(nsf_nodes)), dim3(kBlock), 0, st, nsf_nodes, sf_rows.p, sfex.p,
                       bcfix.p);
    CG_HIP(hipStreamSynchronize(st));
  }
  // the active fixed-head record to the device: clear the old rows' bit,
  // mark the new rows, PRESC by list position
  void apply_dir() {
    if (!bc_dir.set()) return;
    if (np_cur > 0)
      hipLaunchKernelGGL(k_mark_rows, dim3(grid_for(np_cur)), dim3(kBlock), 0, st, np_cur, dir_rows.p, 2, 0, bcfix.p);
    const int k = bc_dir.act, m = bc_dir.count(k);
    if (m > 0) {
      h_dir_rows.assign(bc_dir.row.begin() + bc_dir.off[k], bc_dir.row.begin() + bc_dir.off[k + 1]);
      h_dir_vals.assign(bc_dir.val.begin() + bc_dir.off[k], bc_dir.val.begin() + bc_dir.off[k + 1]);
      dir_rows.upload(h_dir_rows.data(), m, st);
      dir_vals.upload(h_dir_vals.data(), m, st);
      hipLaunchKernelGGL(k_mark_rows, dim3(grid_for(m)), dim3(kBlock), 0, st, m, dir_rows.p, 2, 1, bcfix.p);
    }
    np_cur = m;
  }
  std::vector<double> h_dir_vals, h_neu_vals;
  // the active Neumann record: Q by list position (neumann.f:25-31 takes the
  // window's second record for the listed nodes under free drainage), summed
  // per row in list order; the old rows' fluxes are cleared first
  void apply_neu() {
    if (!bc_neu.set()) return;
    if (nx_cur > 0)
      hipLaunchKernelGGL(k_set_rows, dim3(grid_for(nx_cur)), dim3(kBlock), 0, st, nx_cur, neu_rows.p,
                         (const double *)nullptr, 0.0, qnode.p);
    const int k = bc_neu.act, m = bc_neu.count(k);
    const int k2 = bc_neu.rec[1], m2 = bc_neu.count(k2);
    h_neu_rows.clear();
    h_neu_vals.clear();
    nx_in = nx_out = 0.0;
    for (int e = 0; e < m; ++e) {
      const double v = (prm.free_drainage && e < m2) ? bc_neu.val[bc_neu.off[k2] + e] : bc_neu.val[bc_neu.off[k] + e];
      const int r = bc_neu.row[bc_neu.off[k] + e];
      if (v > 0.0) nx_in += v; else nx_out += v;
      if (pos_tmp[r] < 0) {
        pos_tmp[r] = (int)h_neu_rows.size();
        h_neu_rows.push_back(r);
        h_neu_vals.push_back(v);
      } else {
        h_neu_vals[pos_tmp[r]] = h_neu_vals[pos_tmp[r]] + v;
      }
    }
    for (int r : h_neu_rows) pos_tmp[r] = -1;
    nx_cur = (int)h_neu_rows.size();
    neu_rows.upload(h_neu_rows.data(), nx_cur, st);
    neu_vals.upload(h_neu_vals.data(), nx_cur, st);
  }
  // surface nodes of the active lists: IFATM = -1 (atmnxt.f:80-102)
  std::vector<int> h_atm_off;
  void collect_atm_off() {
    h_atm_off.clear();
    auto take = [&](const BcSeries &b) {
      if (!b.set() || b.act < 0) return;
      for (int e = b.off[b.act]; e < b.off[b.act + 1]; ++e)
        if (b.node[e] < nnod) h_atm_off.push_back(b.node[e]);
    };
    take(bc_dir);
    take(bc_neu);
    for (int v : sf_node)
      if (v < nnod) h_atm_off.push_back(v);
    std::sort(h_atm_off.begin(), h_atm_off.end());
    h_atm_off.erase(std::unique(h_atm_off.begin(), h_atm_off.end()), h_atm_off.end());
    n_atm_off = (int)h_atm_off.size();
    if (n_atm_off > (int)atm_off.n) atm_off.alloc(n_atm_off);
    atm_off.upload(h_atm_off.data(), n_atm_off, st);
  }
  void atm_off_apply(int *fa) {
    if (n_atm_off > 0)
      hipLaunchKernelGGL(k_set_int, dim3(grid_for(n_atm_off)), dim3(kBlock), 0, st, n_atm_off, atm_off.p, -1, fa);
  }
  // NEUMANN (neumann.f) and the listed fluxes into qnode: the listed rows
  // restart from zero, free drainage writes the base rows, the listed
  // fluxes are added
  void neumann_apply(const double *ck) {
    if (nx_cur > 0)
      hipLaunchKernelGGL(k_set_rows, dim3(grid_for(nx_cur)), dim3(kBlock), 0, st, nx_cur, neu_rows.p,
                         (const double *)nullptr, 0.0, qnode.p);
    free_drainage(ck);
    if (nx_cur > 0)
      hipLaunchKernelGGL(k_add_rows, dim3(grid_for(nx_cur)), dim3(kBlock), 0, st, nx_cur, neu_rows.p, neu_vals.p,
                         qnode.p);
  }
  // SHLPIC's PNEW(CONTP) = PRESC (shlpic.f:30-35)
  void shlpic_lists() {
    if (np_cur > 0)
      hipLaunchKernelGGL(k_set_rows, dim3(grid_for(np_cur)), dim3(kBlock), 0, st, np_cur, dir_rows.p, dir_vals.p,
                         0.0, pnew.p);
    if (nsf_nodes > 0)
      hipLaunchKernelGGL(k_zero_exit, dim3(grid_for(nsf_nodes)), dim3(kBlock), 0, st, nsf_nodes, sf_rows.p,
                         sfex.p, pnew.p);
  }

  // ---- routing block ----
  // CELL(5, NCOARSE), INDCEL / INDCELWL / DEM_MAP (I, J) at (I-1) + (J-1)*NROW
  // (I the DEM row from the north), TRIANG(4, NTRI): 1-based as DATIN and
  // TRIANGOLI give them.
  void set_routing(const RouteLink &link, int nrow_, int ncol_, double dx_, double dy_, int ncoarse_,
                   const int *cell, const int *indcel, const int *indcelwl, const double *dem_map, int ntri_,
                   const int *triang) {
    CG_REQUIRE(link.numres == 0, "the routing network has reservoirs; the coupled step routes none");
    CG_REQUIRE(!link.typed, "the routing network has ALTEZZE cell types; the time loop passes none");
    CG_REQUIRE(nrow_ >= 1 && ncol_ >= 1 && link.nrow == nrow_ && link.ncol == ncol_ && link.ngrid == nrow_ * ncol_,
               "the routing network's grid is not NROW x NCOL");
    CG_REQUIRE(ntri_ == ntri, "NTRI differs from the mesh's");
    CG_REQUIRE(cell && indcel && indcelwl && dem_map && triang && ncoarse_ >= 1, "null argument");
    auto R = std::make_unique<Routing>();
    R->link = link;
    R->nrow = nrow_; R->ncol = ncol_; R->ngrid = nrow_ * ncol_;
    R->dx = dx_; R->dy = dy_;
    const int nr = nrow_, nc = ncol_, ng = R->ngrid;
    auto at = [&](const int *a, int i, int j) { return a[(size_t)(i - 1) + (size_t)(j - 1) * nr]; };
    // NOD_CELL (nod_cell.f:44-71) and CELL_NOD's coarse cells (cell_nod.f:22-48): the DEM's active
    // cells row by row, in TRIANGOLI's order
    std::vector<int> corner, target, csrc;
    for (int i = 1; i <= nr; ++i)
      for (int j = 1; j <= nc; ++j) {
        if (dem_map[(size_t)(i - 1) + (size_t)(j - 1) * nr] == 0.0) continue;
        const int c = (int)target.size();
        CG_REQUIRE(c < ncoarse_, "more active DEM cells than NCELL_COARSE");
        for (int k = 0; k < 4; ++k) {
          const int v = cell[5 * (size_t)c + k];
          CG_REQUIRE(v >= 1 && v <= nnod, "CELL holds a node outside 1..NNOD");
          corner.push_back(v - 1);
        }
        const int t = at(indcelwl, i, j);
        CG_REQUIRE(t >= 1 && t <= ng, "an active DEM cell has no INDEX_WITH_LAKES number");
        target.push_back(t - 1);
        csrc.push_back(t - 1);
      }
    CG_REQUIRE((int)target.size() == ncoarse_, "NCELL_COARSE differs from the DEM's active cells");
    R->ncoarse = ncoarse_;
    // TRANSFER_F3D_SURF / TRANSFER_SURF_F3D (transfer_f3d_surf.f:42-63): cell I_BASIN_SUB of the
    // flow grid and I_BASIN_SURF of the routing order, where both numberings hold the cell
    std::vector<int> pdst, psrc;
    for (int icol = 1; icol <= nc; ++icol)
      for (int irow = 1; irow <= nr; ++irow) {
        const int sub = nc * (nr - irow) + icol, surfi = (icol - 1) * nr + irow;
        const int i = nr - irow + 1, j = icol;
        if (at(indcel, i, j) != 0 && at(indcelwl, i, j) != 0) {
          pdst.push_back(surfi - 1);
          psrc.push_back(sub - 1);
        }
      }
    R->npair = (int)pdst.size();
    // CELL_NOD (cell_nod.f:50-68): each node's triangles' coarse cells, in triangle order
    std::vector<int> cnt(nnod + 1, 0);
    for (int k = 1; k <= ntri_; ++k)
      for (int ii = 0; ii < 3; ++ii) {
        const int v = triang[4 * (size_t)(k - 1) + ii];
        CG_REQUIRE(v >= 1 && v <= nnod, "TRIANG holds a node outside 1..NNOD");
        CG_REQUIRE(k / 2 + k % 2 <= ncoarse_, "a triangle beyond the coarse cells");
        cnt[v] += 1;
      }
    std::vector<int> ptr(nnod + 1, 0), list((size_t)3 * ntri_);
    for (int s = 0; s < nnod; ++s) {
      CG_REQUIRE(cnt[s + 1] > 0, "a surface node in no triangle");
      ptr[s + 1] = ptr[s] + cnt[s + 1];
    }
    std::vector<int> fillp(ptr.begin(), ptr.end() - 1);
    for (int k = 1; k <= ntri_; ++k)
      for (int ii = 0; ii < 3; ++ii) {
        const int v = triang[4 * (size_t)(k - 1) + ii] - 1;
        list[fillp[v]++] = k / 2 + k % 2 - 1;
      }
    R->corner.upload(corner, st); R->target.upload(target, st); R->csrc.upload(csrc, st);
    R->pair_dst.upload(pdst, st); R->pair_src.upload(psrc, st);
    R->nptr.upload(ptr, st); R->nlist.upload(list, st);
    for (DBuf<double> *b : {&R->ovflnod, &R->ovflp, &R->pondnod}) b->alloc(nnod);
    R->ovflcel.alloc(ng); R->pondcel.alloc(ng); R->coarse.alloc(ncoarse_);
    R->state.alloc((size_t)10 * ng); R->sav.alloc((size_t)4 * ng); R->pst.alloc((size_t)4 * ng);
    R->cu_ak.alloc(2); R->err.alloc(1);
    CG_HIP(hipStreamSynchronize(st));
    rt = std::move(R);
    routing_reset();
  }
  // Initial ponding heads (IPOND = 1 or 2, datin.f:394-403): with them the
  // run starts ponded (inital.f:201-207), so its first step routes.
  std::vector<double> pond0;
  void set_initial_ponding(const double *p) {
    if (!p) {
      pond0.clear();
      return;
    }
    for (int s = 0; s < nnod; ++s) CG_REQUIRE(std::isfinite(p[s]), "non-finite ponding head");
    pond0.assign(p, p + nnod);
  }
  // INIT_SR (init_sr.f:13-38) and the time loop's own arrays: all clear
  void routing_reset() {
    Routing &R = *rt;
    auto zero = [&](DBuf<double> &b) {
      hipLaunchKernelGGL(k_fill, dim3(grid_for((int)b.n)), dim3(kBlock), 0, st, b.n, 0.0, b.p);
    };
    for (DBuf<double> *b : {&R.ovflnod, &R.ovflp, &R.pondnod, &R.ovflcel, &R.pondcel, &R.coarse, &R.state, &R.sav,
                            &R.pst, &R.cu_ak})
      zero(*b);
    hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, R.err.p, 0);
    R.ak_max = R.ak_max_sav = R.ak_max_p = 0.0;
    R.ak_fresh = false;
    R.pond_host = R.pondp = 0;
    R.routed = 0;
  }
  // cathy_main.f:2993-3058 (TRANSP false): SURF_FLOWTRA (surf_flowtra.f:48-190) then PONDUPD
  void route_block() {
    Routing &R = *rt;
    const size_t g = R.ngrid;
    R.ak_max_sav = R.ak_max;
    CG_HIP(hipMemcpyAsync(R.sav.p, R.state.p + g, sizeof(double) * 4 * g, hipMemcpyDeviceToDevice, st));
    const int gs = grid_for(nnod);
    hipLaunchKernelGGL(k_nod_cell_div, dim3(gs), dim3(kBlock), 0, st, nnod, arenod.p, R.ovflnod.p);
    hipLaunchKernelGGL(k_nod_cell, dim3(grid_for(R.ncoarse)), dim3(kBlock), 0, st, R.ncoarse, R.corner.p,
                       R.target.p, R.dx, R.dy, (const double *)R.ovflnod.p, R.ovflcel.p);
    if (R.npair)
      hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for(R.npair)), dim3(kBlock), 0, st, R.npair, R.pair_dst.p,
                         R.pair_src.p, (const double *)R.ovflcel.p, R.state.p);
    // surf_flowtra.f:82-92: substeps of at most unit Courant number
    const double cu_max = R.ak_max * deltat, cutrgt = 1.0;
    int nsurf = 1;
    double dtsurf = deltat;
    if (cu_max > cutrgt) {
      dtsurf = cutrgt / R.ak_max;
      nsurf = (int)(deltat / dtsurf) + 1;
      dtsurf = deltat / nsurf;
    }
    hipLaunchKernelGGL(k_put_double, dim3(1), dim3(64), 0, st, R.cu_ak.p + 1, R.ak_max);
    for (int i = 1; i <= nsurf; ++i) {
      route_substep(R.link, st, dtsurf, R.dx, R.dy, R.state.p, R.cu_ak.p, R.err.p);
      if (nsurf > 1 && i < nsurf) route_advance_state();
    }
    R.ak_fresh = true;
    if (R.npair)
      hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for(R.npair)), dim3(kBlock), 0, st, R.npair, R.pair_src.p,
                         R.pair_dst.p, (const double *)(R.state.p + 9 * g), R.pondcel.p);
    hipLaunchKernelGGL(k_cell_coarse, dim3(grid_for(R.ncoarse)), dim3(kBlock), 0, st, R.ncoarse, R.csrc.p,
                       (const double *)R.pondcel.p, R.coarse.p);
    hipLaunchKernelGGL(k_cell_nod, dim3(gs), dim3(kBlock), 0, st, nnod, R.nptr.p, R.nlist.p,
                       (const double *)R.coarse.p, R.pondnod.p);
    hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, pond.p, 0);
    hipLaunchKernelGGL(k_pondupd, dim3(gs), dim3(kBlock), 0, st, nnod, surf_row.p, ifatm.p, prm.pondh_min,
                       1.0 / deltat, (const double *)R.pondnod.p, arenod.p, atmpot.p, atmact.p, pnew.p, pond.p);
    R.routed += 1;
  }
  // KK = KKP1, KKP1 = 0 (surf_flowtra.f:150-158, cathy_main.f:3773-3780)
  void route_advance_state() {
    Routing &R = *rt;
    const size_t g = R.ngrid;
    CG_HIP(hipMemcpyAsync(R.state.p + g, R.state.p + 5 * g, sizeof(double) * 4 * g, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_fill, dim3(grid_for((int)(4 * g))), dim3(kBlock), 0, st, 4 * g, 0.0, R.state.p + 5 * g);
  }
  // cathy_main.f:3145-3157: the attempt's saved state is the one BKSTEP restores
  void route_backstep() {
    Routing &R = *rt;
    const size_t g = R.ngrid;
    R.ak_max_p = R.ak_max_sav;
    CG_HIP(hipMemcpyAsync(R.pst.p, R.sav.p, sizeof(double) * 4 * g, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_fill, dim3(grid_for((int)(4 * g))), dim3(kBlock), 0, st, 4 * g, 0.0, R.state.p + 5 * g);
  }

  void switch_old() {
    hipLaunchKernelGGL(k_switch_old, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, surf_row.p, ifatm.p,
                       atmact.p, atmpot.p, pnew.p, prm.pmin);
  }
  // conver.f:61-71: SWITCH_OLD, or SWITCH when the surface is coupled (it
  // clears PONDING first and sets it where a node ponds)
  void switch_conv() {
    if (!surf) { switch_old(); return; }
    hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, pond.p, 0);
    hipLaunchKernelGGL(k_switch_surf, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, surf_row.p, ifatm.p,
                       atmact.p, atmpot.p, veg_on ? (const double *)qtr.p : (const double *)nullptr, arenod.p,
                       deltat, prm.pondh_min, prm.pmin, pnew.p, pond.p,
                       rt ? (const double *)rt->pondnod.p : (const double *)nullptr,
                       rt ? rt->ovflnod.p : (double *)nullptr);
  }
  // cathy_main.f:2942-2946 and bkstep.f:135-139: SWITCH_OLD, or ADRSTN
  void switch_start() {
    if (!surf) { switch_old(); return; }
    hipLaunchKernelGGL(k_adrstn, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, surf_row.p, ifatm.p,
                       atmpot.p, atmact.p, pnew.p, prm.pmin);
  }

  // FLOW3D, Picard branch (flow3d.f:93-294).
  void flow3d() {
    const int gn = grid_for(n);
    while (true) {
      double err = 0.0;
      if (prm.iopt == 2) newton_iteration(err);
      else picard_iteration(err);
      c_picard += 1;
      masbal_norms();
      if (wants_switch()) switch_conv();
      if (picard_continues()) {
        hipLaunchKernelGGL(k_next_iter, dim3(gn), dim3(kBlock), 0, st, n, prm.tetaf, pnew.p, ptimep.p,
                           pold.p, ptnew.p, ptold.p);
        continue;
      }
      return;
    }
  }
  // conver.f: atmospheric switching once the iterates settle within TOLSWI
  bool wants_switch() const {
    return (prm.l2norm == 0 && pinf <= prm.tolswi) || (prm.l2norm != 0 && pl2 <= prm.tolswi);
  }
  // flow3d.f:212-255 after MASBAL / CONVER: true to take another Picard
  // iteration (ITER advances; the caller forms the next iterate), false at
  // the end of the attempt (BCKSTP / NOBACK set).
  bool picard_continues() {
    bckstp = 0;
    const bool itagen = iter < prm.ituns;
    const bool errgmx = pl2 >= prm.ernlmx || pinf >= prm.ernlmx || fl2 >= prm.ernlmx || finf >= prm.ernlmx;
    const bool normcv = prm.l2norm == 0 ? pinf <= prm.toluns : pl2 <= prm.toluns;
    // seepage exit points: SFCHEK (inital.f:218-224), KSFZER (conver.f:90-98)
    const bool sfchek = isfcvg == 1 && nsf_nodes > 0, ksfzer = ksf == 0;
    if ((!lsfail && !errgmx && !normcv && itagen) || (!lsfail && !errgmx && itagen && sfchek && !ksfzer)) {
      if (nsf_nodes > 0)  // flow3d.f:241-246: SFEXIT = SFEX
        CG_HIP(hipMemcpyAsync(sfexit.p, sfex.p, sizeof(int) * nsf_nodes, hipMemcpyDeviceToDevice, st));
      iter += 1;
      return true;
    }
    itrtot += iter;
    if (!lsfail && !errgmx && normcv && (!sfchek || ksfzer)) {
      noback = 0;
    } else if (dtgmin) {
      bckstp = 1;
      noback = 0;
    } else {
      noback = 1;
    }
    return false;
  }
  // picard.f:165-177 after the solve: iteration counts and LSFAIL
  void after_solve(int niter) {
    last_pcg = niter;
    itlin += niter;
    nitert += niter;
    c_pcg += niter;
    if (niter >= prm.itmxcg) {
      lsfail = 1;
      klsfai += 1;
    } else {
      lsfail = 0;
    }
  }

  // ---------------------------------------------- forcing (atmbc file) --
  int arow() const { return prm.hspatm ? 1 : nnod; }
  bool atm_read(double &tin, double *vals) {
    if (atm_next >= (int)atm_t.size()) return false;
    const int k = atm_next++;
    tin = atm_t[k];
    const int w = arow();
    for (int i = 0; i < w; ++i) vals[i] = atm_v[(size_t)k * w + i];
    return true;
  }
  void atm_shift_read() {  // atmnxt.f:30-40
    const int w = arow();
    double *a1 = atminp.data(), *a2 = a1 + w, *a3 = a2 + w;
    while (!(time <= atmtim[2])) {
      atmtim[0] = atmtim[1];
      atmtim[1] = atmtim[2];
      for (int i = 0; i < w; ++i) { a1[i] = a2[i]; a2[i] = a3[i]; }
      double tin;
      if (!atm_read(tin, a3)) { htiatm = 1; break; }
      atmtim[2] = tin;
    }
  }
  // rate(s) at `time` from window slots i0,i1 (atmnxt.f:41-73 / atmbak.f)
  void atm_apply(int i0, int i1, double t0, double t1, bool interp, int set_act) {
    const int w = arow();
    std::vector<double> rate(w);
    const double *a = atminp.data() + (size_t)i0 * w, *b = atminp.data() + (size_t)i1 * w;
    for (int i = 0; i < w; ++i) {
      if (interp) {
        double slope = (b[i] - a[i]) / (t1 - t0);
        if (prm.ieto != 0) slope = 0.0;
        rate[i] = a[i] + slope * (time - t0);
      } else {
        rate[i] = b[i];
      }
    }
    const double *rn = nullptr;
    if (w > 1) {
      rate_dev.upload(rate, st);
      rn = rate_dev.p;
    }
    hipLaunchKernelGGL(k_atm_update, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, rate[0], rn, arenod.p,
                       ifatm.p, prm.scf, set_act, atmpot.p, atmact.p);
  }
  DBuf<double> rate_dev;
  void atmnxt() {
    if (htiatm != 0) return;
    atm_shift_read();
    if (atmtim[2] > atmtim[1]) atm_apply(1, 2, atmtim[1], atmtim[2], true, 1);
    else atm_apply(2, 2, 0, 0, false, 1);
  }
  void atmbak() {
    if (atmtim[0] >= atmtim[1]) return;
    atm_apply(0, 1, atmtim[0], atmtim[1], true, 1);
  }

  void free_drainage(const double *ck) {
    if (!prm.free_drainage) return;
    hipLaunchKernelGGL(k_free_drainage, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, bot_row.p, arenod.p,
                       ck, kznod.p, qbot.p, qnode.p);
  }

  double device_sum_store() {
    const int G = red_grid(n);
    hipLaunchKernelGGL(k_store, dim3(G), dim3(kBlock), 0, st, n, sw.p, volnod.p, pnodi.p, P(17));
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, st, G, P(17), res.p + 15);
    CG_HIP(hipMemcpyAsync(&hbox.p->v, res.p + 15, sizeof(double), hipMemcpyDeviceToHost, st));
    CG_HIP(hipStreamSynchronize(st));
    return hbox.p->v;
  }

  // -------------------------------------------------- run initialisation --
  // inital.f (INIT1, ATMONE, MBINIT) + cathy_main.f:2663-2704.
  void init(const double *ptimep_ref) {
    CG_REQUIRE(!atm_t.empty(), "forcing not set");
    CG_REQUIRE(atm_w == arow() && atm_v.size() == atm_t.size() * (size_t)arow(),
               "forcing records do not match HSPATM (set the parameters before the forcing)");
    std::vector<double> h(n);
    for (int i = 0; i < n; ++i) h[perm[i]] = ptimep_ref[i];
    ptimep.upload(h, st);
    // inital.f:201-205: PONDING starts clear (IPOND = 0); INIT_SR (init_sr.f)
    hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, pond.p, 0);
    if (rt) routing_reset();
    if (!pond0.empty()) {  // PONDING from the start, PONDNOD as DATIN read it
      CG_REQUIRE(surf && rt, "initial ponding needs the coupled surface with its routing (set_surface, set_routing)");
      rt->pondnod.upload(pond0.data(), nnod, st);
      rt->pond_host = rt->pondp = 1;
      hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, pond.p, 1);
    }
    // INIT1 (init1.f:37-60)
    deltat = prm.deltat;
    if (deltat > prm.dtmax) deltat = prm.dtmax;
    if (deltat <= prm.dtmin) { deltat = prm.dtmin; dtgmin = 0; } else dtgmin = 1;
    timep = 0.0;
    time = deltat;
    iter = 1; nstep = 1; nitert = 0; itlin = 0; itrtot = 0; klsfai = 0;
    finished = 0; noback = 0; bckstp = 0;
    store1 = store2 = dstore = cdstor = cvin = cvout = vtot = cerras = caeras = 0.0;
    CG_HIP(hipMemcpyAsync(pnew.p, ptimep.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    // BCONE of the fixed-head and Neumann series at TIME = DELTAT (inital.f:160-181)
    CG_REQUIRE(!(lists_on() || veg_on) || prm.iopt == 1, "listed boundary conditions and root uptake need IOPT=1");
    alloc_lists();
    if (bc_dir.set()) bc_dir.bcone(time);
    if (bc_neu.set()) bc_neu.bcone(time);
    if (lists_on()) collect_atm_off();
    // ATMONE (atmone.f:39-163), host side on the surface values
    const int w = arow();
    atminp.assign((size_t)3 * w, 0.0);
    atm_next = 0; htiatm = 0;
    atmtim[0] = atmtim[1] = atmtim[2] = 0.0;
    std::vector<double> hp(n);
    pnew.download(hp.data(), n, st);
    CG_HIP(hipStreamSynchronize(st));
    std::vector<int> fa(nnod, 0), fap(nnod, 0);
    std::vector<double> apot(nnod, 0.0), aact(nnod, 0.0), aold(nnod, 0.0);
    std::vector<double> are(nnod);
    arenod.download(are.data(), nnod, st);
    CG_HIP(hipStreamSynchronize(st));
    double *a3 = atminp.data() + 2 * (size_t)w;
    double tin;
    auto rate_at = [&](int slot, int s) { return atminp[(size_t)slot * w + (w > 1 ? s : 0)]; };
    if (!atm_read(tin, a3)) {
      htiatm = 1;
    } else {
      atmtim[2] = tin;
      if (atmtim[2] <= 0.0)
        for (int s = 0; s < nnod; ++s) aold[s] = rate_at(2, s) * are[s];
      atm_shift_read();
      for (int s = 0; s < nnod; ++s) {
        double rt;
        if (atmtim[2] > atmtim[1]) {
          double slope = (rate_at(2, s) - rate_at(1, s)) / (atmtim[2] - atmtim[1]);
          if (prm.ieto != 0) slope = 0.0;
          rt = rate_at(1, s) + slope * (time - atmtim[1]);
        } else {
          rt = rate_at(2, s);
        }
        apot[s] = rt * are[s];
      }
    }
    sfinit(hp);                    // inital.f:226-229, ahead of ATMONE
    std::vector<double> hpt = hp;  // PTIMEP == PNEW here
    for (int s : h_atm_off) fa[s] = fap[s] = -1;  // atmone.f:87-109, ahead of the classification
    for (int s = 0; s < nnod; ++s) {
      const int d = perm[s];
      if (fa[s] == -1) continue;
      if (hp[d] >= prm.pondh_min) { fa[s] = 2; fap[s] = 2; continue; }
      if (hp[d] >= 0.0 && apot[s] > 0.0) fa[s] = 1;
      if (hpt[d] >= 0.0 && aold[s] > 0.0) fap[s] = 1;
      if (hp[d] <= prm.pmin && apot[s] < 0.0) { hp[d] = prm.pmin; fa[s] = 1; }
      if (hpt[d] <= prm.pmin && aold[s] < 0.0) { hpt[d] = prm.pmin; fap[s] = 1; }
    }
    ptimep.upload(hpt, st);
    for (int s = 0; s < nnod; ++s) {
      if (fa[s] == 0) aact[s] = apot[s] >= 0.0 ? apot[s] : (1.0 - prm.scf) * apot[s];
      else aact[s] = 0.0;
      if (fap[s] == 1 || fap[s] == 2) aold[s] = 0.0;
    }
    pnew.upload(hp, st);
    ifatm.upload(fa, st); ifatmp.upload(fap, st);
    atmpot.upload(apot, st); atmact.upload(aact, st); atmold.upload(aold, st);
    const int gn = grid_for(n);
    hipLaunchKernelGGL(k_start_iter, dim3(gn), dim3(kBlock), 0, st, n, prm.tetaf, pnew.p, ptimep.p, pold.p,
                       ptnew.p, ptold.p);
    // MBINIT (mbinit.f:52-74)
    adinp = adoutp = ndinp = ndoutp = 0.0;
    aninp = anoutp = 0.0;
    for (int s = 0; s < nnod; ++s)
      if (fap[s] == 0) {
        if (aold[s] > 0.0) aninp += aold[s]; else anoutp += aold[s];
      }
    // CHVELO(PTIMEP), STORCAL, NEUMANN
    hipLaunchKernelGGL(k_sw_kr, dim3(gn), dim3(kBlock), 0, st, n, ptimep.p, soil(), sw.p, ckrw.p);
    store1 = device_sum_store();
    store2 = store1;
    CG_HIP(hipMemcpyAsync(ckrwp.p, ckrw.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    nninp = nnoutp = 0.0;
    apply_dir();
    apply_neu();
    neumann_apply(ckrw.p);
    if (prm.free_drainage) {
      std::vector<double> q(nnod);
      qbot.download(q.data(), nnod, st);
      CG_HIP(hipStreamSynchronize(st));
      for (int s = 0; s < nnod; ++s) { if (q[s] > 0.0) nninp += q[s]; else nnoutp += q[s]; }
    }
    nninp += nx_in;  // mbinit.f:44-51 / cathy_main.f:2696-2706 over the listed fluxes too
    nnoutp += nx_out;
    // SWTIMEP is a function of PTIMEP only
    hipLaunchKernelGGL(k_sw_kr, dim3(gn), dim3(kBlock), 0, st, n, ptimep.p, soil(), swtimep.p,
                       (double *)nullptr);
    dtprev = 0.0;
    if (pprev.p) CG_HIP(hipMemcpyAsync(pprev.p, ptimep.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    CG_HIP(hipStreamSynchronize(st));
    initialised = 1;
  }

  void bkstep() {  // bkstep.f:58-147 (NSF=0)
    CG_HIP(hipMemcpyAsync(pnew.p, ptimep.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if (np_cur > 0) CG_HIP(hipMemcpyAsync(qpnew.p, qpold.p, sizeof(double) * np_cur, hipMemcpyDeviceToDevice, st));
    if (nsf_nodes > 0) {  // bkstep.f:56-67: SFEX = SFEXIT = SFEXP, SFQ = SFQP
      CG_HIP(hipMemcpyAsync(sfex.p, sfexp.p, sizeof(int) * nsf_nodes, hipMemcpyDeviceToDevice, st));
      CG_HIP(hipMemcpyAsync(sfexit.p, sfexp.p, sizeof(int) * nsf_nodes, hipMemcpyDeviceToDevice, st));
      CG_HIP(hipMemcpyAsync(sfq.p, sfqp.p, sizeof(double) * nsf_nodes, hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(k_sync_exit, dim3(grid_for(nsf_nodes)), dim3(kBlock), 0, st, nsf_nodes, sf_rows.p, sfex.p,
                         bcfix.p);
    }
    CG_HIP(hipMemcpyAsync(ifatm.p, ifatmp.p, sizeof(int) * nnod, hipMemcpyDeviceToDevice, st));
    CG_HIP(hipMemcpyAsync(atmact.p, atmold.p, sizeof(double) * nnod, hipMemcpyDeviceToDevice, st));
    time = timep;
    deltat = deltat * prm.dtredm - prm.dtreds;
    if (deltat <= prm.dtmin) { deltat = prm.dtmin; dtgmin = 0; } else dtgmin = 1;
    time = time + deltat;
    iter = 1;
    nitert = 0;
    if (lists_on()) {  // bkstep.f:73-88: BCNXT past the window's second time, else BCBAK
      if (bc_dir.set()) bc_dir.back(time);
      if (bc_neu.set()) bc_neu.back(time);
      apply_dir();
      apply_neu();
      collect_atm_off();
    }
    if (time > atmtim[1]) {
      atmnxt();
      atm_off_apply(ifatm.p);
    } else {
      atmbak();
    }
    switch_start();
    if (surf && rt) {  // bkstep.f:148-157: PONDING = PONDP, OVFLNOD = OVFLP, the routing state of _P
      Routing &R = *rt;
      R.pond_host = R.pondp;
      hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, pond.p, R.pondp);
      CG_HIP(hipMemcpyAsync(R.ovflnod.p, R.ovflp.p, sizeof(double) * nnod, hipMemcpyDeviceToDevice, st));
      R.ak_max = R.ak_max_p;
      R.ak_fresh = false;
      CG_HIP(hipMemcpyAsync(R.state.p + R.ngrid, R.pst.p, sizeof(double) * 4 * R.ngrid, hipMemcpyDeviceToDevice,
                            st));
    } else if (surf) {  // PONDING = PONDP, which is clear (a ponded step ends a run without routing)
      hipLaunchKernelGGL(k_put_int, dim3(1), dim3(64), 0, st, pond.p, 0);
    }
    guess_point();
    hipLaunchKernelGGL(k_start_iter, dim3(grid_for(n)), dim3(kBlock), 0, st, n, prm.tetaf, pnew.p, ptimep.p,
                       pold.p, ptnew.p, ptold.p);
  }

  // One accepted step of cathy_main.f:2882-3830.
  int advance(double *stv) {
    step_check();
    if (finished) return finished;
    step_begin();
    while (true) {
      if (surf && rt) {  // label 100 of cathy_main.f: route a ponded surface, then POLD = PNEW on it
        if (rt->pond_host) route_block();
        hipLaunchKernelGGL(k_copy_surf, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, surf_row.p, pnew.p,
                           pold.p);
      }
      flow3d();
      if (!bckstp) break;
      if (surf && rt) route_backstep();
      back_step();
    }
    return step_end(stv);
  }
  void step_check() const {
    CG_REQUIRE(initialised, "flow3d_init not called");
    CG_REQUIRE(atm_w == arow() && atm_v.size() == atm_t.size() * (size_t)arow(),
               "forcing records do not match HSPATM");
  }
  // start of a step (cathy_main.f:2882-2978): NEUMANN, ATMNXT, SWITCH_OLD,
  // the initial iterate
  std::chrono::steady_clock::time_point step_t0;
  double step_pcg0 = 0, step_pic0 = 0;
  int step_backs = 0;
  void step_begin() {
    step_t0 = std::chrono::steady_clock::now();
    step_pcg0 = c_pcg;
    step_pic0 = c_picard;
    step_backs = 0;
    if (lists_on()) {  // BCNXT of the fixed-head and Neumann series (cathy_main.f:2892-2903)
      if (bc_dir.set()) bc_dir.bcnxt(time);
      if (bc_neu.set()) bc_neu.bcnxt(time);
      apply_dir();
      apply_neu();
      collect_atm_off();
    }
    neumann_apply(ckrw.p);
    atmnxt();
    atm_off_apply(ifatm.p);
    etran();
    switch_start();
    guess_point();
    hipLaunchKernelGGL(k_start_iter, dim3(grid_for(n)), dim3(kBlock), 0, st, n, prm.tetaf, pnew.p, ptimep.p,
                       pold.p, ptnew.p, ptold.p);
  }
  // a failed attempt: BKSTEP and the free-drainage fluxes of the restored state
  void back_step() {
    bkstep();
    neumann_apply(ckrwp.p);
    ++step_backs;
  }
  // end of an accepted step: CHVELO, STORCAL, the cumulative balance,
  // TIMUPD.  Returns 0 while the run continues, 1 at TMAX, 2 on NOBACK.
  int step_end(double *stv) {
    const auto t0 = step_t0;
    const double pcg0 = step_pcg0, pic0 = step_pic0;
    const int backs = step_backs;
    // CHVELO, STORCAL, cumulative balance (cathy_main.f:3233-3262)
    hipLaunchKernelGGL(k_sw_kr, dim3(grid_for(n)), dim3(kBlock), 0, st, n, pnew.p, soil(), sw.p, ckrw.p);
    store1 = device_sum_store();
    store2 = store2 + dstore;
    vtot += vin + vout;
    cvin += vin;
    cvout += vout;
    cdstor += dstore;
    cerras += erras;
    caeras += std::fabs(erras);
    if (stv) {
      const double v[28] = {(double)nstep, deltat, time, (double)iter, (double)nitert, (double)backs,
                            store1, store2, dstore, cdstor, vin, cvin, vout, cvout, vin + vout, vtot, erras,
                            errel, cerras, caeras, pinf, (double)itrtot, (double)itlin, (double)klsfai, pl2,
                            finf, fl2, (double)(ikmax + 1)};
      for (int i = 0; i < 28; ++i) stv[i] = v[i];
      for (int i = 28; i < 32; ++i) stv[i] = 0.0;
    }
    adinp = cur[0]; adoutp = cur[1]; ndinp = cur[2]; ndoutp = cur[3];
    aninp = cur[4]; anoutp = cur[5]; nninp = cur[6]; nnoutp = cur[7];
    if (np_cur > 0) CG_HIP(hipMemcpyAsync(qpold.p, qpnew.p, sizeof(double) * np_cur, hipMemcpyDeviceToDevice, st));
    if (nsf_nodes > 0) {  // cathy_main.f:3282-3299
      sfflwp = sfflw_cur;
      CG_HIP(hipMemcpyAsync(sfexp.p, sfex.p, sizeof(int) * nsf_nodes, hipMemcpyDeviceToDevice, st));
      CG_HIP(hipMemcpyAsync(sfexit.p, sfex.p, sizeof(int) * nsf_nodes, hipMemcpyDeviceToDevice, st));
      CG_HIP(hipMemcpyAsync(sfqp.p, sfq.p, sizeof(double) * nsf_nodes, hipMemcpyDeviceToDevice, st));
    }
    CG_HIP(hipMemcpyAsync(ifatmp.p, ifatm.p, sizeof(int) * nnod, hipMemcpyDeviceToDevice, st));
    CG_HIP(hipMemcpyAsync(atmold.p, atmact.p, sizeof(double) * nnod, hipMemcpyDeviceToDevice, st));
    if (grecord) guess_target(st, nnod, gL, gLp, gcolrow.p, ptimep.p, pnew.p, gtarget.p);
    if (pprev.p) {
      CG_HIP(hipMemcpyAsync(pprev.p, ptimep.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      dtprev = deltat;
    }
    CG_HIP(hipMemcpyAsync(ptimep.p, pnew.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    CG_HIP(hipMemcpyAsync(ckrwp.p, ckrw.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_sw_kr, dim3(grid_for(n)), dim3(kBlock), 0, st, n, ptimep.p, soil(), swtimep.p,
                       (double *)nullptr);
    int ponding = 0;
    double ak_new = 0.0;
    if (surf) pond.download(&ponding, 1, st);
    if (surf && rt) {  // cathy_main.f:3223-3225 and 3770-3782
      Routing &R = *rt;
      hipLaunchKernelGGL(k_pond_reset, dim3(grid_for(nnod)), dim3(kBlock), 0, st, nnod, surf_row.p, pnew.p,
                         R.pondnod.p);
      CG_HIP(hipMemcpyAsync(R.ovflp.p, R.ovflnod.p, sizeof(double) * nnod, hipMemcpyDeviceToDevice, st));
      route_advance_state();
      if (R.ak_fresh)
        CG_HIP(hipMemcpyAsync(&ak_new, R.cu_ak.p + 1, sizeof(double), hipMemcpyDeviceToHost, st));
    }
    CG_HIP(hipStreamSynchronize(st));
    if (surf && rt) {
      Routing &R = *rt;
      R.pondp = ponding;
      R.pond_host = ponding;
      if (R.ak_fresh) { R.ak_max = ak_new; R.ak_fresh = false; }
    }
    last_counters[0] = c_picard - pic0;
    last_counters[1] = c_pcg - pcg0;
    last_counters[4] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stv) {
      stv[28] = last_counters[4];  // wall ms from the step's start to its end
      stv[29] = last_counters[0];  // Picard iterations of the step, back-stepped attempts included
      stv[30] = last_counters[1];  // PCG iterations likewise
      stv[31] = ponding;           // PONDING at the step's end (coupled surface)
    }
    if (noback) return finished = 2;
    if (std::fabs(time - prm.tmax) <= 0.001 * deltat) return finished = 1;
    // a ponded surface without the routing linked: the next step would
    // route first (cathy_main.f:2993)
    if (ponding && !rt) return finished = 3;
    dtgmin = timupd();
    nstep += 1;
    iter = 1;
    nitert = 0;
    return 0;
  }
  double last_counters[6] = {0};

  // Prescribed step sizes (replay of an earlier run's accepted DELTAT):
  // entry k is the first attempt's dt of step k+2 (step 1's is set at
  // once); back-steps still reduce it.  Past the end TIMUPD's rule resumes.
  std::vector<double> dt_sched;
  size_t dt_pos = 0;

  void set_dt_schedule(const double *dts, int nd) {
    CG_REQUIRE(initialised, "flow3d_init not called");
    CG_REQUIRE(nstep == 1 && timep == 0.0, "a step-size schedule is set before the first step");
    CG_REQUIRE(nd >= 1, "empty step-size schedule");
    for (int i = 0; i < nd; ++i) CG_REQUIRE(dts[i] > 0.0, "step sizes must be > 0");
    dt_sched.assign(dts, dts + nd);
    deltat = dt_sched[0];
    time = timep + deltat;
    dt_pos = 1;
  }

  int timupd() {  // timupd.f:21-40
    timep = time;
    if (dt_pos < dt_sched.size()) {
      deltat = dt_sched[dt_pos++];
      if (time + deltat >= prm.tmax) { deltat = prm.tmax - time; time = prm.tmax; }
      else time = time + deltat;
      return deltat <= prm.dtmin ? 0 : 1;
    }
    if (iter < prm.ituns1) {
      deltat = deltat * prm.dtmagm + prm.dtmaga;
      if (deltat > prm.dtmax) deltat = prm.dtmax;
    }
    if (iter >= prm.ituns2) {
      deltat = deltat * prm.dtredm - prm.dtreds;
      if (deltat < prm.dtmin) deltat = prm.dtmin;
    }
    if ((time + deltat) >= prm.tmax) {
      deltat = prm.tmax - time;
      time = prm.tmax;
    } else {
      if ((time + 2 * deltat) > prm.tmax) deltat = (prm.tmax - time) / 2;
      time = time + deltat;
    }
    return deltat <= prm.dtmin ? 0 : 1;
  }

  // VEL3D + VNOD3D on the current PNEW / CKRW (the step-end CHVELO
  // values).  The per-element table (b, c, d, VOLUR, IVOL, Kx, Ky, Kz) is
  // built from the caller's arrays on the first call and kept.
  DBuf<double> vtab, e_uu, e_vv, e_ww, n_u, n_v, n_w;

  // Learned initial iterate.  Columns: surface node s, level l -> CATHY
  // node l*NNOD + s.  At the start of every attempt (after SWITCH_OLD) the
  // features of (PTIMEP, SWTIMEP, forcing, dt) are formed; with a network
  // attached PNEW takes PTIMEP + the predicted increment on all but the
  // Dirichlet surface nodes, and the Picard loop then runs unchanged.
  std::shared_ptr<GuessNet> gnet;
  bool grecord = false;
  DBuf<double> pprev;   // PTIMEP of the previous accepted step
  double dtprev = 0.0;  // its DELTAT (0: no previous step)
  int gL = 0, gLp = 0;
  DBuf<int> gcolrow;
  DBuf<float> gstat, graw, gtarget, gy;
  DBuf<unsigned short> gx;
  long guess_calls = 0;

  void set_guess(const int64_t *net, const float *stat) {
    gL = nstr + 1;
    gLp = (gL + 15) / 16 * 16;
    if (stat) {
      std::vector<int> cr((size_t)nnod * gL);
      std::vector<float> sc((size_t)nnod * gL * 5);
      for (int s = 0; s < nnod; ++s)
        for (int l = 0; l < gL; ++l) {
          const int node = l * nnod + s;
          cr[(size_t)s * gL + l] = perm[node];
          for (int c = 0; c < 5; ++c) sc[((size_t)s * gL + l) * 5 + c] = stat[(size_t)node * 5 + c];
        }
      for (int i = 0; i < nnod * gL; ++i)
        CG_REQUIRE(sc[(size_t)i * 5 + 4] > 0.f, "column features: porosity must be > 0");
      gcolrow.upload(cr, st);
      gstat.upload(sc, st);
      graw.alloc((size_t)nnod * gLp * kGuessIn);
      gtarget.alloc((size_t)nnod * gLp);
      gx.alloc((size_t)nnod * gLp * kGuessIn);
      gy.alloc((size_t)nnod * gLp);
      if (!pprev.p) {
        pprev.alloc(n);
        CG_HIP(hipMemcpyAsync(pprev.p, ptimep.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      }
      CG_HIP(hipStreamSynchronize(st));
    }
    if (net) {
      if (*net == 0) {
        gnet.reset();
      } else {
        CG_REQUIRE(gstat.p, "column features not set");
        auto g = guess_lookup(*net);
        CG_REQUIRE(g->L == gL, "guess net levels differ from NSTR+1");
        gnet = g;
      }
    }
  }

  void guess_point() {
    if (!gnet && !grecord) return;
    CG_REQUIRE(gstat.p, "column features not set");
    GuessScale use{};
    if (gnet) {
      use = gnet->scale;
    } else {
      for (int c = 0; c < kGuessIn; ++c) { use.mu[c] = 0.f; use.sd[c] = 1.f; }
    }
    const double ratio = dtprev > 0.0 ? deltat / dtprev : 0.0;
    guess_features(st, nnod, gL, gLp, gcolrow.p, ptimep.p, pprev.p, ratio, swtimep.p, gstat.p, atmpot.p, arenod.p,
                   deltat, use, grecord ? graw.p : nullptr, gnet ? gx.p : nullptr);
    if (!gnet || dtprev <= 0.0) return;  // the first step starts from PTIMEP
    CG_HIP(hipStreamSynchronize(st));
    {
      std::lock_guard<std::mutex> lk(gnet->mu);
      gnet->forward(gx.p, nullptr, nnod, gy.p);
      CG_HIP(hipStreamSynchronize(gnet->st));
    }
    guess_apply(st, nnod, gL, gLp, gcolrow.p, ifatm.p, ptimep.p, pprev.p, ratio, gy.p, use, pnew.p);
    ++guess_calls;
  }

  void guess_sample(float *feat, float *target) {
    CG_REQUIRE(grecord, "recording is off");
    if (feat) graw.download(feat, (size_t)nnod * gLp * kGuessIn, st);
    if (target) gtarget.download(target, (size_t)nnod * gLp, st);
    CG_HIP(hipStreamSynchronize(st));
  }
  void velocities(const int *tetra, const double *permx, const double *permy, const double *permz,
                  const double *bi, const double *ci, const double *di, const double *volur, const int *ivol,
                  double *uu, double *vv, double *ww, double *unod, double *vnod, double *wnod) {
    if (!vtab.p) {
      CG_REQUIRE(tetra && permx && permy && permz && bi && ci && di && volur && ivol,
                 "the first velocity call needs TETRA, PERMX/Y/Z, BI/CI/DI, VOLUR and IVOL");
      std::vector<double> h((size_t)nt * 17);
      for (int e = 0; e < nt; ++e) {
        const int istr = e / (3 * ntri), mt = tetra[5 * (size_t)e + 4] - 1;
        CG_REQUIRE(mt >= 0 && mt < nzone_, "TETRA zone out of range");
        const size_t ix = (size_t)mt * maxstr_ + istr;  // PERMX(MAXSTR, NZONE), column-major
        double *t = &h[(size_t)e * 17];
        for (int j = 0; j < 4; ++j) {
          t[j] = bi[4 * (size_t)e + j];
          t[4 + j] = ci[4 * (size_t)e + j];
          t[8 + j] = di[4 * (size_t)e + j];
        }
        t[12] = volur[e];
        t[13] = (double)ivol[e];
        t[14] = permx[ix];
        t[15] = permy[ix];
        t[16] = permz[ix];
      }
      vtab.upload(h, st);
      for (DBuf<double> *b : {&e_uu, &e_vv, &e_ww}) b->alloc(nt);
      for (DBuf<double> *b : {&n_u, &n_v, &n_w}) b->alloc(n);
    }
    hipLaunchKernelGGL(k_vel3d, dim3(grid_for(nt)), dim3(kBlock), 0, st, nt, enode.p, pnew.p, ckrw.p, vtab.p,
                       e_uu.p, e_vv.p, e_ww.p);
    hipLaunchKernelGGL(k_vnod3d, dim3(grid_for(n)), dim3(kBlock), 0, st, n, inc_ptr.p, inc.p, e_uu.p, e_vv.p,
                       e_ww.p, n_u.p, n_v.p, n_w.p);
    if (uu) e_uu.download(uu, nt, st);
    if (vv) e_vv.download(vv, nt, st);
    if (ww) e_ww.download(ww, nt, st);
    CG_HIP(hipStreamSynchronize(st));
    std::vector<double> h(n);
    for (auto pr : {std::make_pair(&n_u, unod), std::make_pair(&n_v, vnod), std::make_pair(&n_w, wnod)}) {
      if (!pr.second) continue;
      pr.first->download(h.data(), n, st);
      CG_HIP(hipStreamSynchronize(st));
      for (int i = 0; i < n; ++i) pr.second[i] = h[perm[i]];
    }
  }
  int nzone_ = 0, maxstr_ = 0;

  // After an analysis rewrote PNEW: the run continues from it as from an
  // accepted step (PTIMEP = PNEW, Sw and kr at it, the storage baseline
  // of the mass balance reset; the previous-step history of the learned
  // start is dropped).
  void adopt_state() {
    CG_REQUIRE(initialised, "flow3d_init not called");
    CG_HIP(hipMemcpyAsync(ptimep.p, pnew.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_sw_kr, dim3(grid_for(n)), dim3(kBlock), 0, st, n, pnew.p, soil(), sw.p, ckrw.p);
    CG_HIP(hipMemcpyAsync(ckrwp.p, ckrw.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_sw_kr, dim3(grid_for(n)), dim3(kBlock), 0, st, n, ptimep.p, soil(), swtimep.p,
                       (double *)nullptr);
    store1 = device_sum_store();
    store2 = store1;
    dtprev = 0.0;
    mbinit_fluxes();
    CG_HIP(hipStreamSynchronize(st));
  }
  // MBINIT (mbinit.f:28-58) at an adopted state: the previous step's
  // boundary fluxes become those a run started from this state begins
  // with -- no Dirichlet fluxes, the atmospheric fluxes ATMOLD at the
  // IFATMP = 0 nodes, and the Neumann fluxes Q (free drainage at the
  // state's kr, then the listed ones).  The cumulative totals go on.
  void mbinit_fluxes() {
    adinp = adoutp = ndinp = ndoutp = 0.0;
    aninp = anoutp = 0.0;
    std::vector<double> aold(nnod);
    std::vector<int> fap(nnod);
    atmold.download(aold.data(), nnod, st);
    ifatmp.download(fap.data(), nnod, st);
    CG_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < nnod; ++s)
      if (fap[s] == 0) {
        if (aold[s] > 0.0) aninp += aold[s]; else anoutp += aold[s];
      }
    nninp = nnoutp = 0.0;
    neumann_apply(ckrw.p);
    if (prm.free_drainage) {
      std::vector<double> q(nnod);
      qbot.download(q.data(), nnod, st);
      CG_HIP(hipStreamSynchronize(st));
      for (int s = 0; s < nnod; ++s) { if (q[s] > 0.0) nninp += q[s]; else nnoutp += q[s]; }
    }
    nninp += nx_in;
    nnoutp += nx_out;
  }

  void get_state(double *pn, double *swo, double *ck, int *fa) {
    std::vector<double> h(n);
    auto fetch = [&](const DBuf<double> &b, double *out) {
      b.download(h.data(), n, st);
      CG_HIP(hipStreamSynchronize(st));
      for (int i = 0; i < n; ++i) out[i] = h[perm[i]];
    };
    if (pn) fetch(pnew, pn);
    if (swo) fetch(sw, swo);
    if (ck) fetch(ckrw, ck);
    if (fa) {
      ifatm.download(fa, nnod, st);
      CG_HIP(hipStreamSynchronize(st));
    }
  }
};

// Flow of a handle (cathy_gpu_flow3d_create); throws on an unknown handle.
std::shared_ptr<Flow> lookup(const int64_t *h);

}  // namespace cg
