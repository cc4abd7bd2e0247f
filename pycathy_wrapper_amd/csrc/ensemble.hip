// ensemble.hip -- an ensemble of flow runs advanced in lockstep on one GPU.
//
// pyCATHY's data-assimilation forecast runs every member as its own
// ./cathy process (pyCATHY/DA/cathy_DA.py:1373-1390).  Here the members are
// flow handles (Flow, flow.h) on one device, and their Picard iterations
// advance together: each round, every member still iterating takes one
// Picard iteration, and every kernel of that iteration is ONE launch whose
// grid holds all members' work-groups (the kb_* kernels).  A member finds
// its arrays in a device table (MemberDev) and runs the single-member body
// with the group numbers of its own launch, so its results are bitwise
// those of a run alone.  Per-member convergence masks: the PCG kernels of a
// member whose solve has converged exit at once (its done flag), and the
// rounds continue until every member's solve is done; the host then reads
// every member's norms with one copy, and each member takes the decision
// of flow3d.f:212-255 on its own (next iteration, back-step, end of step,
// next step), so members of uneven size and uneven iteration counts share
// the launches without waiting for each other's steps.
//
// The work of one round between host decisions -- the Picard iteration up
// to the PCG loop, a batch of PCG iterations, the post-solve balance -- is a
// fixed launch sequence whose per-round values live in device tables, so it
// can be captured as HIP graphs once and replayed every round
// (cathy_gpu_ensemble_set_graphs) until the set of members taking part
// changes.
#include "flow.h"

namespace cg {
namespace {

constexpr int kMaxMembers = 512;

// launch shapes of the batched kernels: groups per member
enum Shape { kShRows = 0, kShRed, kShNnod, kShFlux, kShOne, kShVec, kShCol0, kNShape = kShCol0 + 4 };

int shape_groups(const Flow &f, int sh) {
  switch (sh) {
    case kShRows: return grid_for(f.n);
    case kShRed: return std::min(grid_for(f.n), 1024);
    case kShNnod: return grid_for(f.nnod);
    case kShFlux: return grid_for(f.nnod, kBlock, kMaxRedBlocks);
    case kShOne: return 1;
    case kShVec: return std::min(grid_for(f.n), kVecGroups);
    default: {
      const int c = sh - kShCol0;
      return c < (int)f.sched.cols.size() ? grid_for(f.sched.cols[c].nc, kChains) : 0;
    }
  }
}

struct Ensemble {
  std::vector<std::shared_ptr<Flow>> mem;
  int nm = 0;
  std::mutex mu;  // one advance at a time
  hipStream_t st = nullptr;
  bool f32 = true;
  int mr = 16, ncol = 0;
  int wfac[4] = {0, 0, 0, 0}, wf[4] = {0, 0, 0, 0}, wb[4] = {0, 0, 0, 0};
  size_t sm_col[4] = {0, 0, 0, 0}, sm_fac[4] = {0, 0, 0, 0};
  bool fuse_last = false;
  DBuf<MemberDev> dargs;
  Pinned<MemberDev> hargs;
  DBuf<int> dtab;
  Pinned<int> htab;
  int gtot[kNShape] = {0};
  DBuf<double> dres, dpoll;
  Pinned<double> hres, hpoll;
  DBuf<int> dact;
  Pinned<int> hact;
  std::vector<char> active, tab_active;
  // statistics of the last advance
  long rounds = 0, polls = 0, pcg_batches = 0, graph_captures = 0;
  // HIP graphs of a round's fixed launch sequences: pre-solve (Picard
  // assembly, factor, PCG start), one PCG batch of `g_follow` iterations,
  // post-solve (update, MASBAL / NORMS).  Their launches read every
  // per-round value from the device tables, so they stay valid until the
  // set of members taking part changes (the group counts of the grids).
  bool use_graphs = false;
  hipGraphExec_t g_pre = nullptr, g_iter = nullptr, g_post = nullptr;
  int g_follow = 0;

  ~Ensemble() {
    drop_graphs();
    if (st) (void)hipStreamDestroy(st);
  }
  void drop_graphs() {
    for (hipGraphExec_t *g : {&g_pre, &g_iter, &g_post})
      if (*g) {
        (void)hipGraphExecDestroy(*g);
        *g = nullptr;
      }
  }
  // capture what `body` launches on the stream as an executable graph (no
  // memset or copy nodes: only kernel launches)
  template <typename F>
  hipGraphExec_t capture(F &&body) {
    hipGraph_t graph = nullptr;
    CG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      body();
    } catch (...) {
      (void)hipStreamEndCapture(st, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    CG_HIP(hipStreamEndCapture(st, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    CG_HIP(e);
    ++graph_captures;
    return exec;
  }
  // run `body`'s launches directly, or through the graph *g (captured on
  // first use)
  template <typename F>
  void run(hipGraphExec_t *g, F &&body) {
    if (!use_graphs) {
      body();
      return;
    }
    if (!*g) *g = capture(body);
    CG_HIP(hipGraphLaunch(*g, st));
  }

  const int *tab(int sh) const { return dtab.p + (size_t)sh * (nm + 1); }

  void setup(const int64_t *handles, int m) {
    CG_REQUIRE(m >= 1 && m <= kMaxMembers, "ensemble size out of range (1..512)");
    nm = m;
    for (int j = 0; j < m; ++j) {
      mem.push_back(lookup(&handles[j]));
      for (int k = 0; k < j; ++k) CG_REQUIRE(mem[k] != mem[j], "a member appears twice");
    }
    const Flow &f0 = *mem[0];
    f32 = f0.f32;
    ncol = (int)f0.sched.cols.size();
    for (int j = 0; j < m; ++j) {
      const Flow &f = *mem[j];
      const std::string who = "member " + std::to_string(j) + ": ";
      CG_REQUIRE(f.initialised, who + "flow3d_init not called");
      CG_REQUIRE(f.column_order(), who + "the ensemble needs the column ordering");
      CG_REQUIRE(f.prm.iopt == 1, who + "the ensemble runs the Picard linearisation (IOPT = 1)");
      CG_REQUIRE(!f.lists_on() && !f.veg_on && !f.surf,
                 who + "listed boundary series, root uptake and the coupled surface switching run per member "
                       "(cathy_gpu_flow3d_advance)");
      CG_REQUIRE(f.use_plan && f.maxrow <= kUnroll, who + "the ensemble needs the lumped plan assembly (LUMP = 1)");
      CG_REQUIRE(f.ftab_ok, who + "the IC(0) chain tables do not apply to this mesh");
      CG_REQUIRE(f.f32 == f32, who + "members differ in IC(0) factor precision");
      CG_REQUIRE((int)f.sched.cols.size() == ncol && ncol >= 1 && ncol <= 4, who + "members differ in colour count");
      for (int c = 0; c < ncol; ++c) {
        const ColDesc &C = f.sched.cols[c];
        CG_REQUIRE(sweep_window(C.nf) <= 12, who + "a row has more earlier-colour slots than the table factor takes");
        wfac[c] = std::max(wfac[c], sweep_window(C.nf));
        wf[c] = std::max(wf[c], sweep_window(C.nf));
        wb[c] = std::max(wb[c], sweep_window(C.nb));
        sm_col[c] = std::max(sm_col[c], f.col_smem(C));
        sm_fac[c] = std::max(sm_fac[c], f.fac_smem(C));
      }
    }
    fuse_last = true;
    for (int j = 0; j < m; ++j) fuse_last = fuse_last && mem[j]->sched.cols[ncol - 1].nb == 0;
    CG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    dargs.alloc(m);
    hargs.alloc(m);
    dtab.alloc((size_t)kNShape * (m + 1));
    htab.alloc((size_t)kNShape * (m + 1));
    dres.alloc((size_t)16 * m);
    hres.alloc((size_t)16 * m);
    dpoll.alloc((size_t)3 * m);
    hpoll.alloc((size_t)3 * m);
    dact.alloc(m);
    hact.alloc(m);
    active.assign(m, 0);
  }

  // group tables of the members taking part in this round
  void build_tables() {
    for (int sh = 0; sh < kNShape; ++sh) {
      int *t = htab.p + (size_t)sh * (nm + 1);
      t[0] = 0;
      for (int j = 0; j < nm; ++j) t[j + 1] = t[j] + (active[j] ? shape_groups(*mem[j], sh) : 0);
      gtot[sh] = t[nm];
    }
    CG_HIP(hipMemcpyAsync(dtab.p, htab.p, sizeof(int) * kNShape * (nm + 1), hipMemcpyHostToDevice, st));
    tab_active = active;
    drop_graphs();  // the grids change with the members taking part
  }

  MemberDev member_dev(int j) const {
    const Flow &f = *mem[j];
    MemberDev d{};
    d.n = f.n;
    d.nnod = f.nnod;
    d.nq = f.prm.free_drainage ? f.nnod : 0;
    d.g_rows = grid_for(f.n);
    d.g_red = std::min(grid_for(f.n), 1024);
    d.g_vec = std::min(grid_for(f.n), kVecGroups);
    d.g_flux = grid_for(f.nnod, kBlock, kMaxRedBlocks);
    d.tetaf = f.prm.tetaf;
    d.rdt = 1.0 / f.deltat;
    d.tol = f.prm.tolcg;
    d.pmin = f.prm.pmin;
    d.imax = f.prm.itmxcg;
    d.ncol = (int)f.sched.cols.size();
    d.sl = f.sell();
    d.soil = f.soil();
    d.a = f.asm_args();
    d.P = f.asm_plan();
    d.ft = FacTab{f.ftab.p, f.ftab_nch};
    const bool tr = f.spmv_transposed && f.ftab_ok && d.sl.coded && f.maxrow <= kUnroll;
    d.fts = FacTab{tr ? f.ftab.p : nullptr, f.ftab_nch};
    for (int c = 0; c < d.ncol; ++c) d.col[c] = f.sched.cols[c];
    d.F32 = f.F32.p;
    d.F64 = f.F64.p;
    d.pnew = f.pnew.p; d.pold = f.pold.p; d.ptnew = f.ptnew.p; d.ptold = f.ptold.p; d.ptimep = f.ptimep.p;
    d.pdiff = f.pdiff.p; d.sw = f.sw.p; d.ckrw = f.ckrw.p; d.et1 = f.et1.p; d.et2 = f.et2.p;
    d.swnew = f.swnew.p; d.swtimep = f.swtimep.p; d.tnoti = f.tnoti.p; d.xt5 = f.xt5.p; d.bcflux = f.bcflux.p;
    d.qnode = f.qnode.p; d.qbot = f.qbot.p; d.atmact = f.atmact.p; d.atmold = f.atmold.p; d.atmpot = f.atmpot.p;
    d.A = f.A.p; d.dinv = f.dinv.p; d.r = f.r.p; d.p = f.p.p; d.Bv = f.Bv.p; d.z = f.z.p; d.part = f.part.p;
    d.res = dres.p + (size_t)16 * j;  // MASBAL / NORMS of every member land in one array
    d.isdir = f.isdir.p; d.ifatm = f.ifatm.p; d.flags = f.flags.p; d.ipart = f.ipart.p;
    d.row_surf = f.row_surf.p; d.surf_row = f.surf_row.p; d.refid = f.refid.p;
    d.volnod = f.volnod.p; d.snodi = f.snodi.p; d.pnodi = f.pnodi.p;
    d.scal = f.scal.p;
    return d;
  }

  // ---------------------------------------------------- batched launches --
  template <typename K, typename... A>
  void launch(K kern, int sh, size_t smem, A... args) {
    if (gtot[sh] == 0) return;
    hipLaunchKernelGGL(kern, dim3(gtot[sh]), dim3(kBlock), smem, st, tab(sh), nm, (const MemberDev *)dargs.p,
                       args...);
  }

  template <typename FT>
  void factor_t() {
    for (int c = 0; c < ncol; ++c) {
      const int sh = kShCol0 + c;
      switch (wfac[c]) {
        case 0: launch(kb_ic_factor_col3t<FT, 0>, sh, sm_fac[c], c); break;
        case 4: launch(kb_ic_factor_col3t<FT, 4>, sh, sm_fac[c], c); break;
        case 8: launch(kb_ic_factor_col3t<FT, 8>, sh, sm_fac[c], c); break;
        default: launch(kb_ic_factor_col3t<FT, 12>, sh, sm_fac[c], c); break;
      }
    }
  }
  template <typename FT, bool LAST>
  void fwd_t(int c, int vin, int vout, int use_done) {
    const int sh = kShCol0 + c;
    const size_t sm = (LAST ? 2 : 1) * sm_col[c];
    switch (wf[c]) {
      case 0: launch(kb_fwd_col3<FT, 0, LAST>, sh, sm, c, vin, vout, use_done); break;
      case 4: launch(kb_fwd_col3<FT, 4, LAST>, sh, sm, c, vin, vout, use_done); break;
      case 8: launch(kb_fwd_col3<FT, 8, LAST>, sh, sm, c, vin, vout, use_done); break;
      case 12: launch(kb_fwd_col3<FT, 12, LAST>, sh, sm, c, vin, vout, use_done); break;
      default: launch(kb_fwd_col3<FT, 16, LAST>, sh, sm, c, vin, vout, use_done); break;
    }
  }
  template <typename FT>
  void bwd_t(int c, int vio, int use_done) {
    const int sh = kShCol0 + c;
    switch (wb[c]) {
      case 0: launch(kb_bwd_col3<FT, 0>, sh, sm_col[c], c, vio, use_done); break;
      case 4: launch(kb_bwd_col3<FT, 4>, sh, sm_col[c], c, vio, use_done); break;
      case 8: launch(kb_bwd_col3<FT, 8>, sh, sm_col[c], c, vio, use_done); break;
      case 12: launch(kb_bwd_col3<FT, 12>, sh, sm_col[c], c, vio, use_done); break;
      default: launch(kb_bwd_col3<FT, 16>, sh, sm_col[c], c, vio, use_done); break;
    }
  }
  // z = (L L^T)^-1 t for every member (Flow::precond_apply_t)
  template <typename FT>
  void precond_t(int vin, int vout, int use_done) {
    for (int c = 0; c < ncol; ++c) {
      if (fuse_last && c + 1 == ncol) fwd_t<FT, true>(c, vin, vout, use_done);
      else fwd_t<FT, false>(c, vin, vout, use_done);
    }
    for (int c = fuse_last ? ncol - 1 : ncol; c-- > 0;) bwd_t<FT>(c, vout, use_done);
  }
  void factor() {
    if (f32) factor_t<float>();
    else factor_t<double>();
  }
  void precond(int vin, int vout, int use_done) {
    if (f32) precond_t<float>(vin, vout, use_done);
    else precond_t<double>(vin, vout, use_done);
  }

  // SYMSLV of every active member on its current A / TNOTI (Flow::pcg):
  // the same launches and batch policy, one launch per kernel for all.
  // one PCG iteration for every member still iterating (graddp.f:1325-1377)
  void pcg_iteration() {
    launch(kb_spmv_dot2, kShRows, 0);
    launch(kb_pcg_alpha, kShVec, 0);
    precond(kVecR, kVecZ, 1);
    launch(kb_dot, kShVec, 0);
    launch(kb_pcg_beta, kShVec, 0);
  }
  // the Picard iteration up to the PCG loop: CHPIC0, BCPIC rows, the fused
  // assembly (picard.f:74-165), IC(0), x0 = M^-1 b, r, p = M^-1 r
  void pre_solve() {
    launch(kb_constitutive, kShRows, 0);
    launch(kb_bc_rows, kShRows, 0);
    launch(kb_assemble_plan<16>, kShRows, 0);
    factor();
    precond(kVecTnoti, kVecPdiff, 0);
    launch(kb_residual0, kShRed, 0);
    launch(kb_pcg_init, kShOne, 0);
    precond(kVecR, kVecP, 0);
  }
  // PNEW update, then MASBAL + NORMS of every member into one array
  void post_solve() {
    launch(kb_update_pnew, kShRows, 0);
    launch(kb_bkpic, kShNnod, 0);
    launch(kb_norms, kShRed, 0);
    launch(kb_flux, kShFlux, 0);
    launch(kb_norms_final, kShOne, 0);
  }

  void pcg(std::vector<int> &niter) {
    run(&g_pre, [&] { pre_solve(); });
    int batch = 1, imax = 1, followup = 1;
    for (int j = 0; j < nm; ++j)
      if (active[j]) {
        const Flow &f = *mem[j];
        batch = std::max(batch, std::min(f.pcg_expect(), f.prm.itmxcg));
        imax = std::max(imax, f.prm.itmxcg);
        followup = std::max(followup, f.pcg_followup);
      }
    if (use_graphs && g_iter && g_follow != followup) {
      (void)hipGraphExecDestroy(g_iter);
      g_iter = nullptr;
    }
    g_follow = followup;
    int launched = 0;
    while (true) {
      if (use_graphs) {
        // whole follow-up batches (iterations past convergence exit at once)
        const int reps = (batch + followup - 1) / followup;
        for (int k = 0; k < reps; ++k)
          run(&g_iter, [&] {
            for (int b = 0; b < followup; ++b) pcg_iteration();
          });
        batch = reps * followup;
      } else {
        for (int b = 0; b < batch; ++b) pcg_iteration();
      }
      launched += batch;
      ++pcg_batches;
      launch(kb_poll, kShOne, 0, dpoll.p);
      CG_HIP(hipMemcpyAsync(hpoll.p, dpoll.p, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost, st));
      CG_HIP(hipStreamSynchronize(st));
      ++polls;
      bool all = true;
      int least = imax;
      for (int j = 0; j < nm; ++j)
        if (active[j] && hpoll.p[3 * j] == 0.0) {
          all = false;
          least = std::min(least, (int)hpoll.p[3 * j + 1]);
        }
      if (all) break;
      CG_REQUIRE(launched <= imax + followup + 2, "PCG did not stop");
      batch = std::max(1, std::min(followup, imax - least + 1));
    }
    for (int j = 0; j < nm; ++j) niter[j] = active[j] ? (int)hpoll.p[3 * j + 1] : 0;
  }

  // Advance every member by nsteps accepted steps (or to the end of its run).
  // rc[j]: 0 after nsteps steps, 1 at TMAX, 2 on NOBACK; taken[j] steps
  // taken; stats[(j*nsteps + s)*32 ..] those of step s (Flow::step_end).
  void advance(int nsteps, int *rc, int *taken, double *stats) {
    std::lock_guard<std::mutex> lk(mu);
    CG_REQUIRE(nsteps >= 1, "NSTEPS must be >= 1");
    for (int j = 0; j < nm; ++j) {
      mem[j]->step_check();
      CG_HIP(hipStreamSynchronize(mem[j]->st));
    }
    // every member's own work goes to the ensemble's stream meanwhile
    std::vector<hipStream_t> own(nm);
    for (int j = 0; j < nm; ++j) {
      own[j] = mem[j]->st;
      mem[j]->st = st;
    }
    struct Restore {
      Ensemble *e;
      std::vector<hipStream_t> &own;
      ~Restore() {
        (void)hipStreamSynchronize(e->st);
        for (int j = 0; j < e->nm; ++j) e->mem[j]->st = own[j];
      }
    } restore{this, own};
    rounds = polls = pcg_batches = graph_captures = 0;
    std::vector<int> niter(nm, 0), cont(nm, 0);
    for (int j = 0; j < nm; ++j) {
      rc[j] = 0;
      taken[j] = 0;
      Flow &f = *mem[j];
      if (f.finished) {
        rc[j] = f.finished;
        active[j] = 0;
        continue;
      }
      active[j] = 1;
      f.step_begin();
    }
    tab_active.assign(nm, 2);  // forces the first table build
    while (std::find(active.begin(), active.end(), 1) != active.end()) {
      ++rounds;
      if (active != tab_active) build_tables();
      for (int j = 0; j < nm; ++j)
        if (active[j]) hargs.p[j] = member_dev(j);
      CG_HIP(hipMemcpyAsync(dargs.p, hargs.p, sizeof(MemberDev) * nm, hipMemcpyHostToDevice, st));
      // PICARD (picard.f:74-200) for every member, one read-back of all
      // their MASBAL / NORMS values
      pcg(niter);
      for (int j = 0; j < nm; ++j)
        if (active[j]) mem[j]->after_solve(niter[j]);
      run(&g_post, [&] { post_solve(); });
      CG_HIP(hipMemcpyAsync(hres.p, dres.p, sizeof(double) * 16 * nm, hipMemcpyDeviceToHost, st));
      CG_HIP(hipStreamSynchronize(st));
      for (int j = 0; j < nm; ++j) {
        hact.p[j] = 0;
        if (!active[j]) continue;
        Flow &f = *mem[j];
        f.c_picard += 1;
        f.masbal_host(hres.p + 16 * j);
        const bool sw = f.wants_switch();
        cont[j] = f.picard_continues();
        hact.p[j] = (sw ? 1 : 0) | (cont[j] ? 2 : 0);
      }
      CG_HIP(hipMemcpyAsync(dact.p, hact.p, sizeof(int) * nm, hipMemcpyHostToDevice, st));
      launch(kb_switch_old, kShNnod, 0, (const int *)dact.p);
      launch(kb_next_iter, kShRows, 0, (const int *)dact.p);
      // members whose attempt ended: back-step, or end the step and start
      // the next one
      for (int j = 0; j < nm; ++j) {
        if (!active[j] || cont[j]) continue;
        Flow &f = *mem[j];
        if (f.bckstp) {
          f.back_step();
          continue;
        }
        const int r = f.step_end(stats ? stats + ((size_t)j * nsteps + taken[j]) * 32 : nullptr);
        taken[j] += 1;
        if (r != 0 || taken[j] == nsteps) {
          rc[j] = r;
          active[j] = 0;
        } else {
          f.step_begin();
        }
      }
    }
  }
};

std::mutex g_ens_mu;
std::map<int64_t, std::shared_ptr<Ensemble>> g_ens;
int64_t g_ens_next = 1;

std::shared_ptr<Ensemble> ens_lookup(const int64_t *h) {
  std::lock_guard<std::mutex> lk(g_ens_mu);
  if (!h) throw Error(CATHY_GPU_EHANDLE, "null ensemble handle");
  auto it = g_ens.find(*h);
  if (it == g_ens.end()) throw Error(CATHY_GPU_EHANDLE, "unknown ensemble handle");
  return it->second;
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

int cathy_gpu_ensemble_create(int64_t *ens, const int *nmember, const int64_t *members) {
  return guarded([&] {
    CG_REQUIRE(ens && nmember && members, "null argument");
    auto e = std::make_shared<Ensemble>();
    e->setup(members, *nmember);
    std::lock_guard<std::mutex> lk(g_ens_mu);
    *ens = g_ens_next++;
    g_ens[*ens] = std::move(e);
    return 0;
  });
}

int cathy_gpu_ensemble_advance(int64_t *ens, const int *nsteps, int *rc, int *taken, double *stats) {
  return guarded([&] {
    CG_REQUIRE(nsteps && rc && taken, "null argument");
    ens_lookup(ens)->advance(*nsteps, rc, taken, stats);
    return 0;
  });
}

int cathy_gpu_ensemble_set_graphs(int64_t *ens, const int *on) {
  return guarded([&] {
    CG_REQUIRE(on, "null argument");
    const std::shared_ptr<Ensemble> e = ens_lookup(ens);
    std::lock_guard<std::mutex> lk(e->mu);
    e->use_graphs = *on != 0;
    e->drop_graphs();
    return 0;
  });
}

int cathy_gpu_ensemble_info(int64_t *ens, int64_t *out5) {
  return guarded([&] {
    CG_REQUIRE(out5, "null argument");
    const std::shared_ptr<Ensemble> e = ens_lookup(ens);
    out5[0] = e->nm;
    out5[1] = e->rounds;
    out5[2] = e->polls;
    out5[3] = e->pcg_batches;
    out5[4] = e->graph_captures;
    return 0;
  });
}

int cathy_gpu_ensemble_destroy(int64_t *ens) {
  return guarded([&] {
    std::shared_ptr<Ensemble> e;
    {
      std::lock_guard<std::mutex> lk(g_ens_mu);
      if (!ens) throw Error(CATHY_GPU_EHANDLE, "null ensemble handle");
      auto it = g_ens.find(*ens);
      if (it == g_ens.end()) throw Error(CATHY_GPU_EHANDLE, "unknown ensemble handle");
      e = std::move(it->second);
      g_ens.erase(it);
    }
    *ens = 0;
    return 0;
  });
}

}  // extern "C"
