// cathy_gpu.hip -- C ABI and host orchestration of the CATHY subsurface
// Picard step on one MI355X (gfx950).
//
// Layering:
//   Flow (flow.h)        one mesh: device-resident state, the Picard loop of
//                        flow3d.f/picard.f, IC(0)-PCG of solscal-extended.f,
//                        and the per-time-step host logic of cathy_main.f.
//   Ensemble             members advanced in lockstep with batched launches
//   (ensemble.hip)       (cathy_gpu_ensemble_*).
//   extern "C" entries   the FLOW3D / SYMSLV drop-in boundary
//                        (include/cathy_gpu.h).
#include "flow.h"

namespace cg {

static thread_local std::string g_last_error;

// ----------------------------------------------------------- handles --
// Handles own their Flow jointly with the calls in flight: a destroy from
// another thread removes the handle, and the Flow ends with the last call.
static std::mutex g_mu;
static std::map<int64_t, std::shared_ptr<Flow>> g_flows;
static int64_t g_next = 1;

std::shared_ptr<Flow> lookup(const int64_t *h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!h) throw Error(CATHY_GPU_EHANDLE, "null handle");
  auto it = g_flows.find(*h);
  if (it == g_flows.end()) throw Error(CATHY_GPU_EHANDLE, "unknown handle");
  return it->second;
}

void set_last_error(const std::string &m) { g_last_error = m; }

// Device staging of the allocation-free EnKF analysis: W (nens x nens FP64)
// then the members' read and write state pointers, for nens <= kEnkfMax.
// One array per GPU, held until the process ends; every call on a GPU uses
// its GPU's copy, so calls that overlap on one GPU race.
constexpr size_t kScratchW = sizeof(double) * kEnkfMax * kEnkfMax;
constexpr size_t kScratchTab = sizeof(double *) * kEnkfMax;
constexpr size_t kScratchBytes = kScratchW + 2 * kScratchTab;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current GPU, looked up once per GPU
// with that GPU current (the host-side name is only a stand-in).
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch[kMaxDevices] = {};
static unsigned char *scratch_ptr() {
  int dev = -1;
  CG_HIP(hipGetDevice(&dev));
  CG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    CG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    CG_REQUIRE(ptr != nullptr, "no device address for the analysis staging array");
    g_scratch[dev] = static_cast<unsigned char *>(ptr);
  }
  return g_scratch[dev];
}

// X_a = X + X' W on the members' PNEW in place (then inflation about the
// analysis mean), with W and the state pointer tables staged in device
// memory the caller provides (dw: m*m doubles, dx / dxa: m pointers each);
// every member then continues from its analysis state.
using Members = std::array<std::shared_ptr<Flow>, kEnkfMax>;

// The members of an analysis: initialised, distinct, on one mesh and node
// ordering; their queued work has ended when this returns.
static Members analysis_members(int m, int64_t *handles) {
  Members fl;
  for (int j = 0; j < m; ++j) fl[j] = lookup(&handles[j]);
  Flow *f0 = fl[0].get();
  for (int j = 0; j < m; ++j) {
    CG_REQUIRE(fl[j]->initialised, "flow3d_init not called on every member");
    for (int k = 0; k < j; ++k) CG_REQUIRE(fl[k] != fl[j], "a member appears twice");
    CG_REQUIRE(fl[j]->n == f0->n && fl[j]->perm == f0->perm,
               "members differ in mesh or node ordering; their states do not align");
  }
  // every member's queued work ends before its state is read
  for (int j = 0; j < m; ++j) CG_HIP(hipStreamSynchronize(fl[j]->st));
  return fl;
}

// register width of the member loops: the smallest of 8, 16, 32, 64 >= m
static int member_width(int m) { return m <= 8 ? 8 : m <= 16 ? 16 : m <= 32 ? 32 : 64; }

static void enkf_apply(int m, int64_t *handles, const double *w, double inflation, double *dw,
                       const double **dx, double **dxa) {
  for (int j = 0; j < m * m; ++j) CG_REQUIRE(std::isfinite(w[j]), "non-finite analysis weight");
  Members fl = analysis_members(m, handles);
  Flow *f0 = fl[0].get();
  std::array<const double *, kEnkfMax> hx;
  std::array<double *, kEnkfMax> hxa;
  for (int j = 0; j < m; ++j) {
    hx[j] = fl[j]->pnew.p;
    hxa[j] = fl[j]->pnew.p;  // in place: each thread reads its row of every member before writing
  }
  hipStream_t s = f0->st;
  CG_HIP(hipMemcpyAsync(dx, hx.data(), sizeof(double *) * m, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(dxa, hxa.data(), sizeof(double *) * m, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(dw, w, sizeof(double) * m * m, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_enkf_update, dim3(grid_for(f0->n)), dim3(kBlock), 0, s, f0->n, m, dx, dw, inflation, dxa);
  CG_HIP(hipGetLastError());
  CG_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < m; ++j) fl[j]->adopt_state();
}

}  // namespace cg

using namespace cg;

extern "C" {

const char *cathy_gpu_last_error(void) { return g_last_error.c_str(); }

int cathy_gpu_device_info(char *name, int *name_len, int *cu_count) {
  return guarded([&] {
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) throw Error(CATHY_GPU_EDEVICE, "no HIP device");
    hipDeviceProp_t pr;
    CG_HIP(hipGetDeviceProperties(&pr, 0));
    if (name && name_len && *name_len > 0) {
      std::strncpy(name, pr.gcnArchName, *name_len - 1);
      name[*name_len - 1] = 0;
    }
    if (cu_count) *cu_count = pr.multiProcessorCount;
    return 0;
  });
}

int cathy_gpu_flow3d_create(int64_t *handle, const int *n, const int *nt, const int *ntri, const int *nnod,
                            const int *nstr, const int *nzone, const int *maxstr, const int *tetra,
                            const double *bi, const double *ci, const double *di, const double *volu,
                            const int *ivol, const double *volnod, const double *arenod, const double *permx,
                            const double *permy, const double *permz, const double *snodi,
                            const double *pnodi, const double *vgn, const double *vgrmc,
                            const double *vgpsat, const double *kznod, const int *lump,
                            const int *ordering) {
  return guarded([&] {
    CG_REQUIRE(handle && n && nt && ntri && nnod && nstr && nzone && maxstr && tetra && bi && ci && di &&
                   volu && ivol && volnod && arenod && permx && permy && permz && snodi && pnodi && vgn &&
                   vgrmc && vgpsat && lump && ordering,
               "null argument");
    auto f = std::make_unique<Flow>();
    f->setup(*n, *nt, *ntri, *nnod, *nstr, *nzone, *maxstr, tetra, bi, ci, di, volu, ivol, volnod, arenod,
             permx, permy, permz, snodi, pnodi, vgn, vgrmc, vgpsat, kznod, *lump, *ordering);
    std::lock_guard<std::mutex> lk(g_mu);
    *handle = g_next++;
    g_flows[*handle] = std::move(f);
    return 0;
  });
}

int cathy_gpu_flow3d_destroy(int64_t *handle) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!handle || !g_flows.count(*handle)) throw Error(CATHY_GPU_EHANDLE, "unknown handle");
    g_flows.erase(*handle);
    *handle = 0;
    return 0;
  });
}

int cathy_gpu_flow3d_set_params(int64_t *handle, const int *ip, const double *dp) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(ip && dp, "null argument");
    CG_REQUIRE(ip[0] == 1 || ip[0] == 2, "IOPT must be 1 (Picard) or 2 (Newton)");
    CG_REQUIRE(ip[0] == 1 || f->use_plan, "IOPT=2 on the device needs LUMP=1 and rows of at most 31 entries");
    Params &p = f->prm;
    p.iopt = ip[0]; p.l2norm = ip[1]; p.ituns = ip[2]; p.ituns1 = ip[3]; p.ituns2 = ip[4];
    p.itmxcg = ip[5]; p.ieto = ip[6]; p.free_drainage = ip[7]; p.hspatm = ip[8];
    p.tetaf = dp[0]; p.toluns = dp[1]; p.tolswi = dp[2]; p.ernlmx = dp[3]; p.tolcg = dp[4];
    p.deltat = dp[5]; p.dtmin = dp[6]; p.dtmax = dp[7]; p.tmax = dp[8]; p.dtmaga = dp[9];
    p.dtmagm = dp[10]; p.dtreds = dp[11]; p.dtredm = dp[12]; p.pmin = dp[13]; p.pondh_min = dp[14];
    p.scf = dp[15];
    CG_REQUIRE(p.tetaf > 0.0 && p.deltat > 0.0 && p.itmxcg >= 1 && p.ituns >= 1, "bad solver parameters");
    return 0;
  });
}

int cathy_gpu_flow3d_set_forcing(int64_t *handle, const int *natm, const double *atmtim,
                                 const double *atmval) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(natm && *natm >= 1 && atmtim && atmval, "forcing needs at least one record");
    const int w = f->arow();  // one value per record (HSPATM=1) or one per surface node
    f->atm_t.assign(atmtim, atmtim + *natm);
    f->atm_v.assign(atmval, atmval + (size_t)(*natm) * w);
    f->atm_w = w;
    return 0;
  });
}

int cathy_gpu_flow3d_init(int64_t *handle, const double *ptimep) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(ptimep, "null argument");
    f->init(ptimep);
    return 0;
  });
}

int cathy_gpu_flow3d_advance(int64_t *handle, double *stats) {
  return guarded([&] { return lookup(handle)->advance(stats); });
}

int cathy_gpu_flow3d_get_state(int64_t *handle, double *pnew, double *sw, double *ckrw, int *ifatm) {
  return guarded([&] {
    lookup(handle)->get_state(pnew, sw, ckrw, ifatm);
    return 0;
  });
}

int cathy_gpu_flow3d_vel3d(int64_t *handle, const int *tetra, const double *permx, const double *permy,
                           const double *permz, const double *bi, const double *ci, const double *di,
                           const double *volur, const int *ivol, double *uu, double *vv, double *ww,
                           double *unod, double *vnod, double *wnod) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(f->initialised, "flow3d_init not called");
    f->velocities(tetra, permx, permy, permz, bi, ci, di, volur, ivol, uu, vv, ww, unod, vnod, wnod);
    return 0;
  });
}

int cathy_gpu_flow3d_set_spmv_transposed(int64_t *handle, const int *on, int *active) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(on, "null argument");
    f->spmv_transposed = *on != 0;
    if (active) *active = (f->spmv_transposed && f->ftab_ok && f->sell().coded && f->maxrow <= kUnroll) ? 1 : 0;
    return 0;
  });
}

int cathy_gpu_flow3d_set_initial_ponding(int64_t *handle, const double *pondnod) {
  return guarded([&] {
    lookup(handle)->set_initial_ponding(pondnod);
    return 0;
  });
}

int cathy_gpu_flow3d_set_factor_fp64(int64_t *handle, const int *on) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(on, "null argument");
    f->set_factor_fp64(*on != 0);
    return 0;
  });
}

int cathy_gpu_flow3d_set_dt_schedule(int64_t *handle, const int *nd, const double *dts) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(nd && dts, "null argument");
    f->set_dt_schedule(dts, *nd);
    return 0;
  });
}

int cathy_gpu_flow3d_guess(int64_t *handle, const int64_t *net, const float *stat) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    f->set_guess(net, stat);
    return 0;
  });
}

int cathy_gpu_flow3d_guess_record(int64_t *handle, const int *on) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(on, "null argument");
    CG_REQUIRE(!*on || f->gstat.p, "column features not set");
    f->grecord = *on != 0;
    return 0;
  });
}

int cathy_gpu_flow3d_guess_sample(int64_t *handle, float *feat, float *target) {
  return guarded([&] {
    lookup(handle)->guess_sample(feat, target);
    return 0;
  });
}

int cathy_gpu_enkf_update(const int *nens, int64_t *handles, const double *w, const double *inflation) {
  return guarded([&] {
    CG_REQUIRE(nens && handles && w && inflation, "null argument");
    const int m = *nens;
    CG_REQUIRE(m >= 2 && m <= kEnkfMax, "ensemble size out of range (2..64)");
    DBuf<const double *> dx(m);
    DBuf<double *> dxa(m);
    DBuf<double> dw((size_t)m * m);
    enkf_apply(m, handles, w, *inflation, dw.p, dx.p, dxa.p);
    return 0;
  });
}

int cathy_gpu_enkf_update_noalloc(const int *nens, int64_t *handles, const double *w, const double *inflation) {
  return guarded([&] {
    CG_REQUIRE(nens && handles && w && inflation, "null argument");
    const int m = *nens;
    CG_REQUIRE(m >= 2 && m <= kEnkfMax,
               "ensemble size out of range (2..64): the allocation-free analysis stages at most 64 members");
    unsigned char *base = scratch_ptr();
    double *dw = reinterpret_cast<double *>(base);
    const double **dx = reinterpret_cast<const double **>(base + kScratchW);
    double **dxa = reinterpret_cast<double **>(base + kScratchW + kScratchTab);
    enkf_apply(m, handles, w, *inflation, dw, dx, dxa);
    return 0;
  });
}

}  // extern "C"

namespace cg {
// the localised analyses: one of loc (dense), row_zone + obs_zone (zones),
// node_xy + obs_xy + radius (Gaspari-Cohn)
static void enkf_localized_apply(int m, int64_t *handles, int mo, const double *s, const double *invd,
                                 const double *loc, const double *node_xy, const double *obs_xy, double radius,
                                 const int *row_zone, const int *obs_zone, double inflation) {
  CG_REQUIRE(m >= 2 && m <= kEnkfMax, "ensemble size out of range (2..64)");
  CG_REQUIRE(mo >= 1, "at least one observation");
  CG_REQUIRE(std::isfinite(inflation), "non-finite inflation");
  for (size_t j = 0; j < (size_t)mo * m; ++j)
    CG_REQUIRE(std::isfinite(s[j]) && std::isfinite(invd[j]), "non-finite observation term");
  Members fl = analysis_members(m, handles);
  Flow *f0 = fl[0].get();
  const int n = f0->n;
  hipStream_t st = f0->st;
  std::array<const double *, kEnkfMax> hx;
  std::array<double *, kEnkfMax> hxa;
  for (int j = 0; j < m; ++j) hx[j] = hxa[j] = fl[j]->pnew.p;
  DBuf<const double *> dx(m);
  DBuf<double *> dxa(m);
  DBuf<double> ds((size_t)mo * m), dinv((size_t)mo * m), dloc, dnxy, doxy;
  DBuf<int> dip(n), drz, doz;
  dx.upload(hx.data(), m, st);
  dxa.upload(hxa.data(), m, st);
  ds.upload(s, (size_t)mo * m, st);
  dinv.upload(invd, (size_t)mo * m, st);
  dip.upload(f0->iperm.data(), n, st);
  if (loc) {
    dloc.alloc((size_t)n * mo);
    dloc.upload(loc, (size_t)n * mo, st);
  } else if (row_zone) {
    drz.alloc(n);
    doz.alloc(mo);
    drz.upload(row_zone, n, st);
    doz.upload(obs_zone, mo, st);
  } else {
    dnxy.alloc((size_t)2 * n);
    doxy.alloc((size_t)2 * mo);
    dnxy.upload(node_xy, (size_t)2 * n, st);
    doxy.upload(obs_xy, (size_t)2 * mo, st);
  }
  const LocArgs a{n, m, mo, dx.p, dxa.p, ds.p, dinv.p, dip.p, dloc.p, dnxy.p, doxy.p, drz.p, doz.p,
                  radius, inflation};
  const dim3 g(grid_for(n)), b(kBlock);
  switch (member_width(m)) {
    case 8: hipLaunchKernelGGL(k_enkf_localized<8>, g, b, 0, st, a); break;
    case 16: hipLaunchKernelGGL(k_enkf_localized<16>, g, b, 0, st, a); break;
    case 32: hipLaunchKernelGGL(k_enkf_localized<32>, g, b, 0, st, a); break;
    default: hipLaunchKernelGGL(k_enkf_localized<64>, g, b, 0, st, a); break;
  }
  CG_HIP(hipGetLastError());
  CG_HIP(hipStreamSynchronize(st));
  for (int j = 0; j < m; ++j) fl[j]->adopt_state();
}
}  // namespace cg

extern "C" {

int cathy_gpu_enkf_update_localized(const int *nens, int64_t *handles, const int *nobs, const double *s,
                                    const double *invd, const double *loc, const double *node_xy,
                                    const double *obs_xy, const double *radius, const double *inflation) {
  return guarded([&] {
    CG_REQUIRE(nens && handles && nobs && s && invd && inflation, "null argument");
    CG_REQUIRE(loc || (node_xy && obs_xy && radius),
               "give the localisation matrix, or node and observation positions with a radius");
    if (!loc) CG_REQUIRE(std::isfinite(*radius) && *radius > 0.0, "the localisation radius must be positive");
    enkf_localized_apply(*nens, handles, *nobs, s, invd, loc, node_xy, obs_xy, loc ? 0.0 : *radius, nullptr,
                         nullptr, *inflation);
    return 0;
  });
}

int cathy_gpu_enkf_update_zones(const int *nens, int64_t *handles, const int *nobs, const double *s,
                                const double *invd, const int *row_zone, const int *obs_zone,
                                const double *inflation) {
  return guarded([&] {
    CG_REQUIRE(nens && handles && nobs && s && invd && row_zone && obs_zone && inflation, "null argument");
    enkf_localized_apply(*nens, handles, *nobs, s, invd, nullptr, nullptr, nullptr, 0.0, row_zone, obs_zone,
                         *inflation);
    return 0;
  });
}

int cathy_gpu_enkf_gram(const int *nens, int64_t *handles, double *gram) {
  return guarded([&] {
    CG_REQUIRE(nens && handles && gram, "null argument");
    const int m = *nens;
    CG_REQUIRE(m >= 2 && m <= kEnkfMax, "ensemble size out of range (2..64)");
    Members fl = analysis_members(m, handles);
    Flow *f0 = fl[0].get();
    const int n = f0->n, ne = member_width(m);
    hipStream_t st = f0->st;
    std::array<const double *, kEnkfMax> hx;
    for (int j = 0; j < m; ++j) hx[j] = fl[j]->pnew.p;
    const int nb = std::max(1, std::min(kGramGroups, (n + kGramRows - 1) / kGramRows));
    DBuf<const double *> dx(m);
    DBuf<double> part((size_t)nb * ne * ne), g((size_t)ne * ne);
    dx.upload(hx.data(), m, st);
    switch (ne) {
      case 8: hipLaunchKernelGGL(k_enkf_gram<8>, dim3(nb), dim3(256), 0, st, n, m, dx.p, part.p); break;
      case 16: hipLaunchKernelGGL(k_enkf_gram<16>, dim3(nb), dim3(256), 0, st, n, m, dx.p, part.p); break;
      case 32: hipLaunchKernelGGL(k_enkf_gram<32>, dim3(nb), dim3(256), 0, st, n, m, dx.p, part.p); break;
      default: hipLaunchKernelGGL(k_enkf_gram<64>, dim3(nb), dim3(256), 0, st, n, m, dx.p, part.p); break;
    }
    hipLaunchKernelGGL(k_enkf_gram_sum, dim3((ne * ne + 255) / 256), dim3(256), 0, st, nb, ne * ne, part.p, g.p);
    CG_HIP(hipGetLastError());
    std::vector<double> h((size_t)ne * ne);
    g.download(h.data(), h.size(), st);
    CG_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) gram[(size_t)i * m + j] = h[(size_t)i * ne + j];
    return 0;
  });
}

int cathy_gpu_pf_resample(const int *nens, int64_t *handles, const int *idx, const double *jitter_std,
                          const int64_t *seed) {
  return guarded([&] {
    CG_REQUIRE(nens && handles && idx && jitter_std && seed, "null argument");
    const int m = *nens;
    CG_REQUIRE(m >= 2 && m <= kEnkfMax, "ensemble size out of range (2..64)");
    for (int j = 0; j < m; ++j) CG_REQUIRE(idx[j] >= 0 && idx[j] < m, "resampled member index out of range");
    CG_REQUIRE(std::isfinite(*jitter_std) && *jitter_std >= 0.0, "the jitter deviation must be finite and >= 0");
    Members fl = analysis_members(m, handles);
    Flow *f0 = fl[0].get();
    const int n = f0->n;
    hipStream_t st = f0->st;
    std::array<const double *, kEnkfMax> hs;
    std::array<double *, kEnkfMax> hd;
    for (int j = 0; j < m; ++j) {
      hs[j] = fl[idx[j]]->pnew.p;
      hd[j] = fl[j]->pnew.p;
    }
    DBuf<const double *> dsrc(m);
    DBuf<double *> ddst(m);
    DBuf<int> dip(n);
    dsrc.upload(hs.data(), m, st);
    ddst.upload(hd.data(), m, st);
    dip.upload(f0->iperm.data(), n, st);
    const dim3 g(grid_for(n)), b(kBlock);
    const double sg = *jitter_std;
    const uint64_t sd = (uint64_t)*seed;
    switch (member_width(m)) {
      case 8: hipLaunchKernelGGL(k_pf_resample<8>, g, b, 0, st, n, m, dsrc.p, dip.p, sg, sd, ddst.p); break;
      case 16: hipLaunchKernelGGL(k_pf_resample<16>, g, b, 0, st, n, m, dsrc.p, dip.p, sg, sd, ddst.p); break;
      case 32: hipLaunchKernelGGL(k_pf_resample<32>, g, b, 0, st, n, m, dsrc.p, dip.p, sg, sd, ddst.p); break;
      default: hipLaunchKernelGGL(k_pf_resample<64>, g, b, 0, st, n, m, dsrc.p, dip.p, sg, sd, ddst.p); break;
    }
    CG_HIP(hipGetLastError());
    CG_HIP(hipStreamSynchronize(st));
    for (int j = 0; j < m; ++j) fl[j]->adopt_state();
    return 0;
  });
}

int cathy_gpu_flow3d_adopt_state(int64_t *handle) {
  return guarded([&] {
    lookup(handle)->adopt_state();
    return 0;
  });
}

int cathy_gpu_flow3d_set_pnew(int64_t *handle, const double *pnew) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(pnew, "null argument");
    std::vector<double> h(f->n);
    for (int i = 0; i < f->n; ++i) h[f->perm[i]] = pnew[i];
    f->pnew.upload(h, f->st);
    CG_HIP(hipStreamSynchronize(f->st));
    return 0;
  });
}

int cathy_gpu_flow3d_profile(int64_t *handle, const int *enable, double *spmv_ms, int64_t *launches) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    if (spmv_ms) *spmv_ms = f->spmv_ms;
    if (launches) *launches = f->spmv_launches;
    if (enable) {
      f->prof = *enable != 0;
      f->spmv_ms = 0.0;
      f->spmv_launches = 0;
      f->pev_used = 0;
    }
    return 0;
  });
}

int cathy_gpu_flow3d_spmv_bytes(int64_t *handle, double *bytes) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(bytes, "null argument");
    // values (8 B) + column ids (4 B) per entry, row pointers, x read, y write
    *bytes = 12.0 * f->nnz + 4.0 * (f->n + 1) + 16.0 * f->n;
    return 0;
  });
}

int cathy_gpu_flow3d_layout(int64_t *handle, int64_t *out6) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(out6, "null argument");
    out6[0] = f->n;
    out6[1] = f->nnz;
    out6[2] = f->nnz_sell;
    out6[3] = f->maxrow;
    out6[4] = f->coded;
    out6[5] = f->code_mismatch;
    return 0;
  });
}

int cathy_gpu_calib_copy(const int64_t *nbytes, const int *reps, double *ms_per_launch) {
  return guarded([&] {
    CG_REQUIRE(nbytes && reps && ms_per_launch && *nbytes >= 8 && *reps >= 1, "bad argument");
    const size_t n = (size_t)(*nbytes / 8);
    DBuf<double> a(n), b(n);
    hipStream_t s;
    CG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    CG_HIP(hipEventCreate(&e0));
    CG_HIP(hipEventCreate(&e1));
    const int g = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(kBlock), 0, s, n, 1.0, a.p);
    CG_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < *reps; ++i) hipLaunchKernelGGL(k_copy, dim3(g), dim3(kBlock), 0, s, n, a.p, b.p);
    CG_HIP(hipEventRecord(e1, s));
    CG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CG_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_launch = ms / *reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    return 0;
  });
}

// One SYMSLV set-up per calling thread, rebuilt when N, NTERM or the
// pattern (TOPOL / JA checksum) changes.
static uint64_t pattern_key(int n, int nterm, const int *topol, const int *ja) {
  uint64_t h = 1469598103934665603ull ^ (uint64_t)n ^ ((uint64_t)nterm << 32);
  for (int i = 0; i <= n; ++i) h = (h ^ (uint32_t)topol[i]) * 1099511628211ull;
  for (int i = 0; i < nterm; ++i) h = (h ^ (uint32_t)ja[i]) * 1099511628211ull;
  return h;
}
static thread_local std::unique_ptr<Flow> t_sym;
static thread_local uint64_t t_sym_key = 0;

int cathy_gpu_symslv(const int *n, const int *nterm, const int *numdir, int *niter, const int *itmxcg,
                     const double *tolcg, const double *rmin, const int *noddir, const int *ja, const int *topol,
                     double *pnew, const double *tnoti, const double *coef1, double *err, double *res) {
  return guarded([&] {
    CG_REQUIRE(n && nterm && numdir && niter && itmxcg && tolcg && rmin && ja && topol && pnew && tnoti && coef1 &&
                   err && res,
               "null argument");
    CG_REQUIRE(*numdir >= 0 && (*numdir == 0 || noddir), "NUMDIR / NODDIR");
    CG_REQUIRE(*itmxcg >= 1, "ITMXCG must be >= 1");
    const uint64_t key = pattern_key(*n, *nterm, topol, ja);
    if (!t_sym || key != t_sym_key) {
      t_sym.reset();
      auto f = std::make_unique<Flow>();
      f->setup_csr(*n, *nterm, topol, ja);
      t_sym = std::move(f);
      t_sym_key = key;
    }
    return t_sym->symslv(*numdir, noddir, *itmxcg, *tolcg, *rmin, tnoti, coef1, pnew, niter, err, res);
  });
}

int cathy_gpu_symslv_release(void) {
  return guarded([&] {
    t_sym.reset();
    t_sym_key = 0;
    return 0;
  });
}

int cathy_gpu_flow3d_set_bc(int64_t *handle, const int *kind, const int *nrec, const double *times,
                            const int *counts, const int *nodes, const double *values) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(kind && nrec && times && counts, "null argument");
    f->set_bc(*kind, *nrec, times, counts, nodes, values);
    return 0;
  });
}

int cathy_gpu_flow3d_set_seepage(int64_t *handle, const int *nsf, const int *counts, const int *nodes,
                                 const int *isfcvg) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(nsf && isfcvg, "null argument");
    f->set_seepage(*nsf, counts, nodes, *isfcvg);
    return 0;
  });
}

int cathy_gpu_flow3d_get_seepage(int64_t *handle, int *sfex_out, double *sfq_out) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    const int m = f->nsf_nodes;
    if (m > 0 && sfex_out) f->sfex.download(sfex_out, m, f->st);
    if (m > 0 && sfq_out) f->sfq.download(sfq_out, m, f->st);
    CG_HIP(hipStreamSynchronize(f->st));
    return m;
  });
}

int cathy_gpu_flow3d_set_vegetation(int64_t *handle, const int *nveg, const int *veg_type, const double *pcana,
                                    const double *pcref, const double *pcwlt, const double *zroot,
                                    const double *pz, const double *omgc, const double *z) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(nveg, "null argument");
    const double *tab[6] = {pcana, pcref, pcwlt, zroot, pz, omgc};
    f->set_vegetation(*nveg, veg_type, tab, z);
    return 0;
  });
}

int cathy_gpu_flow3d_set_surface(int64_t *handle, const int *surf) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(surf && (*surf == 0 || *surf == 1), "SURF must be 0 or 1");
    f->surf = *surf == 1;
    f->initialised = 0;
    return 0;
  });
}

int cathy_gpu_flow3d_set_routing(int64_t *handle, const int64_t *route, const int *nrow, const int *ncol,
                                 const double *delta_x, const double *delta_y, const int *ncell_coarse,
                                 const int *cell, const int *indcel, const int *indcelwl, const double *dem_map,
                                 const int *ntri, const int *triang) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(route && nrow && ncol && delta_x && delta_y && ncell_coarse && ntri, "null argument");
    const RouteLink link = route_link(route);
    f->set_routing(link, *nrow, *ncol, *delta_x, *delta_y, *ncell_coarse, cell, indcel, indcelwl, dem_map, *ntri,
                   triang);
    f->initialised = 0;
    return 0;
  });
}

int cathy_gpu_flow3d_routing_state(int64_t *handle, double *pondnod, double *ovflnod, double *h_water, int *info) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(f->rt, "no routing linked (cathy_gpu_flow3d_set_routing)");
    Flow::Routing &R = *f->rt;
    if (pondnod) R.pondnod.download(pondnod, f->nnod, f->st);
    if (ovflnod) R.ovflnod.download(ovflnod, f->nnod, f->st);
    if (h_water) CG_HIP(hipMemcpyAsync(h_water, R.state.p + 9 * (size_t)R.ngrid, sizeof(double) * R.ngrid,
                                       hipMemcpyDeviceToHost, f->st));
    CG_HIP(hipStreamSynchronize(f->st));
    if (info) {
      info[0] = R.routed;
      info[1] = R.pondp;
    }
    return 0;
  });
}

int cathy_gpu_flow3d_bench_spmv(int64_t *handle, const int *nrep, const double *flush_mb, double *out2) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(nrep && flush_mb && out2 && *flush_mb >= 0.0, "bad argument");
    f->bench_spmv(*nrep, (size_t)(*flush_mb * 1e6), out2);
    return 0;
  });
}

int cathy_gpu_flow3d_set_pcg_batches(int64_t *handle, const int *first_pct, const int *followup) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(first_pct && followup, "null argument");
    CG_REQUIRE(*first_pct >= 10 && *first_pct <= 400, "FIRST_PCT must lie in 10..400");
    CG_REQUIRE(*followup >= 1 && *followup <= 64, "FOLLOWUP must lie in 1..64");
    f->pcg_first_pct = *first_pct;
    f->pcg_followup = *followup;
    return 0;
  });
}

int cathy_gpu_flow3d_counters(int64_t *handle, double *out6) {
  return guarded([&] {
    const std::shared_ptr<Flow> f = lookup(handle);
    CG_REQUIRE(out6, "null argument");
    for (int i = 0; i < 6; ++i) out6[i] = f->last_counters[i];
    return 0;
  });
}

}  // extern "C"
