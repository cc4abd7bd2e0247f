// route.hip -- surface routing substep (ISIMGR = 2) on gfx950: ROUTE
// (route.f:4-234, Muskingum-Cunge MC mc.f:8-63) and ALTEZZE
// (altezze.f:4-60) for cells without reservoirs (TIPO_R = 0).
//
// The reference visits the cells in QOI_SN order (highest first); each
// cell routes a cardinal (W_1) and a diagonal (W_2) branch and adds each
// branch's outflow to its downstream cell's Q_IN_KKP1_SN.  A cell's routing
// reads only the inflows of cells before it, so the network is a DAG in
// QOI_SN order.  The host assigns each cell a level (1 + the deepest level
// among the cells that feed it) and lists each cell's feeders in QOI_SN
// order, branch 1 before 2.  One work-group routes one ensemble member:
// level by level, each thread gathers its cell's inflow from the feeders'
// outflows in that order (the reference's summation order, so the inflow
// is bitwise the reference's for the same outflows), routes both branches
// and stores them; a barrier separates the levels.  ALTEZZE follows over
// every grid cell, with its own cell-type array (the reference's time loop
// passes CELTYPE, which is zero everywhere, so by default every cell takes
// the volume balance).  CU_MAX / AK_MAX keep the reference's rule (the last
// branch in visiting order whose CU is >= the running maximum, starting
// from 0) as a (CU, position) maximum.

#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <memory>
#include <mutex>

#include "cathy_gpu.h"
#include "common.h"
#include "route_link.h"

namespace cg {
namespace {

// per-cell network parameters, in level order (structure of arrays), two
// branches each: W, EPL and the flow-independent factors of MC (set-up)
enum RouteParam { kW1, kW2, kEpl1, kEpl2, kCk1, kCk2, kEq1, kEq2, kEb1, kEb2, kDh1, kDh2, kNrc, kNParam };
// DTM grids given to cathy_gpu_route_create, in argument order
enum DtmGrid { gW1, gW2, gSlope1, gSlope2, gEpl1, gEpl2, gKs1, gKs2, gWs1, gWs2, gB1, gY1, gNrc, kNGrid };

constexpr int kRouteBlock = 512;
// A cell is fed by at most its 4 cardinal neighbours' W_1 branches and its
// 4 diagonal neighbours' W_2 branches.
constexpr int kMaxFeed = 8;

// Reservoirs (SERBATOIO / CALCOLI); res == nullptr: none.
struct ResDev {
  int maxres;               // leading dimension of the height-area tables
  const int *n_ha;          // [numres] table entries
  const double *hres, *ares;  // HRES / ARES (MAXRES, 30) column-major
  const double *h_soglia, *h_fondo;
  const int *rid;           // [ncell] reservoir of order index k, -1 if none
  const int *rflag;         // [ncell] bit br: branch br goes through the reservoir; bit 2+br: it sets the pool level
  const int *alt;           // [ngrid] ALTEZZE type: -1 volume balance, 2r pool depth of reservoir r
                            // (type 1..999), 2r+1 lake depth (type >= 1000); null: all -1
  const double *elevation;  // [ngrid]
};

struct RouteDev {
  int ncell;                // routed cells
  int ngrid;                // cells of the grid (nrow * ncol): state arrays
  int nlev;
  const int *lev_ptr;       // [nlev + 1] level k holds order indices lev_ptr[k]..lev_ptr[k+1]-1
  const int *cell;          // [ncell] grid index (I_BASIN - 1) of order index k
  const int *key;           // [ncell] 2 * QOI_SN position
  const int *pred;          // [ncell][kMaxFeed] feeders in order: grid index * 2 + branch (0: W_1,
                            // 1: W_2), then -1
  const double *prm;        // [kNParam][ncell]
  int numres;
  ResDev res;
  // branches draining to a cell QOI_SN does not route (a no-data cell, or
  // where the outlet's diagonal branch lands past the grid's edge): pairs
  // (source grid index * 2 + branch, target grid index), in QOI_SN order
  int nsink;
  const int *sink;
};

// The flow-independent factors of MC for one branch (mc.f:40-49), in the
// reference's evaluation order: CK = ck0 * QC**eq with
// ck0 = ((5/(3G) * KS**0.6) * W**-0.4) * SIN(BETA)**0.3, eq = 1 - 3G/5;
// DH = QC**eb / dh0 with eb = 1 - B1, dh0 = ((2G) W) TAN(BETA).
struct McConst {
  double ck0, eq, eb, dh0;
};

McConst mc_const(double local_slope, double ks, double w, double b1, double y1) {
  const double beta = atan(local_slope);
  const double g = (1.0 - y1 + 2.0 / 3.0 * b1);
  McConst c;
  c.ck0 = 5.0 / (3.0 * g) * pow(ks, 3.0 / 5.0) * pow(w, -2.0 / 5.0) * pow(sin(beta), 3.0 / 1.0e1);
  c.eq = 1.0 - 3.0 * g / 5.0;
  c.eb = 1.0 - b1;
  c.dh0 = 2 * g * w * tan(beta);
  return c;
}

// MC (mc.f:8-63): coefficients of one branch and its outflow.
__device__ __forceinline__ double mc_outflow(double ck0, double eq, double eb, double dh0, double epl,
                                             double deltat, double q_in_kk, double q_in_kkp1, double q_out_kk,
                                             double q_overland, double &cu, double &ak) {
  double qc = 1.0 / 3.0 * (q_in_kk + q_in_kkp1 + q_out_kk);
  if (qc <= 1.0e-05) qc = 1.0e-05;
  const double ck = ck0 * pow(qc, eq);
  ak = ck / epl;
  cu = ck * deltat / epl;
  double dh = pow(qc, eb) / dh0;
  if (dh < (1.0 - cu)) dh = 1.0 - cu;
  const double x = 0.50 - dh / (ck * epl);
  const double den = 2.0 * (1.0 - x) + cu;
  const double c1 = (cu - 2.0 * x) / den;
  const double c2 = (cu + 2.0 * x) / den;
  const double c3 = (2.0 * (1.0 - x) - cu) / den;
  const double c4 = (2.0 * ck * deltat) / den;
  return c1 * q_in_kkp1 + c2 * q_in_kk + c3 * q_out_kk + c4 * q_overland;
}

// CALCOLI (calcoli.f:4-76): rate of change of the pool level; false where
// the reference stops (level at or below the table's first height).
__device__ bool calcoli(double hh, double &dd, double delta_t, int n_ha, const double *ah_r, const double *ha_r,
                        int ld, double q_in_kk, double t, double h_sfioro, double q_in_kkp1, double surface_water,
                        double h_fondo_rr) {
  double q_inflow_p = q_in_kk + (q_in_kkp1 - q_in_kk) / delta_t * t;
  q_inflow_p = q_inflow_p + surface_water;
  if (hh == h_fondo_rr && q_inflow_p < 0.0) {
    dd = 0.0;
    return true;
  }
  double ah_r_p = 0.0;
  bool found = false;
  for (int j = 0; j < n_ha - 1 && !found; ++j) {
    const double h0 = ha_r[j * ld], h1 = ha_r[(j + 1) * ld];
    if (h0 < hh && h1 >= hh) {
      ah_r_p = ah_r[j * ld] + (ah_r[(j + 1) * ld] - ah_r[j * ld]) / (h1 - h0) * (hh - h0);
      found = true;
    }
  }
  if (!found) {
    if (hh > ha_r[(n_ha - 1) * ld]) ah_r_p = ah_r[(n_ha - 1) * ld];
    else return false;
  }
  const double qh_r_p = hh <= h_sfioro ? 0.0 : (q_inflow_p >= 0.0 ? q_inflow_p : 0.0);
  dd = 1 / ah_r_p * (q_inflow_p - qh_r_p);
  return true;
}

// SERBATOIO (serbatoio.f:6-119): RK4 level-pool step of reservoir r, whose
// fourth stage takes half of the third increment as the reference does.
__device__ bool serbatoio(const ResDev &R, int r, double q_in_kk, double q_in_kkp1, double h_pool_kk,
                          double delta_t, double surface_water, double &q_out_kkp1, double &h_pool_kkp1) {
  const int n_ha = R.n_ha[r], ld = R.maxres;
  const double *ha_r = R.hres + r, *ah_r = R.ares + r;
  const double h_sfioro = R.h_soglia[r], h_fondo = R.h_fondo[r];
  double dd, hh = h_pool_kk;
  if (!calcoli(hh, dd, delta_t, n_ha, ah_r, ha_r, ld, q_in_kk, 0.0, h_sfioro, q_in_kkp1, surface_water, h_fondo))
    return false;
  const double dh1 = delta_t * dd;
  hh = h_pool_kk + 0.50 * dh1;
  if (hh < h_fondo) hh = h_fondo;
  if (!calcoli(hh, dd, delta_t, n_ha, ah_r, ha_r, ld, q_in_kk, 0.0 + 0.50 * delta_t, h_sfioro, q_in_kkp1,
               surface_water, h_fondo))
    return false;
  const double dh2 = delta_t * dd;
  hh = h_pool_kk + 0.50 * dh2;
  if (hh < h_fondo) hh = h_fondo;
  if (!calcoli(hh, dd, delta_t, n_ha, ah_r, ha_r, ld, q_in_kk, 0.0 + 0.50 * delta_t, h_sfioro, q_in_kkp1,
               surface_water, h_fondo))
    return false;
  const double dh3 = delta_t * dd;
  hh = h_pool_kk + 0.50 * dh3;
  if (hh < h_fondo) hh = h_fondo;
  if (!calcoli(hh, dd, delta_t, n_ha, ah_r, ha_r, ld, q_in_kk, delta_t, h_sfioro, q_in_kkp1, surface_water,
               h_fondo))
    return false;
  const double dh4 = delta_t * dd;
  double h = h_pool_kk + dh1 / 6 + dh2 / 3 + dh3 / 3 + dh4 / 6;
  if (h < h_fondo) h = h_fondo;
  h_pool_kkp1 = h;
  q_out_kkp1 = h <= h_sfioro ? 0.0 : (q_in_kkp1 + surface_water >= 0.0 ? q_in_kkp1 + surface_water : 0.0);
  return true;
}

struct CuBest {
  double cu, ak;
  int key;  // -1: no branch routed (AK_MAX unchanged)
};

__device__ __forceinline__ bool cu_beats(double cu, int key, const CuBest &b) {
  return cu > b.cu || (cu == b.cu && key > b.key);
}

__global__ __launch_bounds__(kRouteBlock) void k_route(
    RouteDev d, double deltat, double dx, double dy, const double *__restrict__ sw_all,
    const double *__restrict__ qin_kk_all, const double *__restrict__ qout_kk1_all,
    const double *__restrict__ qout_kk2_all, const double *__restrict__ vol_kk_all, double *qin_kkp1_all,
    double *qout1_all, double *qout2_all, double *__restrict__ vol_kkp1_all, double *__restrict__ h_all,
    double *__restrict__ cu_ak, const double *__restrict__ hpool_kk_all, double *__restrict__ hpool_kkp1_all,
    int *__restrict__ err) {
  const size_t off = (size_t)blockIdx.x * d.ngrid;
  const double *sw = sw_all + off, *qin_kk = qin_kk_all + off;
  const double *qout_kk[2] = {qout_kk1_all + off, qout_kk2_all + off};
  double *qin = qin_kkp1_all + off;
  double *qout[2] = {qout1_all + off, qout2_all + off};
  const double *hpool_kk = d.numres ? hpool_kk_all + (size_t)blockIdx.x * d.numres : nullptr;
  double *hpool = d.numres ? hpool_kkp1_all + (size_t)blockIdx.x * d.numres : nullptr;
  __shared__ int s_err;
  if (threadIdx.x == 0) s_err = 0;
  __syncthreads();
  CuBest best{0.0, 0.0, -1};
  int kbeg = d.lev_ptr[0], kend = d.lev_ptr[1];
  for (int lv = 0; lv < d.nlev; ++lv) {
    const int knext = lv + 2 <= d.nlev ? d.lev_ptr[lv + 2] : 0;  // consumed after the barrier
    for (int k = kbeg + (int)threadIdx.x; k < kend; k += blockDim.x) {
      const int c = d.cell[k];
      // the feeder slots, then all their outflows, as two batches of loads
      // (empty slots read entry 0); the sum runs over the feeders in order
      int src[kMaxFeed];
#pragma unroll
      for (int u = 0; u < kMaxFeed; ++u) src[u] = d.pred[(size_t)k * kMaxFeed + u];
      double vin[kMaxFeed];
#pragma unroll
      for (int u = 0; u < kMaxFeed; ++u) {
        const int sidx = src[u] >= 0 ? src[u] : 0;
        vin[u] = qout[sidx & 1][sidx >> 1];
      }
      double q_in = qin[c];
#pragma unroll
      for (int u = 0; u < kMaxFeed; ++u)
        if (src[u] >= 0) q_in = q_in + vin[u];
      const double *P = d.prm;
      const int n = d.ncell;
      const double nrc = P[kNrc * n + k];
      const double swc = sw[c] / nrc;
      const int rflag = d.numres ? d.res.rflag[k] : 0;
      for (int br = 0; br < 2; ++br) {
        const double w = P[(kW1 + br) * n + k];
        if (w == 0) continue;
        if (rflag & (1 << br)) {  // route.f:101-126: through the reservoir, no NRC scaling of the outflow
          const int r = d.res.rid[k];
          double q, h;
          if (!serbatoio(d.res, r, qin_kk[c] * w / nrc, q_in * w / nrc, hpool_kk[r], deltat, swc, q, h)) {
            s_err = 1;
            q = h = 0.0;
          }
          if (q < 0.0) q = 0.0;
          if (rflag & (4 << br)) hpool[r] = h;
          qout[br][c] = q;
          continue;
        }
        const double epl = P[(kEpl1 + br) * n + k];
        const double cr = 1.0 / epl;
        const double q_overland = swc * w * cr;
        const double qi_kk = qin_kk[c] * w / nrc;
        const double qo_kk = qout_kk[br][c] / nrc;
        const double qi_kkp1 = q_in * w / nrc;
        double cu, ak;
        double q = mc_outflow(P[(kCk1 + br) * n + k], P[(kEq1 + br) * n + k], P[(kEb1 + br) * n + k],
                              P[(kDh1 + br) * n + k], epl, deltat, qi_kk, qi_kkp1, qo_kk, q_overland, cu, ak);
        if (q < 0.0) q = 0.0;
        qout[br][c] = q * nrc;
        const int key = d.key[k] + br;
        if (cu >= 0.0 && cu_beats(cu, key, best)) best = CuBest{cu, ak, key};
      }
      qin[c] = q_in;
    }
    __syncthreads();
    kbeg = kend;
    kend = knext;
  }
  // route.f:133-137, 244-246: the outflows into unrouted cells, in the
  // order the reference adds them
  if (d.nsink) {
    if (threadIdx.x == 0)
      for (int u = 0; u < d.nsink; ++u) {
        const int src = d.sink[2 * u], tgt = d.sink[2 * u + 1];
        qin[tgt] = qin[tgt] + qout[src & 1][src >> 1];
      }
    __syncthreads();
  }
  // ALTEZZE (altezze.f:33-105) over every grid cell; reservoir and lake
  // depths only where the caller's type array asks for them
  for (int c = threadIdx.x; c < d.ngrid; c += blockDim.x) {
    const int alt = d.res.alt ? d.res.alt[c] : -1;
    if (alt >= 0) {  // reservoir outlet (depth clamped below at 0) or lake cell (positive depths only)
      const double hw = hpool[alt >> 1] - d.res.elevation[c];
      h_all[off + c] = (alt & 1) ? (hw > 0.0 ? hw : 0.0) : (hw < 0 ? 0.0 : hw);
      continue;
    }
    const double dv = (qin_kk[c] + qin[c]) / 2 * deltat + sw[c] * deltat -
                      (qout_kk[0][c] + qout_kk[1][c]) / 2 * deltat - (qout[0][c] + qout[1][c]) / 2 * deltat;
    const double v = vol_kk_all[off + c] + dv;
    vol_kkp1_all[off + c] = v >= 0.0 ? v : 0.0;
    h_all[off + c] = v >= 0.0 ? v / (dx * dy) : 0.0;
  }
  // (CU, key) maximum over the work-group
  __shared__ double s_cu[kRouteBlock / 64], s_ak[kRouteBlock / 64];
  __shared__ int s_key[kRouteBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    const double cu2 = __shfl_down(best.cu, o), ak2 = __shfl_down(best.ak, o);
    const int key2 = __shfl_down(best.key, o);
    if (key2 >= 0 && cu_beats(cu2, key2, best)) best = CuBest{cu2, ak2, key2};
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    s_cu[wv] = best.cu;
    s_ak[wv] = best.ak;
    s_key[wv] = best.key;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    CuBest b{0.0, 0.0, -1};
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
      if (s_key[w] >= 0 && cu_beats(s_cu[w], s_key[w], b)) b = CuBest{s_cu[w], s_ak[w], s_key[w]};
    cu_ak[2 * blockIdx.x] = b.key >= 0 ? b.cu : 0.0;
    if (b.key >= 0) cu_ak[2 * blockIdx.x + 1] = b.ak;
    err[blockIdx.x] = s_err;
  }
}

struct Route {
  int nrow = 0, ncol = 0, ncell = 0, ngrid = 0, nlev = 0, maxpred = 0;
  std::mutex mu;  // steps and reservoir set-up on one handle share its device buffers
  hipStream_t st = nullptr;
  DBuf<int> lev_ptr, cell, key, pred, sink;
  int nsink = 0;
  DBuf<double> prm;
  // per-step state [member][ngrid]: 5 inputs, 4 in/out (volume_kkp1 keeps
  // reservoir cells), 1 output; cu_ak; pool levels [member][numres] x 2
  DBuf<double> state, cu_ak, hpool;
  DBuf<int> err;
  int cap_members = 0;
  // reservoirs
  int numres = 0, maxres = 0;
  std::vector<int> pos_of_k, cellh;  // QOI_SN position and grid index of order index k
  std::vector<double> wgrid[2];      // W_1 / W_2 by grid index
  std::vector<int> reservr_h;        // RESERVR by grid index (NUM_R of ALTEZZE)
  bool typed = false;                // ALTEZZE types set (some cell not 0)
  DBuf<int> r_nha, r_rid, r_rflag, r_alt;
  DBuf<double> r_hres, r_ares, r_soglia, r_fondo, r_elev;

  ~Route() {
    if (st) (void)hipStreamDestroy(st);
  }

  RouteDev dev() const {
    return RouteDev{ncell, ngrid, nlev, lev_ptr.p, cell.p, key.p, pred.p, prm.p, numres,
                    ResDev{maxres, r_nha.p, r_hres.p, r_ares.p, r_soglia.p, r_fondo.p, r_rid.p, r_rflag.p,
                           typed ? r_alt.p : nullptr, r_elev.p},
                    nsink, sink.p};
  }
};

// A call in flight holds its Route: a destroy from another thread removes
// the handle, and the Route (its buffers and mutex) ends with the last call.
std::mutex g_route_mu;
std::map<int64_t, std::shared_ptr<Route>> g_routes;
int64_t g_route_next = 1;

std::shared_ptr<Route> route_lookup(const int64_t *h) {
  std::lock_guard<std::mutex> lk(g_route_mu);
  if (!h) throw Error(CATHY_GPU_EHANDLE, "null route handle");
  auto it = g_routes.find(*h);
  if (it == g_routes.end()) throw Error(CATHY_GPU_EHANDLE, "unknown route handle");
  return it->second;
}

// Cell (I, J) of I_BASIN (route.f:43-51).
inline void basin_ij(int ib, int nrow, int &i, int &j) {
  const int jr = ib % nrow;
  j = jr != 0 ? jr : nrow;
  i = jr != 0 ? (ib - j) / nrow + 1 : ib / nrow;
}

}  // namespace

RouteLink route_link(const int64_t *handle) {
  std::shared_ptr<Route> r = route_lookup(handle);
  RouteLink l;
  l.route = r.get();
  l.nrow = r->nrow;
  l.ncol = r->ncol;
  l.ngrid = r->ngrid;
  l.ncell = r->ncell;
  l.numres = r->numres;
  l.typed = r->typed;
  l.keep = std::static_pointer_cast<void>(r);
  return l;
}

void route_substep(const RouteLink &l, hipStream_t st, double deltat, double dx, double dy, double *state,
                   double *cu_ak, int *err) {
  Route *r = static_cast<Route *>(l.route);
  CG_REQUIRE(r->numres == 0, "the coupled step routes networks without reservoirs");
  std::lock_guard<std::mutex> lk(r->mu);
  const size_t g = r->ngrid;
  double *S = state;
  hipLaunchKernelGGL(k_route, dim3(1), dim3(kRouteBlock), 0, st, r->dev(), deltat, dx, dy, S, S + g, S + 2 * g,
                     S + 3 * g, S + 4 * g, S + 5 * g, S + 6 * g, S + 7 * g, S + 8 * g, S + 9 * g, cu_ak,
                     (const double *)nullptr, (double *)nullptr, err);
  CG_HIP(hipGetLastError());
}
}  // namespace cg

using namespace cg;

extern "C" {

int cathy_gpu_route_create(int64_t *handle, const int *nrow, const int *ncol, const int *ldc, const int *ncell,
                           const int *qoi_sn, const int *p_outflow_1, const int *p_outflow_2, const double *w_1,
                           const double *w_2, const double *local_slope_1, const double *local_slope_2,
                           const double *epl_1, const double *epl_2, const double *kss1_sf_1,
                           const double *kss1_sf_2, const double *ws1_sf_1, const double *ws1_sf_2,
                           const double *b1_sf, const double *y1_sf, const double *nrc) {
  return guarded([&] {
    CG_REQUIRE(handle && nrow && ncol && ldc && ncell && qoi_sn && p_outflow_1 && p_outflow_2 && w_1 && w_2 &&
                   local_slope_1 && local_slope_2 && epl_1 && epl_2 && kss1_sf_1 && kss1_sf_2 && ws1_sf_1 &&
                   ws1_sf_2 && b1_sf && y1_sf && nrc,
               "null argument");
    const int nr = *nrow, nc = *ncol, ld = *ldc, n = *ncell;
    CG_REQUIRE(nr >= 1 && nc >= 1 && ld >= nc && (int64_t)nr * nc <= (1 << 29), "grid size out of range");
    const int ngrid = nr * nc;
    CG_REQUIRE(n >= 1 && n <= ngrid, "NCELL must lie in 1..NROW*NCOL");
    std::vector<int> pos(ngrid, -1);
    for (int p = 0; p < n; ++p) {
      const int ib = qoi_sn[p];
      CG_REQUIRE(ib >= 1 && ib <= ngrid, "QOI_SN entry outside the grid");
      CG_REQUIRE(pos[ib - 1] < 0, "QOI_SN lists a cell twice");
      pos[ib - 1] = p;
    }
    const double *grids[kNGrid] = {w_1, w_2, local_slope_1, local_slope_2, epl_1, epl_2, kss1_sf_1,
                                   kss1_sf_2, ws1_sf_1, ws1_sf_2, b1_sf, y1_sf, nrc};
    // feeders of each cell in QOI_SN order, branch 1 before 2 (route.f:133-137, 220-221)
    std::vector<std::vector<int>> feeders(n);
    std::vector<int> level(n, 0), sinks;
    int nlev = 1;
    for (int p = 0; p < n; ++p) {
      const int ib = qoi_sn[p];
      int i, j;
      basin_ij(ib, nr, i, j);
      const size_t g = (size_t)(i - 1) + (size_t)(j - 1) * ld;
      for (int br = 0; br < 2; ++br) {
        if ((br ? w_2[g] : w_1[g]) == 0) continue;
        if (br == 0 && p == n - 1) continue;  // the last cell's cardinal outflow leaves the network
        const int code = br ? p_outflow_2[g] : p_outflow_1[g];
        const int di = (int)lroundf((float)(code - 5) / 3.0f);  // NINT((P-5)/3.)
        const int dj = code - 5 - 3 * di;
        const int i2 = i + di, j2 = j + dj;
        // the downstream I_BASIN as route.f forms it (:86-88), also past the
        // grid's edge, where it names another cell of the grid
        const int64_t tb = (int64_t)(i2 - 1) * nr + j2;
        if (tb < 1 || tb > ngrid) throw Error(-1, "cell " + std::to_string(ib) + " drains off the grid");
        const int t = pos[tb - 1];
        if (t < 0) {  // a cell QOI_SN does not route: its inflow only enters ALTEZZE
          sinks.push_back(((i - 1) * nr + j - 1) * 2 + br);
          sinks.push_back((int)tb - 1);
          continue;
        }
        if (t <= p)
          throw Error(-1, "cell " + std::to_string(ib) + " drains to a cell not after it in QOI_SN");
        feeders[t].push_back(((i - 1) * nr + j - 1) * 2 + br);
        level[t] = std::max(level[t], level[p] + 1);
        nlev = std::max(nlev, level[t] + 1);
      }
    }
    // order: by level, then QOI_SN position
    std::vector<int> lev_ptr(nlev + 1, 0);
    for (int p = 0; p < n; ++p) lev_ptr[level[p] + 1]++;
    for (int l = 0; l < nlev; ++l) lev_ptr[l + 1] += lev_ptr[l];
    std::vector<int> slot(lev_ptr.begin(), lev_ptr.end() - 1), of(n);
    for (int p = 0; p < n; ++p) of[p] = slot[level[p]]++;
    std::vector<int> cell(n), key(n), pred((size_t)n * kMaxFeed, -1);
    std::vector<double> prm((size_t)kNParam * n);
    int maxpred = 0;
    std::vector<int> by_k(n);
    for (int p = 0; p < n; ++p) by_k[of[p]] = p;
    for (int k = 0; k < n; ++k) {
      const int p = by_k[k];
      const int ib = qoi_sn[p];
      int i, j;
      basin_ij(ib, nr, i, j);
      const size_t g = (size_t)(i - 1) + (size_t)(j - 1) * ld;
      cell[k] = ib - 1;
      key[k] = 2 * p;
      auto put = [&](int q, double v) { prm[(size_t)q * n + k] = v; };
      put(kNrc, grids[gNrc][g]);
      for (int br = 0; br < 2; ++br) {
        const double w = grids[gW1 + br][g];
        put(kW1 + br, w);
        put(kEpl1 + br, grids[gEpl1 + br][g]);
        const McConst c = w != 0 ? mc_const(grids[gSlope1 + br][g], grids[gKs1 + br][g], grids[gWs1 + br][g],
                                            grids[gB1][g], grids[gY1][g])
                                 : McConst{0.0, 0.0, 0.0, 1.0};
        put(kCk1 + br, c.ck0);
        put(kEq1 + br, c.eq);
        put(kEb1 + br, c.eb);
        put(kDh1 + br, c.dh0);
      }
      maxpred = std::max(maxpred, (int)feeders[p].size());
      for (size_t u = 0; u < feeders[p].size() && u < (size_t)kMaxFeed; ++u) pred[(size_t)k * kMaxFeed + u] = feeders[p][u];
    }
    CG_REQUIRE(maxpred <= kMaxFeed, "a cell has more than 8 feeders");
    auto r = std::make_unique<Route>();
    r->nrow = nr;
    r->ncol = nc;
    r->ncell = n;
    r->ngrid = ngrid;
    r->nlev = nlev;
    r->maxpred = maxpred;
    r->pos_of_k.resize(n);
    for (int k = 0; k < n; ++k) r->pos_of_k[k] = key[k] / 2;
    r->cellh = cell;
    for (int br = 0; br < 2; ++br) {
      r->wgrid[br].assign(ngrid, 0.0);
      for (int jj = 1; jj <= nr; ++jj)
        for (int ii = 1; ii <= nc; ++ii)
          r->wgrid[br][(size_t)(ii - 1) * nr + jj - 1] = grids[gW1 + br][(size_t)(ii - 1) + (size_t)(jj - 1) * ld];
    }
    CG_HIP(hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking));
    r->lev_ptr.upload(lev_ptr, r->st);
    r->cell.upload(cell, r->st);
    r->key.upload(key, r->st);
    r->pred.upload(pred, r->st);
    r->prm.upload(prm, r->st);
    r->nsink = (int)sinks.size() / 2;
    if (r->nsink) r->sink.upload(sinks, r->st);
    CG_HIP(hipStreamSynchronize(r->st));
    std::lock_guard<std::mutex> lk(g_route_mu);
    *handle = g_route_next++;
    g_routes[*handle] = std::move(r);
    return 0;
  });
}

int cathy_gpu_route_set_reservoirs(int64_t *handle, const int *tipo_r, const int *reservr, const int *numres,
                                   const int *maxres, const int *n_ha, const double *hres, const double *ares,
                                   const double *h_soglia, const double *h_fondo, const double *elevation) {
  return guarded([&] {
    CG_REQUIRE(tipo_r && reservr && numres && maxres && n_ha && hres && ares && h_soglia && h_fondo && elevation,
               "null argument");
    const std::shared_ptr<Route> r = route_lookup(handle);
    std::lock_guard<std::mutex> lk(r->mu);
    const int nres = *numres, ld = *maxres, ng = r->ngrid, n = r->ncell;
    CG_REQUIRE(nres >= 1 && ld >= nres, "NUMRES must lie in 1..MAXRES");
    for (int q = 0; q < nres; ++q) CG_REQUIRE(n_ha[q] >= 2 && n_ha[q] <= 30, "N_HA must lie in 2..30");
    auto res_of = [&](int c) {  // reservoir (0-based) of grid cell c
      const int v = reservr[c] - 1;
      if (v < 0 || v >= nres) throw Error(-1, "cell " + std::to_string(c + 1) + ": RESERVR outside 1..NUMRES");
      return v;
    };
    std::vector<int> rid(n, -1), rflag(n, 0);
    std::vector<int> last_key(nres, -1), last_k(nres, -1), last_br(nres, 0);
    for (int k = 0; k < n; ++k) {
      const int c = r->cellh[k], t = tipo_r[c];
      if (t == 0) continue;
      const bool routes = r->wgrid[0][c] != 0 || r->wgrid[1][c] != 0;
      if (t >= 1000) {
        if (routes)
          throw Error(-1, "cell " + std::to_string(c + 1) + " has TIPO_R >= 1000 and routes water (route.f stops)");
        continue;
      }
      const int q = res_of(c);
      rid[k] = q;
      for (int br = 0; br < 2; ++br) {
        if (r->wgrid[br][c] == 0) continue;
        rflag[k] |= 1 << br;
        const int key = 2 * r->pos_of_k[k] + br;  // the last branch in visiting order sets the level
        if (key > last_key[q]) {
          last_key[q] = key;
          last_k[q] = k;
          last_br[q] = br;
        }
      }
    }
    for (int q = 0; q < nres; ++q)
      if (last_k[q] >= 0) rflag[last_k[q]] |= 4 << last_br[q];
    for (int c = 0; c < ng; ++c) {  // lake cells name a grid cell of their reservoir
      const int t = tipo_r[c];
      if (t >= 1000) {
        CG_REQUIRE(t - 1000 >= 1 && t - 1000 <= ng, "TIPO_R - 1000 outside the grid");
        (void)res_of(t - 1000 - 1);
      } else if (t != 0) {
        (void)res_of(c);
      }
    }
    r->numres = nres;
    r->reservr_h.assign(reservr, reservr + ng);
    r->typed = false;  // ALTEZZE types refer to these reservoirs: set them again
    r->maxres = ld;
    r->r_nha.upload(std::vector<int>(n_ha, n_ha + nres), r->st);
    r->r_hres.upload(std::vector<double>(hres, hres + (size_t)ld * 30), r->st);
    r->r_ares.upload(std::vector<double>(ares, ares + (size_t)ld * 30), r->st);
    r->r_soglia.upload(std::vector<double>(h_soglia, h_soglia + nres), r->st);
    r->r_fondo.upload(std::vector<double>(h_fondo, h_fondo + nres), r->st);
    r->r_elev.upload(std::vector<double>(elevation, elevation + ng), r->st);
    r->r_rid.upload(rid, r->st);
    r->r_rflag.upload(rflag, r->st);
    CG_HIP(hipStreamSynchronize(r->st));
    r->cap_members = 0;  // pool buffers follow NUMRES
    return 0;
  });
}

int cathy_gpu_route_set_celtype(int64_t *handle, const int *celtype) {
  return guarded([&] {
    CG_REQUIRE(celtype, "null argument");
    const std::shared_ptr<Route> r = route_lookup(handle);
    std::lock_guard<std::mutex> lk(r->mu);
    const int ng = r->ngrid;
    std::vector<int> alt(ng, -1);
    bool any = false;
    auto res_of = [&](int c) {  // NUM_R(c) - 1 (altezze.f:78, :89)
      CG_REQUIRE(r->numres > 0, "ALTEZZE types other than 0 need the reservoirs (set_reservoirs) first");
      const int v = r->reservr_h[c] - 1;
      if (v < 0 || v >= r->numres) throw Error(-1, "cell " + std::to_string(c + 1) + ": RESERVR outside 1..NUMRES");
      return v;
    };
    for (int c = 0; c < ng; ++c) {
      const int t = celtype[c];
      if (t == 0) continue;
      any = true;
      if (t < 1000) {
        alt[c] = 2 * res_of(c);
      } else {
        CG_REQUIRE(t - 1000 >= 1 && t - 1000 <= ng, "ALTEZZE type - 1000 outside the grid");
        alt[c] = 2 * res_of(t - 1000 - 1) + 1;
      }
    }
    if (any) {
      r->r_alt.upload(alt, r->st);
      CG_HIP(hipStreamSynchronize(r->st));
    }
    r->typed = any;
    return 0;
  });
}

int cathy_gpu_route_info(int64_t *handle, int *info) {
  return guarded([&] {
    CG_REQUIRE(info, "null argument");
    const std::shared_ptr<const Route> r = route_lookup(handle);
    info[0] = r->nlev;
    info[1] = r->maxpred;
    info[2] = r->ncell;
    info[3] = r->ngrid;
    return 0;
  });
}

int cathy_gpu_route_step(int64_t *handle, const int *nmember, const double *deltat, const double *delta_x,
                         const double *delta_y, const double *surface_water_sn, const double *q_in_kk_sn,
                         const double *q_out_kk_sn_1, const double *q_out_kk_sn_2, const double *volume_kk_sn,
                         double *q_in_kkp1_sn, double *q_out_kkp1_sn_1, double *q_out_kkp1_sn_2,
                         double *volume_kkp1_sn, double *h_water_kkp1_sn, double *cu_ak,
                         const double *h_pool_kk_vec, double *h_pool_kkp1_vec) {
  return guarded([&] {
    CG_REQUIRE(nmember && deltat && delta_x && delta_y && surface_water_sn && q_in_kk_sn && q_out_kk_sn_1 &&
                   q_out_kk_sn_2 && volume_kk_sn && q_in_kkp1_sn && q_out_kkp1_sn_1 && q_out_kkp1_sn_2 &&
                   volume_kkp1_sn && h_water_kkp1_sn && cu_ak,
               "null argument");
    const std::shared_ptr<Route> r = route_lookup(handle);
    std::lock_guard<std::mutex> lk(r->mu);
    CG_REQUIRE(!r->numres || (h_pool_kk_vec && h_pool_kkp1_vec), "reservoirs need H_POOL_KK_VEC / H_POOL_KKP1_VEC");
    const int m = *nmember;
    CG_REQUIRE(m >= 1 && m <= 65535, "NMEMBER must lie in 1..65535");
    const size_t len = (size_t)m * r->ngrid;
    const size_t lres = (size_t)m * r->numres;
    if (m > r->cap_members) {
      r->state.alloc(10 * (size_t)m * r->ngrid);
      r->cu_ak.alloc(2 * (size_t)m);
      r->err.alloc(m);
      r->hpool.alloc(2 * std::max<size_t>(lres, 1));
      r->cap_members = m;
    }
    double *S = r->state.p;
    const double *in[9] = {surface_water_sn, q_in_kk_sn,      q_out_kk_sn_1,   q_out_kk_sn_2, volume_kk_sn,
                           q_in_kkp1_sn,     q_out_kkp1_sn_1, q_out_kkp1_sn_2, volume_kkp1_sn};
    for (int a = 0; a < 9; ++a)
      CG_HIP(hipMemcpyAsync(S + a * len, in[a], len * sizeof(double), hipMemcpyHostToDevice, r->st));
    CG_HIP(hipMemcpyAsync(r->cu_ak.p, cu_ak, 2 * (size_t)m * sizeof(double), hipMemcpyHostToDevice, r->st));
    if (lres) {
      CG_HIP(hipMemcpyAsync(r->hpool.p, h_pool_kk_vec, lres * sizeof(double), hipMemcpyHostToDevice, r->st));
      CG_HIP(hipMemcpyAsync(r->hpool.p + lres, h_pool_kkp1_vec, lres * sizeof(double), hipMemcpyHostToDevice,
                            r->st));
    }
    hipLaunchKernelGGL(k_route, dim3(m), dim3(kRouteBlock), 0, r->st, r->dev(), *deltat, *delta_x, *delta_y,
                       S, S + len, S + 2 * len, S + 3 * len, S + 4 * len, S + 5 * len, S + 6 * len, S + 7 * len,
                       S + 8 * len, S + 9 * len, r->cu_ak.p, r->hpool.p, r->hpool.p + lres, r->err.p);
    CG_HIP(hipGetLastError());
    double *out[5] = {q_in_kkp1_sn, q_out_kkp1_sn_1, q_out_kkp1_sn_2, volume_kkp1_sn, h_water_kkp1_sn};
    for (int a = 0; a < 5; ++a)
      CG_HIP(hipMemcpyAsync(out[a], S + (5 + a) * len, len * sizeof(double), hipMemcpyDeviceToHost, r->st));
    CG_HIP(hipMemcpyAsync(cu_ak, r->cu_ak.p, 2 * (size_t)m * sizeof(double), hipMemcpyDeviceToHost, r->st));
    if (lres)
      CG_HIP(hipMemcpyAsync(h_pool_kkp1_vec, r->hpool.p + lres, lres * sizeof(double), hipMemcpyDeviceToHost,
                            r->st));
    std::vector<int> herr(m);
    CG_HIP(hipMemcpyAsync(herr.data(), r->err.p, m * sizeof(int), hipMemcpyDeviceToHost, r->st));
    CG_HIP(hipStreamSynchronize(r->st));
    for (int j = 0; j < m; ++j)
      CG_REQUIRE(!herr[j], "member " + std::to_string(j) +
                               ": a reservoir level left its height table (CALCOLI stops there)");
    return 0;
  });
}

int cathy_gpu_route_destroy(int64_t *handle) {
  return guarded([&] {
    std::shared_ptr<Route> r;
    {
      std::lock_guard<std::mutex> lk(g_route_mu);
      if (!handle) throw Error(CATHY_GPU_EHANDLE, "null route handle");
      auto it = g_routes.find(*handle);
      if (it == g_routes.end()) throw Error(CATHY_GPU_EHANDLE, "unknown route handle");
      r = std::move(it->second);
      g_routes.erase(it);
    }
    *handle = 0;
    return 0;
  });
}

}  // extern "C"
