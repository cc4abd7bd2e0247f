// newton_kernels.hip -- the Newton linearisation (IOPT=2) of the CATHY
// subsurface step on gfx950: derivative-aware constitutive update, the
// Jacobian assembly, ILU(0) in the solver's elimination order and the
// split-preconditioned BiCGSTAB of the reference's GCSTAS.
//
// Reference: newton.f (NEWUNS/CHNEW0, ASSNEW, RHSNEW, CFMATN, RHSGRV, BCNEW,
// NSYSLV ISOLV=0, SHLNEW).  The matrix is the full SELL-64 pattern of the
// Picard path; the ILU(0) factor lives in the same slots (FP32 in the
// column ordering, FP64 in the natural one): L in the lower slots with its
// diagonal, the unit upper factor U in the upper slots.
#include "flow_kernels.h"

namespace cg {

// chnew0.f (IVGHU=0): Sw, kr, ETAI = Sw Ss + phi dSw/dpsi and the
// derivatives DETAI, DCKRW at PTNEW.
__global__ void k_constitutive_newton(int n, const double *__restrict__ ptnew, SoilSoA s,
                                      double *__restrict__ sw, double *__restrict__ ckrw,
                                      double *__restrict__ etai, double *__restrict__ dckrw,
                                      double *__restrict__ detai) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double psi = ptnew[i];
  const double psat = s.psat[i], vgn = s.vgn[i], vgm = s.vgm[i], vgn1 = s.vgn1[i];
  const double pnot = s.pnot[i];
  const double se = dev_fvgse(psi, psat, vgn, vgm);
  const double dswdp = pnot * dev_fvgdse(psi, psat, vgn, vgn1, s.vgpsn[i], s.vgnr[i]);
  const double swv = pnot * se + s.rmcp[i];
  sw[i] = swv;
  etai[i] = swv * s.snodi[i] + s.pnodi[i] * dswdp;
  detai[i] = dswdp * s.snodi[i] + s.pnodi[i] * pnot * dev_fvgdds(psi, psat, vgn, vgn1, vgm);
  ckrw[i] = dev_fvgkr(psi, se, s.vgmr[i], vgm);
  dckrw[i] = dev_fvgdkr(psi, psat, vgn, vgn1, vgm);
}

// Jacobian of one row from the lumped assembly plan (ASSNEW + RHSNEW +
// CFMATN + RHSGRV + BCNEW's right-hand side):
//   A_rj = TETAF K_rj + delta_rj M_rr / dt + DCKRW_j T_rj + DETAI_j c_r V_rj
// with K = sum_e kr_e S_e, M_rr = sum_e V_e ETAE_e / 4 (LUMP=1),
// T_rj = sum_{e ∋ r,j} TETAF ((S_e psi)_k + Kz d_k IVOL) (k = r's vertex),
// c_r = TETAF (PNEW_r - PTIMEP_r) / (4 dt) and V_rj = sum_{e ∋ r,j} V_e;
// TNOTI = -K PTNEW - M (PNEW - PTIMEP) / dt - gravity.  The weights w2
// (= V/16 per node pair) give M_rr from nodal ETAI and V_rj; gw (= gz/4)
// gives the gravity term from nodal kr.
template <int MR>
__global__ __launch_bounds__(kBlock) void k_assemble_newton(AsmArgs a, AsmPlan P, NewtonArgs nw) {
  const int r = xcd_group() * kBlock + threadIdx.x;
  if (r >= a.n) return;
  const RowPos rp = row_pos(a.sl, r);
  const int s = r >> 6, lane = r & 63;
  const int base = a.sl.so[s] + lane;
  const int w = (a.sl.so[s + 1] - a.sl.so[s]) >> 6;
  const int len = sell_len(a.sl, rp, r), dg = sell_diag(a.sl, rp, r);
  const double tetaf = a.tetaf, rdt = a.rdt;
  double m2 = 0.0, grav = 0.0;
  for (int j = 0; j < w; ++j) {
    const int q = base + 64 * j;
    const int c = sell_col(a.sl, rp, q, j);
    m2 += P.w2[q] * nw.etai[c];
    grav += P.gw[q] * nw.ckrw[c];
  }
  double K[MR], T[MR];
#pragma unroll
  for (int j = 0; j < MR; ++j) {
    K[j] = 0.0;
    T[j] = 0.0;
  }
  const int p0 = P.off[s] + lane;
  const int cw = (P.off[s + 1] - P.off[s]) >> 6;
  for (int m = 0; m < cw; ++m) {
    const int q = p0 + 64 * m;
    const int e = P.ie[q];
    if (e < 0) continue;
    const uint32_t off = P.io[q];
    const double kr = a.ckrwe[e];
    const double4 sv = P.s4[q];
    const double sk[4] = {sv.x, sv.y, sv.z, sv.w};
    int o[4];
    double spsi = 0.0;
    int k = 0;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      o[l] = (off >> (8 * l)) & 0xff;
      if (o[l] == dg) k = l;
      spsi += sk[l] * a.ptnew[sell_col(a.sl, rp, base + 64 * o[l], o[l])];
    }
    const double tsumtd = tetaf * spsi + tetaf * nw.gz[(size_t)e * 4 + k];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const double v = kr * sk[l];
#pragma unroll
      for (int j = 0; j < MR; ++j) {
        K[j] += (j == o[l]) ? v : 0.0;
        T[j] += (j == o[l]) ? tsumtd : 0.0;
      }
    }
  }
  const double dp = a.pnew[r] - a.ptimep[r];
  const double cr = tetaf * (0.25 * dp) * rdt;
  double t = 0.0;
#pragma unroll
  for (int j = 0; j < MR; ++j) {
    if (j < len) {
      const int q = base + 64 * j;
      const int c = sell_col(a.sl, rp, q, j);
      t = t - K[j] * a.ptnew[c];
      double v = tetaf * K[j] + nw.dckrw[c] * T[j] + nw.detai[c] * (cr * (16.0 * P.w2[q]));
      if (j == dg) v = v + m2 * rdt;
      a.A[q] = v;
    }
  }
  t = t - (m2 * dp) * rdt;
  t = t - grav;
  a.xt5[r] = t;
  a.tnoti[r] = a.isdir[r] ? 0.0 : t + a.bcflux[r];
}
template __global__ void k_assemble_newton<16>(AsmArgs, AsmPlan, NewtonArgs);
template __global__ void k_assemble_newton<32>(AsmArgs, AsmPlan, NewtonArgs);

// sum over the common lower neighbours m of rows r and c (m < c) of
//   lu = L_rm U_mc   and   ul = L_cm U_mr
// for the SELL slot q = (r, c): pairs (a, b) = (slot (r,m), slot (c,m)),
// L_rm = F[a], U_mc = F[tpos[b]], L_cm = F[b], U_mr = F[tpos[a]].
template <typename FT>
__device__ __forceinline__ void ilu_pair_sums(const PairList &pl, const int *__restrict__ tpos,
                                              const FT *__restrict__ F, int q, double &lu, double &ul) {
  lu = 0.0;
  ul = 0.0;
  const int m1 = pl.pp[q + 1];
  for (int m = pl.pp[q]; m < m1; ++m) {
    const int2 pr = pl.pr[m];
    lu += (double)F[pr.x] * (double)F[tpos[pr.y]];
    ul += (double)F[pr.y] * (double)F[tpos[pr.x]];
  }
}

// Crout ILU(0), column ordering, one colour class (cf. k_ic_factor_col3).
// Row i, lower neighbours j ascending (earlier colours, row-parallel):
//   L_ij = A_ij - sum_m L_im U_mj          (slot (i,j))
//   U_ji = (A_ji - sum_m L_jm U_mi) / L_jj   (slot (j,i) = tpos of (i,j))
// then the chain: with p the row above (same column),
//   L_ip = q1,  U_pi = q2 / L_pp,  L_ii = s - L_ip U_pi
// where q1, q2, s are the earlier-colour parts.  Dirichlet rows become
// identity rows and their couplings vanish (dinv = 0).
template <typename FT>
__global__ __launch_bounds__(256) void k_ilu_factor_col3(ColDesc C, Sell sl, PairList pl,
                                                         const int *__restrict__ tpos,
                                                         const double *__restrict__ A,
                                                         const uint8_t *__restrict__ isdir, FT *__restrict__ F,
                                                         double *__restrict__ dinv, int *nfall) {
  extern __shared__ double smem[];
  const int tot = C.len * kChains;
  double *Sv = smem, *Q1 = smem + tot, *Q2 = smem + 2 * tot;
  int *Pd = (int *)(smem + 3 * tot);
  int *Pl = Pd + tot, *Pt = Pl + tot;
  const int t0 = xcd_group() * kChains;
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    int k, t;
    chain_elem(C, t0, i, k, t);
    if (t >= C.nc) continue;
    const int r = col_row(C, t, k);
    const RowPos rp = chain_pos(sl, C.g0 + t, k);
    const int jd = sell_diag(sl, rp, r);
    const int jp = k > 0 ? jd - 1 : jd;
    const int base = sl.so[r >> 6] + (r & 63);
    Pl[i] = k > 0 ? base + 64 * jp : -1;
    Pt[i] = k > 0 ? tpos[base + 64 * jp] : -1;
    if (isdir[r]) {
      for (int j = 0; j < jp; ++j) {
        const int pj = base + 64 * j;
        F[pj] = FT(0);
        F[tpos[pj]] = FT(0);
      }
      Pd[i] = -(base + 64 * jd) - 1;
      Sv[i] = 1.0;
      continue;
    }
    Pd[i] = base + 64 * jd;
    double sdiag = A[base + 64 * jd];
    for (int j = 0; j < jp; ++j) {
      const int pj = base + 64 * j;
      const int c = sell_col(sl, rp, pj, j);
      double L = 0.0, U = 0.0;
      if (!isdir[c]) {
        double lu, ul;
        ilu_pair_sums(pl, tpos, F, pj, lu, ul);
        L = A[pj] - lu;
        U = (A[tpos[pj]] - ul) * dinv[c];
      }
      F[pj] = FT(L);
      F[tpos[pj]] = FT(U);
      sdiag -= L * U;
    }
    Sv[i] = sdiag;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    int k, t;
    chain_elem(C, t0, i, k, t);
    if (t >= C.nc || k == 0) continue;
    if (Pd[i] < 0) continue;
    const int q = Pl[i];
    double lu, ul;
    ilu_pair_sums(pl, tpos, F, q, lu, ul);
    Q1[i] = A[q] - lu;
    Q2[i] = A[Pt[i]] - ul;
  }
  __syncthreads();
  if (threadIdx.x < kChains) {
    const int t = t0 + threadIdx.x;
    if (t < C.nc) {
      double dprev = 1.0;
      bool pdir = true;
      for (int k = 0; k < C.len; ++k) {
        const int i = chain_lds(C, threadIdx.x, k);
        const int r = col_row(C, t, k);
        const int pd = Pd[i];
        if (pd < 0) {
          if (k > 0) {
            F[Pl[i]] = FT(0);
            F[Pt[i]] = FT(0);
          }
          F[-pd - 1] = FT(1);
          dinv[r] = 0.0;
          pdir = true;
          continue;
        }
        double l = 0.0, u = 0.0;
        if (k > 0) {
          if (!pdir) {
            l = Q1[i];
            u = Q2[i] / dprev;
          }
          F[Pl[i]] = FT(l);
          F[Pt[i]] = FT(u);
        }
        double v = Sv[i] - l * u;
        if (!(fabs(v) > 0.0) || !isfinite(v)) {  // zero pivot: the previous row's (solscal-extended.f:2263)
          v = pdir ? 1.0 : dprev;
          atomicAdd(nfall, 1);
        }
        F[pd] = FT(v);
        dinv[r] = 1.0 / v;
        dprev = v;
        pdir = false;
      }
    }
  }
}

// Crout ILU(0), natural order, one level: each row of the level computes
// its lower slots in ascending order, then its diagonal.
template <typename FT>
__global__ void k_ilu_factor_list(ListLevel Lv, const int *__restrict__ rows, Sell sl, PairList pl,
                                  const int *__restrict__ tpos, const double *__restrict__ A,
                                  const uint8_t *__restrict__ isdir, FT *__restrict__ F,
                                  double *__restrict__ dinv, int *nfall) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Lv.cnt) return;
  const int r = rows[Lv.off + t];
  const RowPos rp = row_pos(sl, r);
  const int jd = sell_diag(sl, rp, r);
  const int base = sl.so[r >> 6] + (r & 63);
  if (isdir[r]) {
    for (int j = 0; j < jd; ++j) {
      const int pj = base + 64 * j;
      F[pj] = FT(0);
      F[tpos[pj]] = FT(0);
    }
    F[base + 64 * jd] = FT(1);
    dinv[r] = 0.0;
    return;
  }
  double sdiag = A[base + 64 * jd];
  for (int j = 0; j < jd; ++j) {
    const int pj = base + 64 * j;
    const int c = sell_col(sl, rp, pj, j);
    double L = 0.0, U = 0.0;
    if (!isdir[c]) {
      double lu, ul;
      ilu_pair_sums(pl, tpos, F, pj, lu, ul);
      L = A[pj] - lu;
      U = (A[tpos[pj]] - ul) * dinv[c];
    }
    F[pj] = FT(L);
    F[tpos[pj]] = FT(U);
    sdiag -= L * U;
  }
  if (!(fabs(sdiag) > 0.0) || !isfinite(sdiag)) {
    sdiag = 1.0;
    atomicAdd(nfall, 1);
  }
  F[base + 64 * jd] = FT(sdiag);
  dinv[r] = 1.0 / sdiag;
}

#define CG_INST_ILU(FT)                                                                                  \
  template __global__ void k_ilu_factor_col3<FT>(ColDesc, Sell, PairList, const int *, const double *,     \
                                                 const uint8_t *, FT *, double *, int *);                  \
  template __global__ void k_ilu_factor_list<FT>(ListLevel, const int *, Sell, PairList, const int *,      \
                                                 const double *, const uint8_t *, FT *, double *, int *);
CG_INST_ILU(float)
CG_INST_ILU(double)
#undef CG_INST_ILU

// y = U x (pru.f; U unit upper, from the factor's upper slots).
template <typename FT>
__global__ void k_mul_upper(int n, Sell sl, const FT *__restrict__ F, const double *__restrict__ x,
                            double *__restrict__ y) {
  const int r = xcd_group() * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const RowPos rp = row_pos(sl, r);
  const int jd = sell_diag(sl, rp, r), len = sell_len(sl, rp, r);
  const int base = sl.so[r >> 6] + (r & 63);
  double v = x[r];
  for (int j = jd + 1; j < len; ++j) {
    const int q = base + 64 * j;
    v += (double)F[q] * x[sell_col(sl, rp, q, j)];
  }
  y[r] = v;
}
template __global__ void k_mul_upper<float>(int, Sell, const float *, const double *, double *);
template __global__ void k_mul_upper<double>(int, Sell, const double *, const double *, double *);

// y = A x (aperbn.f), one thread per row.
__global__ void k_spmv(int n, Sell sl, const double *__restrict__ A, const double *__restrict__ x,
                       double *__restrict__ y, const int *__restrict__ done) {
  if (done && *done) return;
  const int r = xcd_group() * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const RowPos rp = row_pos(sl, r);
  const int s = r >> 6, lane = r & 63;
  const int o0 = sl.so[s], w = (sl.so[s + 1] - o0) >> 6;
  double acc = 0.0;
  int q = o0 + lane;
  for (int j = 0; j < w; ++j, q += 64) acc += A[q] * x[sell_col(sl, rp, q, j)];
  y[r] = acc;
}

// ------------------------------------------------------------- BiCGSTAB --
// gcstas.f on L^-1 A U^-1 (split ILU(0)), the reference's recurrences:
//   roj = (rt, r); beta = (roj/rojm1)(alfjm1/etaj); p = r + beta(p - etaj v)
//   v = L^-1 A U^-1 p; alfj = roj / (rt, v); s = r - alfj v
//   t = L^-1 A U^-1 s; etaj = (s,t)/(t,t)
//   y += alfj p + etaj s; r = s - etaj t (r = 0 on Dirichlet rows)
//   err = ||r|| / ||b_free||; stop when err <= TOL or NITER > ITMXCG.
// Scalars are reduced from per-group partials in a fixed order (every
// consuming group redundantly), as in the PCG.

__global__ void k_bicg_init(int nxl, const double *__restrict__ part_xl, int nrr, const double *__restrict__ part_rr,
                            BicgScal *sc, int *done, double tol, int imax) {
  __shared__ double sh[kBlock / 64 + 2];
  const double xl = reduce_partials<kBlock>(part_xl, nxl, sh);
  const double roj = reduce_partials<kBlock>(part_rr, nrr, sh);
  if (threadIdx.x == 0) {
    sc->xlung = xl;
    sc->roj = roj;
    sc->rojm1 = 1.0;
    sc->alfjm1 = 1.0;
    sc->etaj = 1.0;
    sc->tol = tol;
    sc->imax = imax;
    sc->niter = 1;
    sc->err = 0.0;
    *done = 0;
  }
}

// p = r + beta (p - etaj v)
__global__ void k_bicg_p(int n, const double *__restrict__ r, const double *__restrict__ v, double *__restrict__ p,
                         const BicgScal *sc, const int *__restrict__ done) {
  if (*done) return;
  const double beta = (sc->roj / sc->rojm1) * (sc->alfjm1 / sc->etaj);
  const double etaj = sc->etaj;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = r[i] + beta * (p[i] - etaj * v[i]);
}

// alfj = roj / (rt, v); s = r - alfj v
__global__ void k_bicg_s(int n, int nparts, const double *__restrict__ part_sig, const double *__restrict__ r,
                         const double *__restrict__ v, double *__restrict__ s, BicgScal *sc,
                         const int *__restrict__ done) {
  if (*done) return;
  __shared__ double sh[kBlock / 64 + 2];
  const double sigma = reduce_partials<kBlock>(part_sig, nparts, sh);
  const double alfj = sc->roj / sigma;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alfj = alfj;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) s[i] = r[i] - alfj * v[i];
}

// partial (s, t) and (t, t)
__global__ void k_bicg_st(int n, const double *__restrict__ s, const double *__restrict__ t,
                          double *__restrict__ part_st, double *__restrict__ part_tt, const int *__restrict__ done) {
  if (*done) return;
  __shared__ double sh[kBlock / 64 + 1];
  double a = 0.0, b = 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    a += s[i] * t[i];
    b += t[i] * t[i];
  }
  const double ta = block_sum<kBlock>(a, sh);
  const double tb = block_sum<kBlock>(b, sh);
  if (threadIdx.x == 0) {
    part_st[blockIdx.x] = ta;
    part_tt[blockIdx.x] = tb;
  }
}

// etaj = (s,t)/(t,t); y += alfj p + etaj s; r = s - etaj t (0 on Dirichlet
// rows); partials of (r, r) and of the next (rt, r).
__global__ void k_bicg_update(int n, int nparts, const double *__restrict__ part_st,
                              const double *__restrict__ part_tt, const double *__restrict__ p,
                              const double *__restrict__ s, const double *__restrict__ t,
                              const double *__restrict__ rt, const uint8_t *__restrict__ isdir,
                              double *__restrict__ y, double *__restrict__ r, double *__restrict__ part_rr,
                              double *__restrict__ part_rtr, BicgScal *sc, const int *__restrict__ done) {
  if (*done) return;
  __shared__ double sh[kBlock / 64 + 2];
  const double aa = reduce_partials<kBlock>(part_st, nparts, sh);
  const double bb = reduce_partials<kBlock>(part_tt, nparts, sh);
  const double etaj = aa / bb;
  const double alfj = sc->alfj;
  double a = 0.0, b = 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    y[i] = y[i] + alfj * p[i] + etaj * s[i];
    const double ri = isdir[i] ? 0.0 : s[i] - etaj * t[i];
    r[i] = ri;
    a += ri * ri;
    b += rt[i] * ri;
  }
  const double ta = block_sum<kBlock>(a, sh);
  const double tb = block_sum<kBlock>(b, sh);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = ta;
    part_rtr[blockIdx.x] = tb;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->etaj_next = etaj;
}

// err and the stopping test; roll the scalars for the next iteration.
__global__ void k_bicg_check(int n, int nparts, const double *__restrict__ part_rr,
                             const double *__restrict__ part_rtr, BicgScal *sc, int *done) {
  if (*done) return;
  __shared__ double sh[kBlock / 64 + 2];
  const double aa = reduce_partials<kBlock>(part_rr, nparts, sh);
  const double rtr = reduce_partials<kBlock>(part_rtr, nparts, sh);
  if (threadIdx.x == 0) {
    const double xl = sc->xlung;
    const double err = xl > 0.0 ? sqrt(aa / xl) : sqrt(aa / n);
    sc->err = err;
    sc->etaj = sc->etaj_next;
    if (sc->niter <= sc->imax && err > sc->tol) {
      sc->niter += 1;
      sc->rojm1 = sc->roj;
      sc->alfjm1 = sc->alfj;
      sc->roj = rtr;
    } else {
      *done = 1;
    }
  }
}

// Sw at a head vector (STORMB's SWNEW / SWTIMEP of the Newton branch).
__global__ void k_sw_only(int n, const double *__restrict__ psi, SoilSoA s, double *__restrict__ sw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sw[i] = s.pnot[i] * dev_fvgse(psi[i], s.psat[i], s.vgn[i], s.vgm[i]) + s.rmcp[i];
}

}  // namespace cg
