// flow_kernels.hip -- gfx950 kernels of one Picard iteration and of the
// IC(0)-PCG solve (device numbering, see kernels.h).
#include "flow_kernels.h"

namespace cg {

// ----------------------------------------------- constitutive (a2) --
// chpic0.f:23-36: Sw, kr, ET1 = Sw*Ss, ET2 = VGPNOT*dSe/dpsi at PTNEW, and
// SWNEW at PNEW.
__device__ __forceinline__ void constitutive_body(Blk bk, int n, const double *__restrict__ ptnew,
                               const double *__restrict__ pnew, SoilSoA s,
                               double *__restrict__ sw, double *__restrict__ ckrw,
                               double *__restrict__ et1, double *__restrict__ et2,
                               double *__restrict__ swnew) {
  int i = bk.b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double psat = s.psat[i], vgn = s.vgn[i], vgm = s.vgm[i];
  double pnot = s.pnot[i], rmcp = s.rmcp[i];
  swnew[i] = pnot * dev_fvgse(pnew[i], psat, vgn, vgm) + rmcp;
  double psi = ptnew[i];
  double se = dev_fvgse(psi, psat, vgn, vgm);
  double swv = pnot * se + rmcp;
  sw[i] = swv;
  et1[i] = swv * s.snodi[i];
  et2[i] = pnot * dev_fvgdse(psi, psat, vgn, s.vgn1[i], s.vgpsn[i], s.vgnr[i]);
  ckrw[i] = dev_fvgkr(psi, se, s.vgmr[i], vgm);
}
__global__ void k_constitutive(int n, const double *__restrict__ ptnew,
                               const double *__restrict__ pnew, SoilSoA s,
                               double *__restrict__ sw, double *__restrict__ ckrw,
                               double *__restrict__ et1, double *__restrict__ et2,
                               double *__restrict__ swnew) {
  constitutive_body(self_blk(), n, ptnew, pnew, s, sw, ckrw, et1, et2, swnew);
}

// Sw at one head vector (SWTIMEP of chpic0.f:24-27, Sw of chvelo.f) and,
// optionally, kr (chvelo.f).
__device__ __forceinline__ void sw_kr_body(Blk bk, int n, const double *__restrict__ psi_in, SoilSoA s,
                        double *__restrict__ sw, double *__restrict__ ckrw) {
  int i = bk.b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double psi = psi_in[i];
  double se = dev_fvgse(psi, s.psat[i], s.vgn[i], s.vgm[i]);
  sw[i] = s.pnot[i] * se + s.rmcp[i];
  if (ckrw) ckrw[i] = dev_fvgkr(psi, se, s.vgmr[i], s.vgm[i]);
}
__global__ void k_sw_kr(int n, const double *__restrict__ psi_in, SoilSoA s,
                        double *__restrict__ sw, double *__restrict__ ckrw) {
  sw_kr_body(self_blk(), n, psi_in, s, sw, ckrw);
}

// nodelt.f: element means of kr and ET1.
__global__ void k_elem_avg(int nt, const int4 *__restrict__ enode,
                           const double *__restrict__ ckrw, const double *__restrict__ et1,
                           double *__restrict__ ckrwe, double *__restrict__ et1e) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  int4 v = enode[e];
  double a = 0.0;
  a = a + ckrw[v.x]; a = a + ckrw[v.y]; a = a + ckrw[v.z]; a = a + ckrw[v.w];
  ckrwe[e] = a * 0.25;
  double b = 0.0;
  b = b + et1[v.x]; b = b + et1[v.y]; b = b + et1[v.z]; b = b + et1[v.w];
  et1e[e] = b * 0.25;
}

// bcpic.f:34-87 classification of each row: Dirichlet (fixed-head nodes,
// seepage-face exit nodes and atmospheric nodes with IFATM = 1 or 2) or not,
// and the flux it receives: the Neumann Q of its node (free drainage and
// listed nodes, bcpic.f:71-74) plus, at IFATM = 0 surface nodes, TETAF*ATMACT
// + (1-TETAF)*ATMOLD (bcpic.f:75-80).  bcfix (null: none) holds per row bit
// 2 = fixed head (the CONTP list), bit 4 = seepage face with SFEX = 1.
// isdir = 1 marks the Dirichlet rows (NODDIR of bcpic.f), which GRADDP
// and RESSYM leave out of their norms (solscal-extended.f:1287-1297,
// 4400-4437: SYMSLV passes them NUMDIR / NODDIR).
__device__ __forceinline__ void bc_rows_body(Blk bk, int n, const int *__restrict__ row_surf,
                          const double *__restrict__ qnode, const int *__restrict__ ifatm,
                          const double *__restrict__ atmact, const double *__restrict__ atmold,
                          double tetaf, const uint8_t *__restrict__ bcfix, uint8_t *__restrict__ isdir,
                          double *__restrict__ bcflux) {
  int r = bk.b * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint8_t d = 0;
  if (bcfix && (bcfix[r] & 6)) d = 1;
  double f = qnode[r];
  int s = row_surf[r];
  if (s >= 0) {
    int fa = ifatm[s];
    if (fa == 1 || fa == 2) d = 1;
    else if (fa == 0) f = f + (tetaf * atmact[s] + (1.0 - tetaf) * atmold[s]);
  }
  isdir[r] = d;
  bcflux[r] = f;
}
__global__ void k_bc_rows(int n, const int *__restrict__ row_surf,
                          const double *__restrict__ qnode, const int *__restrict__ ifatm,
                          const double *__restrict__ atmact, const double *__restrict__ atmold,
                          double tetaf, const uint8_t *__restrict__ bcfix, uint8_t *__restrict__ isdir,
                          double *__restrict__ bcflux) {
  bc_rows_body(self_blk(), n, row_surf, qnode, ifatm, atmact, atmold, tetaf, bcfix, isdir, bcflux);
}

// BCPIC's right-hand side of row r from the assembled value t: Dirichlet
// rows restart from zero (bcpic.f:36,48,62), every row then takes its
// Neumann/atmospheric flux (bcpic.f:71-80) and loses its root uptake
// QTRANIE (bcpic.f:84-86, null: none).
__device__ __forceinline__ double bc_rhs(const AsmArgs &a, int r, double t) {
  double v = (a.isdir[r] ? 0.0 : t) + a.bcflux[r];
  if (a.qtr) v = v - a.qtr[r];
  return v;
}

// Upper-pair slot of (k,l) in an element's 10 stiffness values.
__device__ __forceinline__ int upair(int k, int l) {
  int a = k < l ? k : l, b = k < l ? l : k;
  return a * 4 - (a * (a - 1)) / 2 + (b - a);
}

// Row-gather assembly + RHS + CFMATP + RHSGRV + BCPIC right-hand side.
// asspic.f:26-51 scattered element matrices; here each thread owns one row
// and gathers the contributions of its incident elements (ascending
// element id, the reference's accumulation order) into registers (MR =
// compile-time row width, entries selected by slot), so the assembly is
// deterministic and atomic-free.
//   K      = sum_e kr_e * S_e                       (COEF1 before CFMATP)
//   TNOTI  = -K PTNEW - M2/dt (PNEW-PTIMEP) - M4/dt (SWNEW-SWTIMEP) - grav
//   A      = TETAF*K + (M2 + diag(M4_rr ET2_r))/dt  (cfmatp.f)
// Dirichlet rows keep their true coefficients (BKPIC needs them) and get
// TNOTI = 0; the preconditioner masks them.
template <int MR, bool LUMPED>
__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs a) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= a.n) return;
  const RowPos rp = row_pos(a.sl, r);
  const int len = sell_len(a.sl, rp, r);
  double K[MR];
  double M[LUMPED ? 1 : MR];
#pragma unroll
  for (int j = 0; j < MR; ++j) {
    K[j] = 0.0;
    if (!LUMPED) M[j] = 0.0;
  }
  double m2d = 0.0, grav = 0.0;
  const int i1 = a.inc_ptr[r + 1];
  for (int i = a.inc_ptr[r]; i < i1; ++i) {
    const int ek = a.inc[i];
    const int e = ek >> 2, k = ek & 3;
    const uint32_t off = a.incoff[i];
    const double kr = a.ckrwe[e];
    const double *se = a.S + (size_t)e * 10;
    const double ve1 = a.volu[e] * a.et1e[e];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int o = (off >> (8 * l)) & 0xff;
      const double v = kr * se[upair(k, l)];
      const double w = LUMPED ? 0.0 : ve1 * a.lmass[k * 4 + l];
#pragma unroll
      for (int j = 0; j < MR; ++j) {
        K[j] += (j == o) ? v : 0.0;
        if (!LUMPED) M[j] += (j == o) ? w : 0.0;
      }
    }
    if (LUMPED) m2d = m2d + ve1 * 0.25;
    grav = grav + kr * a.gz[(size_t)e * 4 + k];
  }
  const int dg = sell_diag(a.sl, rp, r);
  const int base = a.sl.so[r >> 6] + (r & 63);
  const double rdt = a.rdt;
  double t = 0.0;
#pragma unroll
  for (int j = 0; j < MR; ++j)
    if (j < len) t = t - K[j] * a.ptnew[sell_col(a.sl, rp, base + 64 * j, j)];
  double m4r;
  if (LUMPED) {
    m4r = a.m4d[r];
    t = t - (m2d * rdt) * (a.pnew[r] - a.ptimep[r]) - (m4r * rdt) * (a.swnew[r] - a.swtimep[r]);
  } else {
    m4r = a.m4full[base + 64 * dg];
#pragma unroll
    for (int j = 0; j < MR; ++j) {
      if (j < len) {
        const int q = base + 64 * j;
        const int c = sell_col(a.sl, rp, q, j);
        t = t - (M[j] * rdt) * (a.pnew[c] - a.ptimep[c]) - (a.m4full[q] * rdt) * (a.swnew[c] - a.swtimep[c]);
      }
    }
  }
  t = t - grav;
  a.xt5[r] = t;
  a.tnoti[r] = bc_rhs(a, r, t);
  const double tetaf = a.tetaf;
  const double m5 = m4r * a.et2[r];
#pragma unroll
  for (int j = 0; j < MR; ++j) {
    if (j < len) {
      double v = tetaf * K[j];
      if (LUMPED) {
        if (j == dg) v = v + m2d * rdt + m5 * rdt;
      } else {
        v = v + M[j] * rdt;
        if (j == dg) v = v + m5 * rdt;
      }
      a.A[base + 64 * j] = v;
    }
  }
}
template __global__ void k_assemble<16, true>(AsmArgs);
template __global__ void k_assemble<16, false>(AsmArgs);
template __global__ void k_assemble<32, true>(AsmArgs);
template __global__ void k_assemble<32, false>(AsmArgs);

// Lumped assembly from the precomputed plan (same quantities as
// k_assemble): plan records, A's slots and the weights are read coalesced;
// the gathers are kr_e (once per incident element) and nodal values of
// the row's neighbours.  Entries accumulate in registers by slot.
template <int MR>
__device__ __forceinline__ void assemble_plan_body(Blk bk, AsmArgs a, AsmPlan P, const double *__restrict__ ckrw_node,
                                                          const double *__restrict__ et1_node) {
  const int r = bk.b * kBlock + threadIdx.x;
  if (r >= a.n) return;
  const RowPos rp = row_pos(a.sl, r);
  const int s = r >> 6, lane = r & 63;
  const int base = a.sl.so[s] + lane;
  const int w = (a.sl.so[s + 1] - a.sl.so[s]) >> 6;
  // Loads are issued in straight-line batches (slots and records past the
  // end re-read the last one and are masked), so each batch is in flight
  // together; a load under a condition makes the compiler wait for every
  // earlier load at the next merge.  Sums keep the slot / record order.
  // nodal kr at the row's slots, for the element means of the records:
  // cks[j][thread] (element kr = NODELT over its 4 nodes, all of which are
  // slots of this row)
  __shared__ double cks[kUnroll * kBlock];
  const bool local_kr = w <= kUnroll;
  double m2d = 0.0, grav = 0.0;
  if (w <= kUnroll) {
    int c[kUnroll];
    double w2v[kUnroll], gwv[kUnroll], e1[kUnroll], ck[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {  // slots past the slice width: one broadcast each
      const int q = j < w ? base + 64 * j : 0;
      w2v[j] = P.w2[q];
      gwv[j] = P.gw[q];
    }
    sell_cols_window(a.sl, rp, base, 0, w - 1, c);
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      e1[j] = et1_node[j < w ? c[j] : 0];
      ck[j] = ckrw_node[j < w ? c[j] : 0];
    }
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      m2d = fma(j < w ? w2v[j] : 0.0, e1[j], m2d);
      grav = fma(j < w ? gwv[j] : 0.0, ck[j], grav);
      cks[j * kBlock + threadIdx.x] = ck[j];
    }
  } else {
    for (int j = 0; j < w; ++j) {
      const int q = base + 64 * j;
      const int c = sell_col(a.sl, rp, q, j);
      m2d += P.w2[q] * et1_node[c];
      grav += P.gw[q] * ckrw_node[c];
    }
  }
  // element contributions accumulate by slot in the thread's own LDS
  // column (slot MR absorbs masked records): one read-modify-write each,
  // in record order per slot, where a register array indexed by the
  // record's slot would take a compare and select per slot
  __shared__ double kacc[(MR + 1) * kBlock];
#pragma unroll
  for (int j = 0; j <= MR; ++j) kacc[j * kBlock + threadIdx.x] = 0.0;
  const int p0 = P.off[s] + lane;
  const int cw = (P.off[s + 1] - P.off[s]) >> 6;
  constexpr int RB = 8;  // element records per batch
  for (int m0 = 0; m0 < cw; m0 += RB) {
    uint32_t off[RB];
    double4 sv[RB];
    double kr[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {  // records past the slice's count: one broadcast each
      const int q = m0 + b < cw ? p0 + 64 * (m0 + b) : 0;
      off[b] = P.io[q];
      sv[b] = P.s4[q];
    }
    // the element ids are read only where kr_e comes from the element
    // means; a padding record's slot word marks it
    if (local_kr) {
      // nodelt.f order: (((k1 + k2) + k3) + k4) / 4, nodes in element order
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        const uint32_t o = off[b];
        const double s1 = cks[(o & 15u) * kBlock + threadIdx.x];
        const double s2 = cks[((o >> 8) & 15u) * kBlock + threadIdx.x];
        const double s3 = cks[((o >> 16) & 15u) * kBlock + threadIdx.x];
        const double s4 = cks[((o >> 24) & 15u) * kBlock + threadIdx.x];
        kr[b] = (((s1 + s2) + s3) + s4) * 0.25;
      }
    } else {
      int e[RB];
#pragma unroll
      for (int b = 0; b < RB; ++b) e[b] = P.ie[p0 + 64 * min(m0 + b, cw - 1)];
#pragma unroll
      for (int b = 0; b < RB; ++b) kr[b] = a.ckrwe[max(e[b], 0)];
    }
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const bool live = m0 + b < cw && off[b] != 0xFFFFFFFFu;
      const double v[4] = {live ? kr[b] * sv[b].x : 0.0, live ? kr[b] * sv[b].y : 0.0,
                           live ? kr[b] * sv[b].z : 0.0, live ? kr[b] * sv[b].w : 0.0};
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int o = live ? (int)((off[b] >> (8 * l)) & 0xff) : MR;
        kacc[o * kBlock + threadIdx.x] += v[l];
      }
    }
  }
  double K[MR];
#pragma unroll
  for (int j = 0; j < MR; ++j) K[j] = kacc[j * kBlock + threadIdx.x];
  const int len = sell_len(a.sl, rp, r), dg = sell_diag(a.sl, rp, r);
  const double rdt = a.rdt, tetaf = a.tetaf;
  const double m4r = a.m4d[r];
  const double m5 = m4r * a.et2[r];
  double t = 0.0;
  {
    int c[MR];
    double pv[MR];
    sell_cols_window(a.sl, rp, base, 0, len - 1, c);
#pragma unroll
    for (int j = 0; j < MR; ++j) pv[j] = a.ptnew[c[j]];
#pragma unroll
    for (int j = 0; j < MR; ++j) t = fma(-(j < len ? K[j] : 0.0), pv[j], t);
  }
#pragma unroll
  for (int j = 0; j < MR; ++j) {
    if (j < len) {
      const int q = base + 64 * j;
      a.A[q] = tetaf * K[j] + (j == dg ? m2d * rdt + m5 * rdt : 0.0);
    }
  }
  t = t - (m2d * rdt) * (a.pnew[r] - a.ptimep[r]) - (m4r * rdt) * (a.swnew[r] - a.swtimep[r]);
  t = t - grav;
  a.xt5[r] = t;
  a.tnoti[r] = bc_rhs(a, r, t);
}
template <int MR>
__global__ __launch_bounds__(kBlock) void k_assemble_plan(AsmArgs a, AsmPlan P, const double *__restrict__ ckrw_node,
                                                          const double *__restrict__ et1_node) {
  assemble_plan_body<MR>(self_xcd_blk(), a, P, ckrw_node, et1_node);
}
template __global__ void k_assemble_plan<16>(AsmArgs, AsmPlan, const double *, const double *);
template __global__ void k_assemble_plan<32>(AsmArgs, AsmPlan, const double *, const double *);

// ------------------------------------------------------------ IC(0) --
// incldp.f restated as an up-looking IC(0) on the full pattern: for each
// row r in elimination order, L_rj = (a_rj - sum_{m<j} L_rm L_jm) / L_jj for
// its lower entries j (ascending), L_rr = sqrt(a_rr - sum L_rj^2).  The
// terms of sum_{m<j} come from the static pair list.  Dirichlet rows and
// columns are removed (identity, zero inverse diagonal) -- what the
// 1e-9*RMAX diagonal penalty of bcpic.f does in exact arithmetic.  The
// factor's transpose is mirrored into the upper slots so both sweeps read
// rows.
template <typename FT>
__device__ __forceinline__ double pair_dot(const PairList &pl, const FT *__restrict__ F, int q) {
  double v = 0.0;
  const int m1 = pl.pp[q + 1];
  for (int m = pl.pp[q]; m < m1; ++m) {
    const int2 pr = pl.pr[m];
    v += (double)F[pr.x] * (double)F[pr.y];
  }
  return v;
}

// Pass 2 of the column factor (k_ic_factor_col3, k_ic_factor_col3t): the
// chain recurrence d_k = sqrt(s_k - (q_k / d_(k-1))^2) in its squared form
// e_k = s_k - q_k^2 / e_(k-1), one division per level, by one thread per
// chain (e_k overwrites Sv); then, for all rows at once, d_k = sqrt(e_k),
// 1/d_k and the predecessor entry l_k = q_k / d_(k-1).  A Dirichlet row is
// an identity row and restarts its chain (its successor's l is 0); a
// non-positive e_k (breakdown) is replaced by |a_kk| (or 1), as incldp.f
// keeps the row's own diagonal.
template <typename FT>
__device__ __forceinline__ void factor_chains(const ColDesc &C, int t0, double *Sv, const double *Qv,
                                              const int *Pd, const int *Pl, const int *Pt,
                                              const double *__restrict__ A, FT *__restrict__ F,
                                              double *__restrict__ dinv, int *nfall) {
  const int tot = C.len * kChains;
  const int ls = chain_step(C);
  if (threadIdx.x < kChains && t0 + threadIdx.x < C.nc) {
    double e = 1.0;
    bool pdir = true;
    for (int k = 0; k < C.len; ++k) {
      const int i = chain_lds(C, threadIdx.x, k);
      const int pd = Pd[i];
      if (pd < 0) {
        pdir = true;
        continue;
      }
      const double qq = (k > 0 && !pdir) ? Qv[i] * Qv[i] : 0.0;
      double v = Sv[i] - qq / e;
      if (!(v > 0.0)) {
        const double a = fabs(A[pd]);
        v = a > 0.0 ? a : 1.0;
        atomicAdd(nfall, 1);
      }
      Sv[i] = v;
      e = v;
      pdir = false;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    int k, t;
    chain_elem(C, t0, i, k, t);
    if (t >= C.nc) continue;
    const int r = col_row(C, t, k);
    const int pd = Pd[i];
    if (pd < 0) {
      if (k > 0) {
        F[Pl[i]] = FT(0);
        F[Pt[i]] = FT(0);
      }
      F[-pd - 1] = FT(1);
      dinv[r] = 0.0;
      continue;
    }
    const double d = sqrt(Sv[i]);
    F[pd] = FT(d);
    dinv[r] = 1.0 / d;
    if (k > 0) {
      const double l = Pd[i - ls] >= 0 ? Qv[i] / sqrt(Sv[i - ls]) : 0.0;
      F[Pl[i]] = FT(l);
      F[Pt[i]] = FT(l);
    }
  }
}

// Column ordering, one colour class.  Each work-group owns kChains chains.
// In this ordering a row's lower neighbours are (a) nodes of earlier
// colours, final before the launch, and (b) its same-column predecessor
// (plane k-1), whose entries common with the row are all of earlier
// colours.  So: pass 1a computes every row's earlier-colour L entries in
// parallel, pass 1b the predecessor's numerator, and pass 2
// (factor_chains) the chain recurrence from LDS.
template <typename FT>
__global__ __launch_bounds__(256) void k_ic_factor_col3(ColDesc C, Sell sl, PairList pl,
                                                        const int *__restrict__ tpos,
                                                        const double *__restrict__ A,
                                                        const uint8_t *__restrict__ isdir, FT *__restrict__ F,
                                                        double *__restrict__ dinv, int *nfall) {
  extern __shared__ double smem[];
  const int tot = C.len * kChains;
  double *Sv = smem;                 // a_rr - sum earlier L^2
  double *Qv = smem + tot;           // a_r,pred - earlier dot
  int *Pd = (int *)(smem + 2 * tot); // slot of the diagonal (-1: Dirichlet)
  int *Pl = Pd + tot;                // slot of the predecessor entry
  int *Pt = Pl + tot;                // its transpose slot
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
    double s = A[base + 64 * jd];
    for (int j = 0; j < jp; ++j) {
      const int pj = base + 64 * j;
      const int c = sell_col(sl, rp, pj, j);
      double L = 0.0;
      if (!isdir[c]) L = (A[pj] - pair_dot(pl, F, pj)) * dinv[c];
      F[pj] = FT(L);
      F[tpos[pj]] = FT(L);
      s -= L * L;
    }
    Sv[i] = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    int k, t;
    chain_elem(C, t0, i, k, t);
    if (t >= C.nc || k == 0) continue;
    if (Pd[i] < 0) continue;
    const int q = Pl[i];
    Qv[i] = A[q] - pair_dot(pl, F, q);
  }
  __syncthreads();
  factor_chains(C, t0, Sv, Qv, Pd, Pl, Pt, A, F, dinv, nfall);
}

// Byte i of a little-endian word array (i a compile-time constant after
// unrolling).
__device__ __forceinline__ int tab_byte(const uint32_t *w, int i) { return (int)((w[i >> 2] >> (8 * (i & 3))) & 0xFFu); }

// k_ic_factor_col3's three passes, with pass 1 (earlier-colour entries)
// and pass 1b (predecessor numerator) from the chain tables.  Each lower
// entry L_rj = (a_rj - sum_a L_ra L_cb) / L_cc sums its pairs in ascending
// a, as the pair lists do, with the row's own L_ra (rounded to FT, as
// stored) held in registers.
template <typename FT, int NW>
__device__ __forceinline__ void ic_factor_col3t_body(Blk bk, double *smem, ColDesc C, Sell sl, FacTab ft,
                                                         const double *__restrict__ A,
                                                         const uint8_t *__restrict__ isdir, FT *__restrict__ F,
                                                         double *__restrict__ dinv, int *nfall) {
  constexpr int NWc = NW > 0 ? NW : 1;
  constexpr int NP = NW * (NW - 1) / 2;  // (j, a) pairs, a < j < NW
  constexpr int NPW = NP > 0 ? (NP + 3) / 4 : 1;
    const int tot = C.len * kChains;
  double *Sv = smem;                 // a_rr - sum earlier L^2
  double *Qv = smem + tot;           // a_r,pred - earlier dot
  int *Pd = (int *)(smem + 2 * tot); // slot of the diagonal (-(slot)-1: Dirichlet)
  int *Pl = Pd + tot;                // slot of the predecessor entry
  int *Pt = Pl + tot;                // its transpose slot
  const int t0 = bk.b * kChains;
  const int i = threadIdx.x;
  int k, t;
  chain_elem(C, t0, i, k, t);
  const bool live = i < tot && t < C.nc;
  double Lr[NWc];
  uint32_t wpre[4];
  int r = 0, base = 0, jd = 0, jp = 0;
  if (live) {
    r = col_row(C, t, k);
    const int g = C.g0 + t;
    const RowPos rp = chain_pos(sl, g, k);
    jd = sell_diag(sl, rp, r);
    jp = k > 0 ? jd - 1 : jd;
    base = sl.so[r >> 6] + (r & 63);
    const uint32_t *tw = ft.w + (size_t)fac_class(k, C.len) * kFacWords * ft.nchain + g;
    uint32_t wp[NPW], wt[4];
#pragma unroll
    for (int w = 0; w < NPW; ++w) wp[w] = tw[(size_t)w * ft.nchain];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      wpre[w] = tw[(size_t)(30 + w) * ft.nchain];
      wt[w] = tw[(size_t)(34 + w) * ft.nchain];
    }
    const bool rdir = isdir[r] != 0;
    const int jl = max(jp - 1, 0);
    int c[NWc];
    double av[NWc], dc[NWc];
    int so_c[NWc];
    uint8_t ic[NWc];
    FT G[NP > 0 ? NP : 1];
    if (NW > 0) {
      sell_cols_window(sl, rp, base, 0, jl, c);
#pragma unroll
      for (int j = 0; j < NW; ++j) {  // slots past jp: index 0 (broadcasts), unused
        const int cj = j < jp ? c[j] : 0;
        av[j] = A[j < jp ? base + 64 * j : 0];
        dc[j] = dinv[cj];
        ic[j] = isdir[cj];
        so_c[j] = sl.so[cj >> 6] + (cj & 63);
      }
#pragma unroll
      for (int j = 1; j < NW; ++j)
#pragma unroll
        for (int a = 0; a < j; ++a) {
          // absent pairs read F[0]: a wave whose rows all lack the pair
          // issues one broadcast load (the value is multiplied by 0)
          const int q = j * (j - 1) / 2 + a, b = tab_byte(wp, q);
          G[q] = F[b == 0xFF ? 0 : so_c[j] + 64 * b];
        }
    }
    if (rdir) {
#pragma unroll
      for (int j = 0; j < NW; ++j)
        if (j < jp) {
          F[base + 64 * j] = FT(0);
          F[so_c[j] + 64 * tab_byte(wt, j)] = FT(0);
        }
#pragma unroll
      for (int j = 0; j < NWc; ++j) Lr[j] = 0.0;
      Pd[i] = -(base + 64 * jd) - 1;
      Sv[i] = 1.0;
    } else {
      double s = A[base + 64 * jd];
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < j; ++a) {
          const int q = j * (j - 1) / 2 + a;
          v = fma(tab_byte(wp, q) != 0xFF ? Lr[a] : 0.0, (double)G[q], v);
        }
        const double L = (j < jp && !ic[j]) ? (av[j] - v) * dc[j] : 0.0;
        const FT Lf = FT(L);
        Lr[j] = (double)Lf;
        if (j < jp) {
          F[base + 64 * j] = Lf;
          F[so_c[j] + 64 * tab_byte(wt, j)] = Lf;
        }
        s -= L * L;
      }
      Sv[i] = s;
      Pd[i] = base + 64 * jd;
    }
    if (k > 0) {
      const int pred = r - col_w(C, t);
      Pl[i] = base + 64 * (jd - 1);
      Pt[i] = sl.so[pred >> 6] + (pred & 63) + 64 * tab_byte(wt, jd - 1);
    } else {
      Pl[i] = -1;
      Pt[i] = -1;
    }
  }
  __syncthreads();
  // the predecessor's numerator: its own earlier-colour entries are final
  // after the barrier
  if (live && k > 0 && Pd[i] >= 0) {
    const int pred = r - col_w(C, t);
    const int sp0 = sl.so[pred >> 6] + (pred & 63);
    FT G2[NWc];
#pragma unroll
    for (int a = 0; a < NW; ++a) {
      const int b = tab_byte(wpre, a);
      G2[a] = F[b == 0xFF ? 0 : sp0 + 64 * b];
    }
    const int q = Pl[i];
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < NW; ++a) v = fma((a < jp && tab_byte(wpre, a) != 0xFF) ? Lr[a] : 0.0, (double)G2[a], v);
    Qv[i] = A[q] - v;
  }
  __syncthreads();
  factor_chains(C, t0, Sv, Qv, Pd, Pl, Pt, A, F, dinv, nfall);
}
template <typename FT, int NW>
__global__ __launch_bounds__(256) void k_ic_factor_col3t(ColDesc C, Sell sl, FacTab ft,
                                                         const double *__restrict__ A,
                                                         const uint8_t *__restrict__ isdir, FT *__restrict__ F,
                                                         double *__restrict__ dinv, int *nfall) {
  extern __shared__ double smem[];
  ic_factor_col3t_body<FT, NW>(self_xcd_blk(), smem, C, sl, ft, A, isdir, F, dinv, nfall);
}

// The chain recurrence v_k = b_k - a_k v_(k-1) (DOWN: k ascending) or
// v_k = b_k - a_k v_(k+1) (up), v = 0 outside the chain, for all chains of
// a work-group at once: element i of the LDS arrays b, a is (chain, plane
// k), one element per thread (len*kChains <= blockDim.x; ON false for
// threads past them), planes LS elements apart.  A Hillis-Steele
// scan of the affine maps: after the round of stride s, (b_k, a_k) is the
// composition of the maps k .. k-2s+1, so after ceil(log2 len) rounds
// b_k = v_k.  Composition: b_k - a_k (b_j - a_j v) = (b_k - a_k b_j) -
// (-a_k a_j) v.
template <bool DOWN>
__device__ __forceinline__ void chain_scan(double *b, double *a, int len, bool on, int k, int ls) {
  const int i = threadIdx.x;
  for (int s = 1; s < len; s <<= 1) {
    const bool act = on && (DOWN ? k >= s : k + s < len);
    double nb = 0.0, na = 0.0;
    if (act) {
      const int j = DOWN ? i - s * ls : i + s * ls;
      const double ai = a[i];
      nb = fma(-ai, b[j], b[i]);
      na = -ai * a[j];
    }
    __syncthreads();
    if (act) {
      b[i] = nb;
      a[i] = na;
    }
    __syncthreads();
  }
}

// Forward sweep L y = t (proddp, solscal-extended.f:4175-4185) on one
// colour: pass 1 gathers the earlier-colour terms of every row and scales
// by 1/L_kk, pass 2 runs y_k = t'_k - l'_k y_k-1 down each chain in LDS.
// LAST (the final colour, whose rows couple to no later colour): pass 2
// then runs the backward sweep L^T x = y of the same chains from LDS
// (x_k = y_k/L_kk - u'_k x_k+1, as k_bwd_col3 forms it) and stores x.
template <typename FT, int NW, bool LAST>
__device__ __forceinline__ void fwd_col3_body(Blk bk, double *smem, ColDesc C, Sell sl, const FT *__restrict__ F,
                                                  const double *__restrict__ dinv, const double *__restrict__ tin,
                                                  double *__restrict__ y, const int *__restrict__ done) {
  if (done && *done) return;
    const int tot = C.len * kChains;
  double *T = smem, *Lp = smem + tot;
  double *U = smem + 2 * tot, *D = smem + 3 * tot;  // LAST only
  const int t0 = bk.b * kChains;
  const int kl = C.len - 1;
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    int k, t;
    chain_elem(C, t0, i, k, t);
    if (t >= C.nc) continue;
    const int r = col_row(C, t, k);
    const RowPos rp = chain_pos(sl, C.g0 + t, k);
    const int jd = sell_diag(sl, rp, r);
    const int jp = k > 0 ? jd - 1 : jd;
    const int base = sl.so[r >> 6] + (r & 63);
    double s = tin[r];
    if (NW == 0 && jp == 0) {
      // no earlier-colour terms in this colour
    } else if (jp <= NW) {
      // straight-line window over slots [0, jp): slots past it re-read
      // slot max(jp - 1, 0) and are masked (see sell_row_dot)
      const int jl = max(jp - 1, 0);
      int c[NW > 0 ? NW : 1];
      float f[NW > 0 ? NW : 1];
      double yv[NW > 0 ? NW : 1];
      // slots past the window read index 0 (a broadcast where the whole
      // wave is past it) and are masked
#pragma unroll
      for (int j = 0; j < NW; ++j) f[j] = (float)F[j < jp ? base + 64 * j : 0];
      sell_cols_window(sl, rp, base, 0, jl, c);
#pragma unroll
      for (int j = 0; j < NW; ++j) yv[j] = y[j < jp ? c[j] : 0];
#pragma unroll
      for (int j = 0; j < NW; ++j) s = fma(-(j < jp ? (double)f[j] : 0.0), yv[j], s);
    } else {
      for (int j = 0; j < jp; ++j) {
        const int pj = base + 64 * j;
        s -= (double)F[pj] * y[sell_col(sl, rp, pj, j)];
      }
    }
    const double di = dinv[r];
    T[i] = s * di;
    Lp[i] = k > 0 ? (double)F[base + 64 * jp] * di : 0.0;
    if (LAST) {
      U[i] = k < kl ? (double)F[base + 64 * (jd + 1)] * di : 0.0;
      D[i] = di;
    }
  }
  __syncthreads();
  if (tot <= (int)blockDim.x) {
    // chains of at most blockDim.x / kChains levels: log-depth scans
    const int i = threadIdx.x;
    int k, t;
    chain_elem(C, t0, i, k, t);
    const bool on = i < tot;
    const bool store = on && t < C.nc;
    chain_scan<true>(T, Lp, C.len, on, k, chain_step(C));
    if (LAST) {
      if (on) T[i] *= D[i];
      __syncthreads();
      chain_scan<false>(T, U, C.len, on, k, chain_step(C));
    }
    if (store) y[col_row(C, t, k)] = T[i];
    return;
  }
  if (threadIdx.x < kChains) {
    const int t = t0 + threadIdx.x;
    if (t < C.nc) {
      double yp = 0.0;
      for (int k = 0; k < C.len; ++k) {
        const int i = chain_lds(C, threadIdx.x, k);
        const double v = T[i] - Lp[i] * yp;
        if (LAST) T[i] = v * D[i];
        else y[col_row(C, t, k)] = v;
        yp = v;
      }
      if (LAST) {
        double xn = 0.0;
        for (int k = kl; k >= 0; --k) {
          const int i = chain_lds(C, threadIdx.x, k);
          const double v = T[i] - U[i] * xn;
          y[col_row(C, t, k)] = v;
          xn = v;
        }
      }
    }
  }
}
template <typename FT, int NW, bool LAST>
__global__ __launch_bounds__(256) void k_fwd_col3(ColDesc C, Sell sl, const FT *__restrict__ F,
                                                  const double *__restrict__ dinv, const double *__restrict__ tin,
                                                  double *__restrict__ y, const int *__restrict__ done) {
  extern __shared__ double smem[];
  fwd_col3_body<FT, NW, LAST>(self_xcd_blk(), smem, C, sl, F, dinv, tin, y, done);
}

// Backward sweep L^T x = y (proddp, solscal-extended.f:4186-4197) on one
// colour, in place: later colours are final; chains run from the base up.
template <typename FT, int NW>
__device__ __forceinline__ void bwd_col3_body(Blk bk, double *smem, ColDesc C, Sell sl, const FT *__restrict__ F,
                                                  const double *__restrict__ dinv, const double *xin, double *x,
                                                  const int *__restrict__ done) {
  if (done && *done) return;
    const int tot = C.len * kChains;
  double *T = smem, *U = smem + tot;
  const int t0 = bk.b * kChains;
  const int kl = C.len - 1;
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    int k, t;
    chain_elem(C, t0, i, k, t);
    if (t >= C.nc) continue;
    const int r = col_row(C, t, k);
    const RowPos rp = chain_pos(sl, C.g0 + t, k);
    const int jd = sell_diag(sl, rp, r);
    const int js = k < kl ? jd + 1 : jd;
    const int len = sell_len(sl, rp, r);
    const int base = sl.so[r >> 6] + (r & 63);
    double s = xin[r];
    const int j0 = js + 1;
    if (NW == 0 && len <= j0) {
      // no later-colour terms in this colour
    } else if (len - j0 <= NW) {
      // straight-line window over slots (js, len), in slot order: window
      // entry j is slot min(js + 1 + j, len - 1), masked past len
      const int jl = len - 1;
      int c[NW > 0 ? NW : 1];
      float f[NW > 0 ? NW : 1];
      double xv[NW > 0 ? NW : 1];
      // slots past the row read index 0 (a broadcast where the whole wave
      // is past it) and are masked
#pragma unroll
      for (int j = 0; j < NW; ++j) f[j] = (float)F[j0 + j < len ? base + 64 * (j0 + j) : 0];
      sell_cols_window(sl, rp, base, j0, jl, c);
#pragma unroll
      for (int j = 0; j < NW; ++j) xv[j] = x[j0 + j < len ? c[j] : 0];
#pragma unroll
      for (int j = 0; j < NW; ++j) s = fma(-(j0 + j < len ? (double)f[j] : 0.0), xv[j], s);
    } else {
      for (int j = js + 1; j < len; ++j) {
        const int pj = base + 64 * j;
        s -= (double)F[pj] * x[sell_col(sl, rp, pj, j)];
      }
    }
    const double di = dinv[r];
    T[i] = s * di;
    U[i] = k < kl ? (double)F[base + 64 * (jd + 1)] * di : 0.0;
  }
  __syncthreads();
  if (tot <= (int)blockDim.x) {
    const int i = threadIdx.x;
    int k, t;
    chain_elem(C, t0, i, k, t);
    const bool on = i < tot;
    chain_scan<false>(T, U, C.len, on, k, chain_step(C));
    if (on && t < C.nc) x[col_row(C, t, k)] = T[i];
    return;
  }
  if (threadIdx.x < kChains) {
    const int t = t0 + threadIdx.x;
    if (t < C.nc) {
      double xn = 0.0;
      for (int k = kl; k >= 0; --k) {
        const int i = chain_lds(C, threadIdx.x, k);
        const double v = T[i] - U[i] * xn;
        x[col_row(C, t, k)] = v;
        xn = v;
      }
    }
  }
}
template <typename FT, int NW>
__global__ __launch_bounds__(256) void k_bwd_col3(ColDesc C, Sell sl, const FT *__restrict__ F,
                                                  const double *__restrict__ dinv, const double *xin, double *x,
                                                  const int *__restrict__ done) {
  extern __shared__ double smem[];
  bwd_col3_body<FT, NW>(self_xcd_blk(), smem, C, sl, F, dinv, xin, x, done);
}

// Natural order: one row per thread, levels of the reference's
// elimination DAG launched in sequence (exact INCLDP/PRODDP order).
template <typename FT>
__global__ void k_ic_factor_list(ListLevel L, const int *__restrict__ rows, Sell sl, PairList pl,
                                 const int *__restrict__ tpos, const double *__restrict__ A,
                                 const uint8_t *__restrict__ isdir, FT *__restrict__ F,
                                 double *__restrict__ dinv, int *nfall) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.cnt) return;
  const int r = rows[L.off + t];
  const int jd = sl.diag[r];
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
  double s = A[base + 64 * jd];
  for (int j = 0; j < jd; ++j) {
    const int pj = base + 64 * j;
    const int c = sl.col[pj];
    double Lv = 0.0;
    if (!isdir[c]) Lv = (A[pj] - pair_dot(pl, F, pj)) * dinv[c];
    F[pj] = FT(Lv);
    F[tpos[pj]] = FT(Lv);
    s -= Lv * Lv;
  }
  if (!(s > 0.0)) {
    const double a = fabs(A[base + 64 * jd]);
    s = a > 0.0 ? a : 1.0;
    atomicAdd(nfall, 1);
  }
  const double d = sqrt(s);
  F[base + 64 * jd] = FT(d);
  dinv[r] = 1.0 / d;
}

template <typename FT>
__global__ void k_fwd_list(ListLevel L, const int *__restrict__ rows, Sell sl, const FT *__restrict__ F,
                           const double *__restrict__ dinv, const double *__restrict__ tin,
                           double *__restrict__ y, const int *__restrict__ done) {
  if (done && *done) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.cnt) return;
  const int r = rows[L.off + t];
  const int jd = sl.diag[r];
  const int base = sl.so[r >> 6] + (r & 63);
  double s = tin[r];
  for (int j = 0; j < jd; ++j) {
    const int pj = base + 64 * j;
    s -= (double)F[pj] * y[sl.col[pj]];
  }
  y[r] = s * dinv[r];
}

template <typename FT>
__global__ void k_bwd_list(ListLevel L, const int *__restrict__ rows, Sell sl, const FT *__restrict__ F,
                           const double *__restrict__ dinv, const double *xin, double *x,
                           const int *__restrict__ done) {
  if (done && *done) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.cnt) return;
  const int r = rows[L.off + t];
  const int jd = sl.diag[r], len = sl.len[r];
  const int base = sl.so[r >> 6] + (r & 63);
  double s = xin[r];
  for (int j = jd + 1; j < len; ++j) {
    const int pj = base + 64 * j;
    s -= (double)F[pj] * x[sl.col[pj]];
  }
  x[r] = s * dinv[r];
}

#define CG_INST_SWEEP(FT, U)                                                                              \
  template __global__ void k_fwd_col3<FT, U, false>(ColDesc, Sell, const FT *, const double *,              \
                                                    const double *, double *, const int *);                 \
  template __global__ void k_fwd_col3<FT, U, true>(ColDesc, Sell, const FT *, const double *,               \
                                                   const double *, double *, const int *);                  \
  template __global__ void k_bwd_col3<FT, U>(ColDesc, Sell, const FT *, const double *, const double *,      \
                                             double *, const int *);
#define CG_INST_FTAB(FT, W)                                                                               \
  template __global__ void k_ic_factor_col3t<FT, W>(ColDesc, Sell, FacTab, const double *, const uint8_t *, \
                                                    FT *, double *, int *);
#define CG_INST_FT(FT)                                                                                      \
  template __global__ void k_ic_factor_col3<FT>(ColDesc, Sell, PairList, const int *, const double *,      \
                                                const uint8_t *, FT *, double *, int *);                    \
  CG_INST_FTAB(FT, 0)                                                                                       \
  CG_INST_FTAB(FT, 4)                                                                                       \
  CG_INST_FTAB(FT, 8)                                                                                       \
  CG_INST_FTAB(FT, 12)                                                                                      \
  CG_INST_SWEEP(FT, 0)                                                                                      \
  CG_INST_SWEEP(FT, 4)                                                                                      \
  CG_INST_SWEEP(FT, 8)                                                                                      \
  CG_INST_SWEEP(FT, 12)                                                                                     \
  CG_INST_SWEEP(FT, 16)                                                                                     \
  template __global__ void k_ic_factor_list<FT>(ListLevel, const int *, Sell, PairList, const int *,        \
                                                const double *, const uint8_t *, FT *, double *, int *);    \
  template __global__ void k_fwd_list<FT>(ListLevel, const int *, Sell, const FT *, const double *,         \
                                          const double *, double *, const int *);                          \
  template __global__ void k_bwd_list<FT>(ListLevel, const int *, Sell, const FT *, const double *,          \
                                          const double *, double *, const int *);
CG_INST_FT(float)
CG_INST_FT(double)
#undef CG_INST_FT
#undef CG_INST_SWEEP

// --------------------------------------------------------------- PCG --
// graddp.f:1275-1385 with the reference's recurrences:
//   alpha = (p,r)/(p,Ap); r -= alpha Ap; x += alpha p; z = M^-1 r;
//   beta = -(Ap,z)/(p,Ap); p = z + beta p; err = ||r_free|| / ||b_free||.
// Dirichlet rows carry p = z = 0, so they drop out of every product.
// One thread per row; a wavefront covers one SELL slice, so the loop count
// (the slice width) is uniform and every load of A / col is coalesced.
// Scalars are summed from per-work-group partials in a fixed order by
// every group of the consuming kernel (deterministic, no atomics).

__device__ __forceinline__ double sell_row_dot(const Sell &sl, const double *__restrict__ A,
                                               const double *__restrict__ x, int r) {
  const int s = r >> 6, lane = r & 63;
  const int o0 = sl.so[s], w = (sl.so[s + 1] - o0) >> 6;
  const RowPos rp = row_pos(sl, r);
  const int q0 = o0 + lane;
  double acc = 0.0;
  if (w <= kUnroll) {
    // Straight-line loads: every slot of the unrolled window is loaded
    // (slots past the row's end re-read its last slot) and masked when
    // summed, so values, column codes and then the gathers are each in
    // flight together -- a load under a condition makes the compiler wait
    // for every earlier load at the next merge.
    int c[kUnroll];
    double a[kUnroll], xv[kUnroll];
    // slots past the slice width (uniform over the wave) read index 0: one
    // broadcast each
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) a[j] = A[j < w ? q0 + 64 * j : 0];
    if (sl.coded) {
      uint32_t cw[kUnroll];
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) cw[j] = code_word(sl, rp, min(j, w - 1));
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) c[j] = code_col(sl, rp.k, cw[j]);
    } else {
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) c[j] = sl.col[q0 + 64 * min(j, w - 1)];
    }
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) xv[j] = x[j < w ? c[j] : 0];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) acc = fma(j < w ? a[j] : 0.0, xv[j], acc);
  } else {
    int q = q0;
    for (int j = 0; j < w; ++j, q += 64) acc += A[q] * x[sell_col(sl, rp, q, j)];
  }
  return acc;
}

// sell_row_dot reading the strict lower triangle through the transpose:
// slot j < diag takes a_jr from row col(r, j) at the slot the chain table
// names (byte 136 + j); the other slots read the row's own values.  The
// products and their order are sell_row_dot's (coded rows of at most
// kUnroll slots; the caller checks).
__device__ __forceinline__ double sell_row_dot_sym(const Sell &sl, const FacTab &ft, const double *__restrict__ A,
                                                   const double *__restrict__ x, int r) {
  const int s = r >> 6, lane = r & 63;
  const int o0 = sl.so[s], w = (sl.so[s + 1] - o0) >> 6;
  const RowPos rp = row_pos(sl, r);
  const int q0 = o0 + lane;
  const int jd = sell_diag(sl, rp, r);
  const uint32_t *tw = ft.w + (size_t)fac_class(rp.k, sl.nplane) * kFacWords * ft.nchain + rp.g;
  uint32_t wt[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) wt[u] = tw[(size_t)(34 + u) * ft.nchain];
  uint32_t cw[kUnroll];
#pragma unroll
  for (int j = 0; j < kUnroll; ++j) cw[j] = code_word(sl, rp, min(j, w - 1));
  int c[kUnroll], sc[kUnroll];
#pragma unroll
  for (int j = 0; j < kUnroll; ++j) c[j] = code_col(sl, rp.k, cw[j]);
#pragma unroll
  for (int j = 0; j < kUnroll; ++j) sc[j] = sl.so[c[j] >> 6] + (c[j] & 63);
  double a[kUnroll], xv[kUnroll];
#pragma unroll
  for (int j = 0; j < kUnroll; ++j) xv[j] = x[j < w ? c[j] : 0];
#pragma unroll
  for (int j = 0; j < kUnroll; ++j) a[j] = A[j < jd ? sc[j] + 64 * tab_byte(wt, j) : (j < w ? q0 + 64 * j : 0)];
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kUnroll; ++j) acc = fma(j < w ? a[j] : 0.0, xv[j], acc);
  return acc;
}

// Partial sum of b_free^2 (XLUNG) and r = b - A x0.
__device__ __forceinline__ void residual0_body(Blk bk, int n, Sell sl, const double *__restrict__ A, const double *__restrict__ x,
                            const double *__restrict__ b, const uint8_t *__restrict__ isdir,
                            double *__restrict__ r, double *__restrict__ part) {
  __shared__ double sh[kBlock / 64 + 1];
  double acc = 0.0;
  for (int i = bk.b * blockDim.x + threadIdx.x; i < n; i += bk.nb * blockDim.x) {
    const double bi = b[i];
    r[i] = bi - sell_row_dot(sl, A, x, i);
    if (!isdir[i]) acc += bi * bi;
  }
  double t = block_sum<kBlock>(acc, sh);
  if (threadIdx.x == 0) part[bk.b] = t;
}
__global__ void k_residual0(int n, Sell sl, const double *__restrict__ A, const double *__restrict__ x,
                            const double *__restrict__ b, const uint8_t *__restrict__ isdir,
                            double *__restrict__ r, double *__restrict__ part) {
  residual0_body(self_blk(), n, sl, A, x, b, isdir, r, part);
}

// B = A p with partial (p,r) and (p,B); one thread per row.
__device__ __forceinline__ void spmv_dot2_body(Blk bk, int n, Sell sl, FacTab ft, const double *__restrict__ A, const double *__restrict__ p,
                            const double *__restrict__ r, double *__restrict__ B,
                            double *__restrict__ part_pr, double *__restrict__ part_pb,
                            const int *__restrict__ done) {
  if (*done) return;
  __shared__ double sh[2 * (kBlock / 64)];
  double apr = 0.0, apb = 0.0;
  const int i = bk.b * blockDim.x + threadIdx.x;
  if (i < n) {
    const double s = ft.w ? sell_row_dot_sym(sl, ft, A, p, i) : sell_row_dot(sl, A, p, i);
    B[i] = s;
    const double pi = p[i];
    apr = pi * r[i];
    apb = pi * s;
  }
  const double2 t = block_sum2<kBlock>(apr, apb, sh);
  if (threadIdx.x == 0) {
    part_pr[bk.b] = t.x;
    part_pb[bk.b] = t.y;
  }
}
__global__ void k_spmv_dot2(int n, Sell sl, FacTab ft, const double *__restrict__ A, const double *__restrict__ p,
                            const double *__restrict__ r, double *__restrict__ B,
                            double *__restrict__ part_pr, double *__restrict__ part_pb,
                            const int *__restrict__ done) {
  spmv_dot2_body(self_xcd_blk(), n, sl, ft, A, p, r, B, part_pr, part_pb, done);
}

// alpha from the SpMV partials (every group sums them in the same fixed
// order); r -= alpha B; x += alpha p; partial r_free^2.
__device__ __forceinline__ void pcg_alpha_body(Blk bk, int n, int nparts, const double *__restrict__ part_pr,
                            const double *__restrict__ part_pb, const double *__restrict__ B,
                            const double *__restrict__ p, const uint8_t *__restrict__ isdir,
                            double *__restrict__ r, double *__restrict__ x, double *__restrict__ part_rr,
                            PcgScal *sc, const int *__restrict__ done) {
  if (*done) return;
  __shared__ double sh[2 * (kBlock / 64)];
  const double2 prb = reduce_partials2<kBlock>(part_pr, part_pb, nparts, sh);
  const double pr = prb.x, pb = prb.y;
  const double alfa = pr / pb;
  if (bk.b == 0 && threadIdx.x == 0) {
    sc->pb = pb;
    sc->alpha = alfa;
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int stride = bk.nb * blockDim.x;
  int i = bk.b * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {  // four rows in flight per thread
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int iu = i + u * stride;
      const double ri = r[iu] - alfa * B[iu];
      r[iu] = ri;
      x[iu] = x[iu] + alfa * p[iu];
      if (!isdir[iu]) acc[u] += ri * ri;
    }
  }
  for (; i < n; i += stride) {
    const double ri = r[i] - alfa * B[i];
    r[i] = ri;
    x[i] = x[i] + alfa * p[i];
    if (!isdir[i]) acc[0] += ri * ri;
  }
  double t = block_sum<kBlock>((acc[0] + acc[1]) + (acc[2] + acc[3]), sh);
  if (threadIdx.x == 0) part_rr[bk.b] = t;
}
__global__ void k_pcg_alpha(int n, int nparts, const double *__restrict__ part_pr,
                            const double *__restrict__ part_pb, const double *__restrict__ B,
                            const double *__restrict__ p, const uint8_t *__restrict__ isdir,
                            double *__restrict__ r, double *__restrict__ x, double *__restrict__ part_rr,
                            PcgScal *sc, const int *__restrict__ done) {
  pcg_alpha_body(self_blk(), n, nparts, part_pr, part_pb, B, p, isdir, r, x, part_rr, sc, done);
}

// partial (B, z)
__device__ __forceinline__ void dot_body(Blk bk, int n, const double *__restrict__ u, const double *__restrict__ v,
                      double *__restrict__ part, const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double sh[kBlock / 64 + 1];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int stride = bk.nb * blockDim.x;
  int i = bk.b * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
#pragma unroll
    for (int w = 0; w < 4; ++w) acc[w] += u[i + w * stride] * v[i + w * stride];
  }
  for (; i < n; i += stride) acc[0] += u[i] * v[i];
  double t = block_sum<kBlock>((acc[0] + acc[1]) + (acc[2] + acc[3]), sh);
  if (threadIdx.x == 0) part[bk.b] = t;
}
__global__ void k_dot(int n, const double *__restrict__ u, const double *__restrict__ v,
                      double *__restrict__ part, const int *__restrict__ done) {
  dot_body(self_blk(), n, u, v, part, done);
}

// beta, p = z + beta p, err and the stopping test (graddp.f:1348-1377).
__device__ __forceinline__ void pcg_beta_body(Blk bk, int n, int nparts, const double *__restrict__ part_bz,
                           const double *__restrict__ part_rr, const double *__restrict__ z,
                           double *__restrict__ p, PcgScal *sc, int *done) {
  if (*done) return;
  __shared__ double sh[2 * (kBlock / 64)];
  const double2 crr = reduce_partials2<kBlock>(part_bz, part_rr, nparts, sh);
  const double c = crr.x, rr = crr.y;
  const double beta = -c / sc->pb;
  const int stride = bk.nb * blockDim.x;
  int i = bk.b * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
#pragma unroll
    for (int u = 0; u < 4; ++u) p[i + u * stride] = z[i + u * stride] + beta * p[i + u * stride];
  }
  for (; i < n; i += stride) p[i] = z[i] + beta * p[i];
  if (bk.b == 0 && threadIdx.x == 0) {
    const double xl = sc->xlung;
    const double err = xl > 0.0 ? sqrt(rr / xl) : sqrt(rr / n);
    sc->err = err;
    if (err > sc->tol && sc->niter < sc->imax) {
      sc->niter += 1;
    } else {
      *done = 1;
    }
  }
}
__global__ void k_pcg_beta(int n, int nparts, const double *__restrict__ part_bz,
                           const double *__restrict__ part_rr, const double *__restrict__ z,
                           double *__restrict__ p, PcgScal *sc, int *done) {
  pcg_beta_body(self_blk(), n, nparts, part_bz, part_rr, z, p, sc, done);
}

// XLUNG from k_residual0's partials; reset the iteration state.
__device__ __forceinline__ void pcg_init_body(Blk bk, int nparts, const double *__restrict__ part, PcgScal *sc, int *done,
                           double tol, int imax) {
  __shared__ double sh[kBlock / 64 + 2];
  const double xl = reduce_partials<kBlock>(part, nparts, sh);
  if (threadIdx.x == 0) {
    sc->xlung = xl;
    sc->tol = tol;
    sc->imax = imax;
    sc->niter = 1;
    sc->err = 0.0;
    *done = 0;
  }
}
__global__ void k_pcg_init(int nparts, const double *__restrict__ part, PcgScal *sc, int *done,
                           double tol, int imax) {
  pcg_init_body(self_blk(), nparts, part, sc, done, tol, imax);
}

// ------------------------------------------------ Picard bookkeeping --
// picard.f:185-200: PNEW += PDIFF, then SHLPIC (atmospheric Dirichlet
// nodes keep POLD).
__device__ __forceinline__ void update_pnew_body(Blk bk, int n, const double *__restrict__ pdiff,
                              const int *__restrict__ row_surf, const int *__restrict__ ifatm,
                              const double *__restrict__ pold, double *__restrict__ pnew) {
  int r = bk.b * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double v = pnew[r] + pdiff[r];
  const int s = row_surf[r];
  if (s >= 0) {
    const int fa = ifatm[s];
    if (fa == 1 || fa == 2) v = pold[r];
  }
  pnew[r] = v;
}
__global__ void k_update_pnew(int n, const double *__restrict__ pdiff,
                              const int *__restrict__ row_surf, const int *__restrict__ ifatm,
                              const double *__restrict__ pold, double *__restrict__ pnew) {
  update_pnew_body(self_blk(), n, pdiff, row_surf, ifatm, pold, pnew);
}

// bkpic.f:28-54 at atmospheric Dirichlet nodes: ATMACT = (A dpsi - XT5 -
// (1-TETAF) ATMOLD) / TETAF with the true (restored) diagonal.
__device__ __forceinline__ void bkpic_body(Blk bk, int nnod, const int *__restrict__ surf_row, Sell sl, const double *__restrict__ A,
                        const double *__restrict__ pdiff, const double *__restrict__ xt5,
                        const int *__restrict__ ifatm, const double *__restrict__ atmold, double tetaf,
                        double *__restrict__ atmact) {
  int s = bk.b * blockDim.x + threadIdx.x;
  if (s >= nnod) return;
  const int fa = ifatm[s];
  if (!(fa == 1 || fa == 2)) return;
  const int r = surf_row[s];
  const RowPos rp = row_pos(sl, r);
  double v = 0.0;
  const int len = sell_len(sl, rp, r);
  for (int j = 0; j < len; ++j) {
    const int q = sp(sl.so, r, j);
    v += A[q] * pdiff[sell_col(sl, rp, q, j)];
  }
  v = v - xt5[r];
  atmact[s] = (v - (1.0 - tetaf) * atmold[s]) * (1.0 / tetaf);
}
__global__ void k_bkpic(int nnod, const int *__restrict__ surf_row, Sell sl, const double *__restrict__ A,
                        const double *__restrict__ pdiff, const double *__restrict__ xt5,
                        const int *__restrict__ ifatm, const double *__restrict__ atmold, double tetaf,
                        double *__restrict__ atmact) {
  bkpic_body(self_blk(), nnod, surf_row, sl, A, pdiff, xt5, ifatm, atmold, tetaf, atmact);
}

// norms.f (PINF with its last-index tie-break on CATHY ids, PL2, FINF,
// FL2) and stormb.f's DSTORE, as per-block partials.
__device__ __forceinline__ void norms_body(Blk bk, int n, const int *__restrict__ refid, const double *__restrict__ pnew,
                        const double *__restrict__ pold, const double *__restrict__ tnoti,
                        const double *__restrict__ swnew, const double *__restrict__ swtimep,
                        const double *__restrict__ ptimep, const double *__restrict__ volnod,
                        const double *__restrict__ snodi, const double *__restrict__ pnodi,
                        NormPart part) {
  __shared__ double sh[kBlock / 64 + 1];
  __shared__ double shv[kBlock / 64], shd[kBlock / 64];
  __shared__ int shi[kBlock / 64];
  double pl2 = 0.0, fl2 = 0.0, dst = 0.0, finf = 0.0, pinf = -1.0, pdif = 0.0;
  int pidx = -1;
  for (int i = bk.b * blockDim.x + threadIdx.x; i < n; i += bk.nb * blockDim.x) {
    const double diff = pnew[i] - pold[i];
    const double a = fabs(diff);
    pl2 += diff * diff;
    const double t = tnoti[i];
    fl2 += t * t;
    finf = fmax(finf, fabs(t));
    const int id = refid[i];
    if (a > pinf || (a == pinf && id > pidx)) {
      pinf = a;
      pidx = id;
      pdif = diff;
    }
    dst += volnod[i] * (snodi[i] * (swnew[i] + swtimep[i]) * 0.5 * (pnew[i] - ptimep[i]) +
                        pnodi[i] * (swnew[i] - swtimep[i]));
  }
  double t1 = block_sum<kBlock>(pl2, sh);
  double t2 = block_sum<kBlock>(fl2, sh);
  double t3 = block_sum<kBlock>(dst, sh);
  // max reductions (finf; pinf with index)
  for (int o = 32; o > 0; o >>= 1) {
    finf = fmax(finf, __shfl_down(finf, o, 64));
    const double ov = __shfl_down(pinf, o, 64);
    const int oi = __shfl_down(pidx, o, 64);
    const double od = __shfl_down(pdif, o, 64);
    if (ov > pinf || (ov == pinf && oi > pidx)) {
      pinf = ov;
      pidx = oi;
      pdif = od;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    shv[w] = pinf;
    shi[w] = pidx;
    shd[w] = pdif;
    sh[w] = finf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bv = -1.0, bd = 0.0, bf = 0.0;
    int bi = -1;
    for (int k = 0; k < kBlock / 64; ++k) {
      if (shv[k] > bv || (shv[k] == bv && shi[k] > bi)) {
        bv = shv[k];
        bi = shi[k];
        bd = shd[k];
      }
      bf = fmax(bf, sh[k]);
    }
    part.pl2[bk.b] = t1;
    part.fl2[bk.b] = t2;
    part.dstore[bk.b] = t3;
    part.finf[bk.b] = bf;
    part.pinf[bk.b] = bv;
    part.pidx[bk.b] = bi;
    part.pdif[bk.b] = bd;
  }
}
__global__ void k_norms(int n, const int *__restrict__ refid, const double *__restrict__ pnew,
                        const double *__restrict__ pold, const double *__restrict__ tnoti,
                        const double *__restrict__ swnew, const double *__restrict__ swtimep,
                        const double *__restrict__ ptimep, const double *__restrict__ volnod,
                        const double *__restrict__ snodi, const double *__restrict__ pnodi,
                        NormPart part) {
  norms_body(self_blk(), n, refid, pnew, pold, tnoti, swnew, swtimep, ptimep, volnod, snodi, pnodi, part);
}

// fluxmb.f: atmospheric and Neumann in/out sums on the surface / base.
__device__ __forceinline__ void flux_body(Blk bk, int nnod, const int *__restrict__ ifatm, const double *__restrict__ atmact,
                       const double *__restrict__ qbot, int nq, double *__restrict__ part6) {
  __shared__ double sh[kBlock / 64 + 1];
  double v[6] = {0, 0, 0, 0, 0, 0};  // adin adout anin anout nnin nnout
  for (int s = bk.b * blockDim.x + threadIdx.x; s < nnod; s += bk.nb * blockDim.x) {
    const int fa = ifatm[s];
    const double a = atmact[s];
    if (fa == 1 || fa == 2) {
      if (a > 0.0) v[0] += a; else v[1] += a;
    } else if (fa == 0) {
      if (a > 0.0) v[2] += a; else v[3] += a;
    }
    if (s < nq) {
      const double q = qbot[s];
      if (q > 0.0) v[4] += q; else v[5] += q;
    }
  }
  for (int k = 0; k < 6; ++k) {
    double t = block_sum<kBlock>(v[k], sh);
    if (threadIdx.x == 0) part6[k * kMaxRedBlocks + bk.b] = t;
  }
}
__global__ void k_flux(int nnod, const int *__restrict__ ifatm, const double *__restrict__ atmact,
                       const double *__restrict__ qbot, int nq, double *__restrict__ part6) {
  flux_body(self_blk(), nnod, ifatm, atmact, qbot, nq, part6);
}

// One work-group folds the partials of k_norms and k_flux into res[16]:
// 0 DSTORE 1 PL2 2 FL2 3 FINF 4 PINF 5 PIKMAX 6 IKMAX 7 ADIN 8 ADOUT
// 9 ANIN 10 ANOUT 11 NNIN 12 NNOUT
__device__ __forceinline__ void norms_final_body(Blk bk, int nb, int nbf, NormPart part, const double *__restrict__ part6,
                              double *__restrict__ res) {
  __shared__ double sh[kBlock / 64 + 2];
  __shared__ double shv[kBlock / 64], shd[kBlock / 64], shf[kBlock / 64];
  __shared__ int shi[kBlock / 64];
  const double pl2 = reduce_partials<kBlock>(part.pl2, nb, sh);
  const double fl2 = reduce_partials<kBlock>(part.fl2, nb, sh);
  const double dst = reduce_partials<kBlock>(part.dstore, nb, sh);
  double fx[6];
  for (int k = 0; k < 6; ++k) fx[k] = reduce_partials<kBlock>(part6 + k * kMaxRedBlocks, nbf, sh);
  double bv = -1.0, bd = 0.0, bf = 0.0;
  int bi = -1;
  for (int k = threadIdx.x; k < nb; k += kBlock) {
    const double v = part.pinf[k];
    const int id = part.pidx[k];
    if (v > bv || (v == bv && id > bi)) {
      bv = v;
      bi = id;
      bd = part.pdif[k];
    }
    bf = fmax(bf, part.finf[k]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    bf = fmax(bf, __shfl_down(bf, o, 64));
    const double ov = __shfl_down(bv, o, 64);
    const int oi = __shfl_down(bi, o, 64);
    const double od = __shfl_down(bd, o, 64);
    if (ov > bv || (ov == bv && oi > bi)) {
      bv = ov;
      bi = oi;
      bd = od;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    shv[w] = bv;
    shi[w] = bi;
    shd[w] = bd;
    shf[w] = bf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    bv = -1.0; bd = 0.0; bf = 0.0; bi = -1;
    for (int k = 0; k < kBlock / 64; ++k) {
      if (shv[k] > bv || (shv[k] == bv && shi[k] > bi)) {
        bv = shv[k];
        bi = shi[k];
        bd = shd[k];
      }
      bf = fmax(bf, shf[k]);
    }
    res[0] = dst;
    res[1] = sqrt(pl2);
    res[2] = sqrt(fl2);
    res[3] = bf;
    res[4] = bv;
    res[5] = bd;
    res[6] = (double)bi;
    for (int k = 0; k < 6; ++k) res[7 + k] = fx[k];
  }
}
__global__ void k_norms_final(int nb, int nbf, NormPart part, const double *__restrict__ part6,
                              double *__restrict__ res) {
  norms_final_body(self_blk(), nb, nbf, part, part6, res);
}

// switch_old.f:22-52 on the surface nodes.
__device__ __forceinline__ void switch_old_body(Blk bk, int nnod, const int *__restrict__ surf_row, int *__restrict__ ifatm,
                             double *__restrict__ atmact, const double *__restrict__ atmpot,
                             double *__restrict__ pnew, double pmin) {
  int i = bk.b * blockDim.x + threadIdx.x;
  if (i >= nnod) return;
  int fa = ifatm[i];
  if (fa == -1) return;
  const int r = surf_row[i];
  const double pn = pnew[r], ap = atmpot[i], aa = atmact[i];
  if (fa == 1 && pn >= 0.0) {
    if (ap < 0.0 || aa > ap) {
      ifatm[i] = 0;
      atmact[i] = ap;
      return;
    }
  }
  if (fa == 1 && pn <= pmin) {
    if (ap > 0.0 || aa < ap) {
      ifatm[i] = 0;
      atmact[i] = ap;
      return;
    }
  }
  if (fa == 0 && pn >= 0.0 && ap >= 0.0) {
    ifatm[i] = 1;
    pnew[r] = 0.0;
    return;
  }
  if (fa == 0 && pn <= pmin && ap < 0.0) {
    ifatm[i] = 1;
    pnew[r] = pmin;
  }
}
__global__ void k_switch_old(int nnod, const int *__restrict__ surf_row, int *__restrict__ ifatm,
                             double *__restrict__ atmact, const double *__restrict__ atmpot,
                             double *__restrict__ pnew, double pmin) {
  switch_old_body(self_blk(), nnod, surf_row, ifatm, atmact, atmpot, pnew, pmin);
}

// switch.f:27-423: the switching CONVER calls when the surface is coupled
// (SURF).  PL = PONDNOD + ATMDIF*DELTAT/ARENOD is the head a node would pond
// to, with ATMDIF = ATMPOT - (ATMACT - QTRANIE); the IFATM 0 cases test
// PNEW itself.  Bands, in SWITCH's order: PL >= PONDH_MIN (the node ponds:
// IFATM 2, PNEW = PL, *ponding = 1), 0 <= PL < PONDH_MIN (IFATM 1),
// PMIN < PL < 0 and PL <= PMIN (IFATM 0, ATMACT = ATMPOT); rainfall with
// exfiltration ends every PL < 0 at IFATM 1; a ponded node under
// evaporation zeroes PNEW in the middle two bands; an IFATM 1 node at or
// below PMIN under evaporation stays unless it exfiltrates faster than the
// potential rate.  OVFLNOD (the overland flux the routing takes) is
// written where SWITCH writes it.  pondnod / ovflnod null: no routing (no
// ponded water, OVFLNOD not kept).
__global__ void k_switch_surf(int nnod, const int *__restrict__ surf_row, int *__restrict__ ifatm,
                              double *__restrict__ atmact, const double *__restrict__ atmpot,
                              const double *__restrict__ qtr, const double *__restrict__ arenod, double deltat,
                              double pondh_min, double pmin, double *__restrict__ pnew, int *__restrict__ ponding,
                              const double *__restrict__ pondnod, double *__restrict__ ovflnod) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnod) return;
  const int fa = ifatm[i];
  if (fa == -1) {
    if (ovflnod) ovflnod[i] = 0.0;
    return;
  }
  const int r = surf_row[i];
  const double ap = atmpot[i], aa = atmact[i], pn = pnew[r], ar = arenod[i];
  const double pond = pondnod ? pondnod[i] : 0.0;
  double dif = ap - (aa - (qtr ? qtr[r] : 0.0));
  if (fabs(dif) < 1.0e-14) dif = 0.0;
  const double pl = pond + (dif * deltat / ar);
  const double drain = -pond * ar / deltat;
  auto band = [&](double h) {
    if (h >= pondh_min) return 1;
    if (h >= 0.0 && h < pondh_min) return 2;
    if (h > pmin && h < 0.0) return 3;
    if (h <= pmin) return 4;
    return 0;
  };
  double ov = 0.0;
  bool set_ov = true;
  if (fa == 0) {
    const double over = (pn - pond) * ar / deltat;
    switch (band(pn)) {
      case 1: ifatm[i] = 2; *ponding = 1; ov = over; break;
      case 2: ifatm[i] = 1; ov = over; break;
      case 3: ov = drain; break;
      case 4: ifatm[i] = 1; pnew[r] = pmin; ov = drain; break;
      default: set_ov = false; break;
    }
  } else if (fa == 1 || fa == 2) {
    const bool rain = ap >= 0.0, infil = aa >= 0.0;
    if (fa == 1 && !rain && pn <= pmin) {  // cases G and H
      set_ov = false;
      if (!infil && aa < ap) {
        ifatm[i] = 0;
        atmact[i] = ap;
        pnew[r] = pmin;
        ov = 0.0;
        set_ov = true;
      }
    } else {
      const bool zero_p = fa == 2 && !rain;
      int b = band(pl);
      if (rain && !infil && b >= 3) b = 5;  // cases B and F
      switch (b) {
        case 1: ifatm[i] = 2; pnew[r] = pl; *ponding = 1; ov = dif; break;
        case 2: ifatm[i] = 1; if (zero_p) pnew[r] = 0.0; ov = dif; break;
        case 3: ifatm[i] = 0; atmact[i] = ap; if (zero_p) pnew[r] = 0.0; ov = drain; break;
        case 4: ifatm[i] = 0; atmact[i] = ap; ov = drain; break;
        case 5: ifatm[i] = 1; ov = dif; break;
        default: set_ov = false; break;
      }
    }
  } else {
    set_ov = false;
  }
  if (ovflnod && set_ov) ovflnod[i] = ov;
}

// adrstn.f:22-31: an 'air dry' surface node under rainfall, or whose actual
// flux is below the potential one, returns to the Neumann condition at
// PNEW = PMIN (step starts and back-steps of a coupled run).
__global__ void k_adrstn(int nnod, const int *__restrict__ surf_row, int *__restrict__ ifatm,
                         const double *__restrict__ atmpot, double *__restrict__ atmact, double *__restrict__ pnew,
                         double pmin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnod) return;
  const int r = surf_row[i];
  if (pnew[r] <= pmin && (atmpot[i] > 0.0 || atmact[i] < atmpot[i])) {
    ifatm[i] = 0;
    atmact[i] = atmpot[i];
    pnew[r] = pmin;
  }
}

// ---- the routing block of a coupled step (cathy_main.f:2993-3058) ----
// nod_cell.f:24-27: OVFLNOD(K) = OVFLNOD(K)/ARENOD(K), in place.
__global__ void k_nod_cell_div(int nnod, const double *__restrict__ arenod, double *__restrict__ ovflnod) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nnod) ovflnod[s] = ovflnod[s] / arenod[s];
}
// nod_cell.f:30-71 with DOSTEP = 1: each active DEM cell (in TRIANGOLI's
// order) takes the mean of its four corner nodes' values times its area,
// stored at its INDEX_WITH_LAKES position.
__global__ void k_nod_cell(int ncell, const int *__restrict__ corner, const int *__restrict__ target, double dx,
                           double dy, const double *__restrict__ nodval, double *__restrict__ cellval) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double v = 0.0;
  for (int k = 0; k < 4; ++k) v = v + nodval[corner[4 * c + k]];
  v = v * 0.25;
  cellval[target[c]] = v * dx * dy;
}
// to[dst[i]] = from[src[i]]: TRANSFER_F3D_SURF (cell to routing order) and
// TRANSFER_SURF_F3D (back), for grids without lakes.
__global__ void k_gather_pairs(int m, const int *__restrict__ dst, const int *__restrict__ src,
                               const double *__restrict__ from, double *__restrict__ to) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) to[dst[i]] = from[src[i]];
}
// cell_nod.f:22-48 with DOSTEP = 1: the coarse cells' values (src -1: none).
__global__ void k_cell_coarse(int ncell, const int *__restrict__ src, const double *__restrict__ cellval,
                              double *__restrict__ coarse) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double v = 0.0;
  if (src[c] >= 0) v = v + cellval[src[c]];
  coarse[c] = v / 1.0;
}
// cell_nod.f:50-68: PONDNOD of a node is the mean of its triangles' cells,
// summed in triangle order.
__global__ void k_cell_nod(int nnod, const int *__restrict__ ptr, const int *__restrict__ list,
                           const double *__restrict__ coarse, double *__restrict__ pondnod) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnod) return;
  double v = 0.0;
  for (int j = ptr[s]; j < ptr[s + 1]; ++j) v = v + coarse[list[j]];
  pondnod[s] = v / (double)(ptr[s + 1] - ptr[s]);
}
// pondupd.f:19-35: a node whose ponded head reaches PONDH_MIN keeps it as
// its head (IFATM 1, 2) or adds it to its flux (IFATM 0); PONDING is set
// where one does (the caller clears it first).
__global__ void k_pondupd(int nnod, const int *__restrict__ surf_row, const int *__restrict__ ifatm,
                          double pondh_min, double dtr, const double *__restrict__ pondnod,
                          const double *__restrict__ arenod, const double *__restrict__ atmpot,
                          double *__restrict__ atmact, double *__restrict__ pnew, int *__restrict__ ponding) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnod) return;
  const int fa = ifatm[s];
  if (fa == -1 || !(pondnod[s] >= pondh_min)) return;
  if (fa == 1 || fa == 2) {
    pnew[surf_row[s]] = pondnod[s];
    *ponding = 1;
  } else if (fa == 0) {
    *ponding = 1;
    atmact[s] = atmpot[s] + pondnod[s] * arenod[s] * dtr;
  }
}
// cathy_main.f:3223-3225: no ponded head where the surface is unsaturated.
__global__ void k_pond_reset(int nnod, const int *__restrict__ surf_row, const double *__restrict__ pnew,
                             double *__restrict__ pondnod) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nnod && pnew[surf_row[s]] <= 0.0) pondnod[s] = 0.0;
}
// cathy_main.f:3099: POLD(1..NNOD) = PNEW(1..NNOD).
__global__ void k_copy_surf(int nnod, const int *__restrict__ surf_row, const double *__restrict__ pnew,
                            double *__restrict__ pold) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nnod) pold[surf_row[s]] = pnew[surf_row[s]];
}
// *x = v (one thread).
__global__ void k_put_double(double *__restrict__ x, double v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *x = v;
}

// *x = v (one thread).
__global__ void k_put_int(int *__restrict__ x, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *x = v;
}

// flow3d.f:235-240 / cathy_main.f:2976-2978: POLD = PNEW, PTOLD = PTNEW,
// PTNEW = TETAF*PNEW + (1-TETAF)*PTIMEP (weight.f).
__device__ __forceinline__ void next_iter_body(Blk bk, int n, double tetaf, const double *__restrict__ pnew,
                            const double *__restrict__ ptimep, double *__restrict__ pold,
                            double *__restrict__ ptnew, double *__restrict__ ptold) {
  int i = bk.b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double pn = pnew[i];
  pold[i] = pn;
  ptold[i] = ptnew[i];
  const double w = tetaf * pn + (1.0 - tetaf) * ptimep[i];
  ptnew[i] = w;
}
__global__ void k_next_iter(int n, double tetaf, const double *__restrict__ pnew,
                            const double *__restrict__ ptimep, double *__restrict__ pold,
                            double *__restrict__ ptnew, double *__restrict__ ptold) {
  next_iter_body(self_blk(), n, tetaf, pnew, ptimep, pold, ptnew, ptold);
}

// Start of a step: POLD = PNEW, PTNEW = weight, PTOLD = PTNEW.
__global__ void k_start_iter(int n, double tetaf, const double *__restrict__ pnew,
                             const double *__restrict__ ptimep, double *__restrict__ pold,
                             double *__restrict__ ptnew, double *__restrict__ ptold) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double pn = pnew[i];
  pold[i] = pn;
  const double w = tetaf * pn + (1.0 - tetaf) * ptimep[i];
  ptnew[i] = w;
  ptold[i] = w;
}

// atmnxt.f:41-73 (label 300) for one interpolated rate (HSPATM=1) or a
// per-node rate array: ATMPOT = rate*area, and ATMACT for IFATM = 0.
__global__ void k_atm_update(int nnod, double rate, const double *__restrict__ rate_node,
                             const double *__restrict__ arenod, const int *__restrict__ ifatm,
                             double scf, int set_act, double *__restrict__ atmpot,
                             double *__restrict__ atmact) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnod) return;
  const double rt = rate_node ? rate_node[s] : rate;
  const double ap = rt * arenod[s];
  atmpot[s] = ap;
  if (set_act && ifatm[s] == 0) atmact[s] = ap >= 0.0 ? ap : (1.0 - scf) * ap;
}

// neumann.f:19-24: free drainage Q = -A kr Kz at the base nodes.
__global__ void k_free_drainage(int nnod, const int *__restrict__ bot_row,
                                const double *__restrict__ arenod, const double *__restrict__ ckrw,
                                const double *__restrict__ kznod, double *__restrict__ qbot,
                                double *__restrict__ qnode) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnod) return;
  const int r = bot_row[s];
  const double q = -1.0 * arenod[s] * ckrw[r] * kznod[r];
  qbot[s] = q;
  qnode[r] = q;
}

// ------------------------------------ listed-node boundary conditions --
// x[rows[i]] = vals[i] (vals null: v0).  SHLPIC's PNEW(CONTP) = PRESC and
// seepage PNEW = 0 (shlpic.f:30-45); the rows of a list are distinct.
__global__ void k_set_rows(int m, const int *__restrict__ rows, const double *__restrict__ vals, double v0,
                           double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) x[rows[i]] = vals ? vals[i] : v0;
}
// x[rows[i]] += vals[i] over distinct rows: the listed Neumann fluxes.
__global__ void k_add_rows(int m, const int *__restrict__ rows, const double *__restrict__ vals,
                           double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) x[rows[i]] = x[rows[i]] + vals[i];
}
// set (set != 0) or clear `bit` of bcfix at the listed rows.
__global__ void k_mark_rows(int m, const int *__restrict__ rows, int bit, int set, uint8_t *__restrict__ bcfix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int r = rows[i];
  bcfix[r] = set ? (uint8_t)(bcfix[r] | bit) : (uint8_t)(bcfix[r] & ~bit);
}
// x[idx[i]] = v: IFATM = -1 at surface nodes that carry another boundary
// condition (atmone.f:101-117, atmnxt.f:80-102).
__global__ void k_set_int(int m, const int *__restrict__ idx, int v, int *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) x[idx[i]] = v;
}
// bkpic.f:28-49 at listed rows: the back-calculated flux
// QNEW = (A PDIFF - XT5 - (1-TETAF) QOLD) / TETAF of a fixed-head node
// (QPNEW) or, where mask = 1, of a seepage-face exit node (SFQ).
__global__ void k_bkpic_rows(int m, const int *__restrict__ rows, const int *__restrict__ mask, Sell sl,
                             const double *__restrict__ A, const double *__restrict__ pdiff,
                             const double *__restrict__ xt5, double tetaf, const double *__restrict__ qold,
                             double *__restrict__ qnew) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (mask && mask[i] != 1) return;
  const int r = rows[i];
  const RowPos rp = row_pos(sl, r);
  const int len = sell_len(sl, rp, r);
  double v = 0.0;
  for (int j = 0; j < len; ++j) {
    const int q = sp(sl.so, r, j);
    v += A[q] * pdiff[sell_col(sl, rp, q, j)];
  }
  v = v - xt5[r];
  qnew[i] = (v - (1.0 - tetaf) * qold[i]) * (1.0 / tetaf);
}
// fluxmb.f:30-37 / :53-62: one work-group sums the positive and the
// negative entries of q (where mask = 1, if given) into out2[0], out2[1].
__global__ void k_list_flux(int m, const double *__restrict__ q, const int *__restrict__ mask,
                            double *__restrict__ out2) {
  __shared__ double sh[kBlock / 64 + 1];
  double pos = 0.0, neg = 0.0;
  for (int i = threadIdx.x; i < m; i += kBlock) {
    if (mask && mask[i] != 1) continue;
    const double v = q[i];
    if (v > 0.0) pos += v; else neg += v;
  }
  const double tp = block_sum<kBlock>(pos, sh);
  __syncthreads();
  const double tn = block_sum<kBlock>(neg, sh);
  if (threadIdx.x == 0) {
    out2[0] = tp;
    out2[1] = tn;
  }
}

// extall.f:22-47 + extcvg.f:20-26, one work-group over the seepage-face
// nodes (face by face, top to bottom): an exit node (SFEX = 1) whose
// back-calculated flux SFQ turned inflowing (>= 0) leaves the exit set
// (SFQ = 0); any other node whose head rose above 1e-8 joins it (PNEW = 0).
// bcfix bit 4 follows SFEX; out[0] = KSF, the nodes whose SFEX differs
// from SFEXIT (the previous iteration's).
__global__ void k_extall(int m, const int *__restrict__ rows, int *__restrict__ sfex,
                         const int *__restrict__ sfexit, double *__restrict__ sfq, double *__restrict__ pnew,
                         uint8_t *__restrict__ bcfix, double *__restrict__ out) {
  __shared__ double sh[kBlock / 64 + 1];
  double ksf = 0.0;
  for (int i = threadIdx.x; i < m; i += kBlock) {
    const int r = rows[i];
    int e = sfex[i];
    if (e == 1) {
      if (sfq[i] >= 0.0) {
        e = 0;
        sfq[i] = 0.0;
      }
    } else if (pnew[r] > 1.0e-8) {
      pnew[r] = 0.0;
      e = 1;
    }
    sfex[i] = e;
    bcfix[r] = e == 1 ? (uint8_t)(bcfix[r] | 4) : (uint8_t)(bcfix[r] & ~4);
    if (e != sfexit[i]) ksf += 1.0;
  }
  const double t = block_sum<kBlock>(ksf, sh);
  if (threadIdx.x == 0) out[0] = t;
}
// bcfix bit 4 from SFEX (after a back-step restores SFEX).
__global__ void k_sync_exit(int m, const int *__restrict__ rows, const int *__restrict__ sfex,
                            uint8_t *__restrict__ bcfix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int r = rows[i];
  bcfix[r] = sfex[i] == 1 ? (uint8_t)(bcfix[r] | 4) : (uint8_t)(bcfix[r] & ~4);
}
// SHLPIC's PNEW = 0 at the exit nodes (shlpic.f:36-45).
__global__ void k_zero_exit(int m, const int *__restrict__ rows, const int *__restrict__ sfex,
                            double *__restrict__ pnew) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && sfex[i] == 1) pnew[rows[i]] = 0.0;
}

// etran.f:15-80: Feddes root water uptake QTRANIE.  The potential
// evaporation ETP = -SCF*ATMPOT (ATMPOT < 0, else 0) of a surface column is
// spread over the root-zone nodes (depth d <= ZROOT of the column's
// vegetation type) in proportion to BTRANI = max(0, beta dz g), beta =
// (1 - d/ZROOT) exp(-PZ d/ZROOT), g the water-stress factor of PSI between
// PCWLT, PCREF and PCANA:
//   QTRANIE = ETP BTRANI / BTRAN / max(OMG, OMGC),
// BTRAN = sum beta dz, OMG = sum g beta dz / BTRAN over the root zone.
// One thread per column; zc and the columns' nodes k = (j-1) NNOD + i are
// in CATHY numbering, perm maps them to rows.  Every root zone lies above
// its column's base (checked on the host), so the layer loop ends before
// the base node.
__global__ void k_etran(int nnod, int nstr, const int *__restrict__ perm, const double *__restrict__ zc,
                        const double *__restrict__ pnew, const double *__restrict__ atmpot,
                        const int *__restrict__ vegc, VegTab vt, double scf, double *__restrict__ qtr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnod) return;
  const int v = vegc[i];
  const double zr = vt.zroot[v], pca = vt.pcana[v], pref = vt.pcref[v], pwlt = vt.pcwlt[v], pzv = vt.pz[v];
  const double ap = atmpot[i];
  const double etp = ap < 0.0 ? -1.0 * scf * ap : 0.0;
  const double zsurf = zc[i];
  double btran = 0.0, omg = 0.0, depth = 0.0;
  int j = 1;
  while (depth <= zr && j <= nstr) {
    const int k = (j - 1) * nnod + i;
    const double s1 = pca, s2 = pca + 1.0e-03;
    const double dz = j == 1 ? (zsurf - zc[k + nnod]) / 2.0 : (zc[k - nnod] - zc[k + nnod]) / 2.0;
    const double sh2o = pnew[perm[k]];
    const double gx1 = fmin(1.0, fmax(0.0, (sh2o - pwlt) / (pref - pwlt)));
    const double gx2 = fmin(1.0, fmax(0.0, 1.0 - (sh2o - s1) / (s2 - s1)));
    const double gx = fmin(gx1, gx2);
    const double beta = (1.0 - depth / zr) * exp(-1.0 * pzv * depth / zr);
    qtr[perm[k]] = fmax(0.0, beta * dz * gx);  // BTRANI until the column's totals are known
    btran = btran + beta * dz;
    omg = omg + gx * beta * dz;
    j = j + 1;
    depth = zsurf - zc[(j - 1) * nnod + i];
  }
  const int jroot = j - 1;  // layers 1..jroot carry BTRANI
  btran = fmax(0.0, btran);
  omg = omg / btran;
  const double den = fmax(omg, vt.omgc[v]);
  for (int l = 1; l <= nstr + 1; ++l) {
    const int r = perm[(l - 1) * nnod + i];
    const double b = l <= jroot ? qtr[r] : 0.0;
    qtr[r] = etp * b / btran / den;
  }
}

// storcal.f partial: sum Sw * VOLNOD * PNODI.
__device__ __forceinline__ void store_body(Blk bk, int n, const double *__restrict__ sw, const double *__restrict__ volnod,
                        const double *__restrict__ pnodi, double *__restrict__ part) {
  __shared__ double sh[kBlock / 64 + 1];
  double acc = 0.0;
  for (int i = bk.b * blockDim.x + threadIdx.x; i < n; i += bk.nb * blockDim.x)
    acc += sw[i] * volnod[i] * pnodi[i];
  double t = block_sum<kBlock>(acc, sh);
  if (threadIdx.x == 0) part[bk.b] = t;
}
__global__ void k_store(int n, const double *__restrict__ sw, const double *__restrict__ volnod,
                        const double *__restrict__ pnodi, double *__restrict__ part) {
  store_body(self_blk(), n, sw, volnod, pnodi, part);
}

__device__ __forceinline__ void sum_final_body(Blk bk, int nb, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[kBlock / 64 + 2];
  const double v = reduce_partials<kBlock>(part, nb, sh);
  if (threadIdx.x == 0) *out = v;
}
__global__ void k_sum_final(int nb, const double *__restrict__ part, double *__restrict__ out) {
  sum_final_body(self_blk(), nb, part, out);
}

// Set-up check of the coded column source: every slot of every row must
// decode to the explicit column id (and the length / diagonal match).
__global__ void k_check_codes(int n, Sell sl, int *__restrict__ bad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const RowPos rp = row_pos(sl, r);
  const int s = r >> 6, w = (sl.so[s + 1] - sl.so[s]) >> 6;
  int nb = 0;
  if (sell_len(sl, rp, r) != sl.len[r] || sell_diag(sl, rp, r) != sl.diag[r]) ++nb;
  for (int j = 0; j < w; ++j) {
    const int q = sl.so[s] + 64 * j + (r & 63);
    if (sell_col(sl, rp, q, j) != sl.col[q]) ++nb;
  }
  if (nb) atomicAdd(bad, nb);
}

// A[q] = COEF1[map[q]] (upper CSR value of the slot's symmetric entry).
__global__ void k_gather_coef(int nslots, const int *__restrict__ map, const double *__restrict__ coef,
                              double *__restrict__ A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nslots) {
    const int m = map[q];
    A[q] = m >= 0 ? coef[m] : 0.0;
  }
}

// partial sums of u^2 over the free (non-Dirichlet) rows
__global__ void k_sumsq_free(int n, const double *__restrict__ u, const uint8_t *__restrict__ isdir,
                             double *__restrict__ part) {
  __shared__ double sh[kBlock / 64 + 1];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (!isdir[i]) acc += u[i] * u[i];
  double t = block_sum<kBlock>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_copy(size_t n, const double *__restrict__ src, double *__restrict__ dst) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void k_fill(size_t n, double v, double *__restrict__ dst) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// ----------------------------------------------------- velocities --
// ------------------------------------------------------------ EnKF --
// Ensemble analysis on member states held in place (enkf.py:16-224,
// enkf_analysis): with the N x N weight W = S^T COV^-1 D / (N-1) formed
// from the observation-space quantities, P_xo COV^-1 D = X' W, so per row
//   xa_j = x_j + sum_k (x_k - xbar) W(k, j),
// then, for inflation alpha != 1 (enkf.py:316-318), xa = mean(xa) +
// alpha (xa - mean(xa)).  One thread per row; the member values of the
// row live in registers (N <= kEnkfMax).
__global__ void k_enkf_update(int n, int nens, const double *const *__restrict__ X, const double *__restrict__ W,
                              double alpha, double *const *__restrict__ Xa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[kEnkfMax], xa[kEnkfMax];
  double mean = 0.0;
  for (int k = 0; k < nens; ++k) {
    x[k] = X[k][i];
    mean += x[k];
  }
  mean /= nens;
  double meana = 0.0;
  for (int j = 0; j < nens; ++j) {
    double s = x[j];
    for (int k = 0; k < nens; ++k) s = fma(x[k] - mean, W[(size_t)j * nens + k], s);
    xa[j] = s;
    meana += s;
  }
  meana /= nens;
  for (int j = 0; j < nens; ++j) Xa[j][i] = alpha == 1.0 ? xa[j] : meana + alpha * (xa[j] - meana);
}

// Gaspari-Cohn weight of distance d for radius L (DA/localisation.py:
// 136-151): the fifth-order piecewise rational function in z = |d|/L, in
// the reference's nesting; 0 from z = 2 on.
__device__ __forceinline__ double gaspari_cohn(double d, double L) {
  const double z = fabs(d) / L;
  if (z <= 1.0) return (((-0.25 * z + 0.5) * z + 0.625) * z - 5.0 / 3.0) * (z * z) + 1.0;
  if (z <= 2.0) return ((((z / 12.0 - 0.5) * z + 0.625) * z + 5.0 / 3.0) * z - 5.0) * z + 4.0 - 2.0 / (3.0 * z);
  return 0.0;
}

// Localised analysis of the states (enkf.py:304-321), one thread per state
// row (device order; ri its CATHY row):
//   c_j = (sum_k x'_k S(j,k)) / (N-1),   x_a,k = x_k + sum_j c_j L(ri,j) invD(j,k),
// S(j,k) at s[j*nens + k], invD likewise; L(ri,j) from loc[ri*m + j], or
// 1 where row ri and observation j lie in one zone and 0 elsewhere, or the
// Gaspari-Cohn weight of the (x, y) distance from node ri to observation
// j.  Then inflation about the analysis mean.  Member
// values stay in registers (NE >= nens, the rest masked); every row of
// every member is read before the thread writes, so the update is in place.
template <int NE>
__global__ __launch_bounds__(256) void k_enkf_localized(LocArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int nens = a.nens;
  const int ri = a.iperm[i];
  double xp[NE], dx[NE];
  double mean = 0.0;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    xp[k] = k < nens ? a.X[k][i] : 0.0;
    mean += xp[k];
  }
  mean /= nens;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    xp[k] = k < nens ? xp[k] - mean : 0.0;
    dx[k] = 0.0;
  }
  double px = 0.0, py = 0.0;
  const int zi = a.row_zone ? a.row_zone[ri] : 0;
  if (!a.loc && !a.row_zone) {
    px = a.node_xy[2 * (size_t)ri];
    py = a.node_xy[2 * (size_t)ri + 1];
  }
  const double nm1 = (double)(nens - 1);
  bool hit = false;  // zones: an observation in the row's zone
  for (int j = 0; j < a.m; ++j) {
    const double *sj = a.s + (size_t)j * nens, *dj = a.invd + (size_t)j * nens;
    double c = 0.0;
#pragma unroll
    for (int k = 0; k < NE; ++k)
      if (k < nens) c = fma(xp[k], sj[k], c);
    c = c / nm1;
    double l;
    if (a.loc) {
      l = a.loc[(size_t)ri * a.m + j];
    } else if (a.row_zone) {
      const bool in = a.obs_zone[j] == zi;
      hit = hit || in;
      l = in ? 1.0 : 0.0;
    } else {
      const double ex = a.obs_xy[2 * j] - px, ey = a.obs_xy[2 * j + 1] - py;
      l = gaspari_cohn(sqrt(ex * ex + ey * ey), a.radius);
    }
    const double cl = c * l;
#pragma unroll
    for (int k = 0; k < NE; ++k)
      if (k < nens) dx[k] = fma(cl, dj[k], dx[k]);
  }
  if (a.row_zone && !hit) return;  // a zone with no observation is not analysed
  double meana = 0.0;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    dx[k] = k < nens ? a.X[k][i] + dx[k] : 0.0;  // x_a
    meana += dx[k];
  }
  meana /= nens;
#pragma unroll
  for (int k = 0; k < NE; ++k)
    if (k < nens) a.Xa[k][i] = a.alpha == 1.0 ? dx[k] : meana + a.alpha * (dx[k] - meana);
}
template __global__ void k_enkf_localized<8>(LocArgs);
template __global__ void k_enkf_localized<16>(LocArgs);
template __global__ void k_enkf_localized<32>(LocArgs);
template __global__ void k_enkf_localized<64>(LocArgs);

// Gram matrix of the member perturbations, G(i,j) = sum_rows x'_i x'_j
// (x' = x - row mean), for the dual analysis.  Each work-group stages
// kGramRows rows of all members in LDS with their row means, then thread t
// accumulates the pairs p = t, t + 256, ... of the NE x NE matrix over its
// rows; one partial matrix per work-group (part[b * NE * NE + p]), summed
// in work-group order by k_enkf_gram_sum.
template <int NE>
__global__ __launch_bounds__(256) void k_enkf_gram(int n, int nens, const double *const *__restrict__ X,
                                                   double *__restrict__ part) {
  constexpr int R = kGramRows;
  constexpr int NP = NE * NE / 256 > 0 ? NE * NE / 256 : 1;
  __shared__ double xs[R * NE];
  __shared__ double mu[R];
  double acc[NP];
#pragma unroll
  for (int u = 0; u < NP; ++u) acc[u] = 0.0;
  for (int r0 = blockIdx.x * R; r0 < n; r0 += gridDim.x * R) {
    for (int e = threadIdx.x; e < R * NE; e += blockDim.x) {
      const int r = e / NE, k = e % NE, row = r0 + r;
      xs[e] = (row < n && k < nens) ? X[k][row] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < R) {
      double m = 0.0;
      for (int k = 0; k < nens; ++k) m += xs[threadIdx.x * NE + k];
      mu[threadIdx.x] = m / nens;
    }
    __syncthreads();
    for (int r = 0; r < R; ++r) {
      if (r0 + r >= n) break;
      const double m = mu[r];
#pragma unroll
      for (int u = 0; u < NP; ++u) {
        const int p = threadIdx.x + 256 * u;
        const int i = p / NE, j = p % NE;
        if (p < NE * NE && i < nens && j < nens)
          acc[u] = fma(xs[r * NE + i] - m, xs[r * NE + j] - m, acc[u]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    const int p = threadIdx.x + 256 * u;
    if (p < NE * NE) part[(size_t)blockIdx.x * NE * NE + p] = acc[u];
  }
}
template __global__ void k_enkf_gram<8>(int, int, const double *const *, double *);
template __global__ void k_enkf_gram<16>(int, int, const double *const *, double *);
template __global__ void k_enkf_gram<32>(int, int, const double *const *, double *);
template __global__ void k_enkf_gram<64>(int, int, const double *const *, double *);

__global__ void k_enkf_gram_sum(int nb, int nn, const double *__restrict__ part, double *__restrict__ g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * nn + p];
  g[p] = s;
}

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC 2011): ten rounds of two 32x32 -> 64-bit
// multiplies with the key bumped by the Weyl constants between rounds.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k.x, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k.y, (uint32_t)p0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// One standard normal per (CATHY row, member, seed): Philox of counter
// (row, member, 0, 0) under the 64-bit seed, two 32-bit words to uniforms
// u1 in (0, 1], u2 in [0, 1), then Box-Muller sqrt(-2 ln u1) cos(2 pi u2).
__device__ __forceinline__ double jitter_normal(uint32_t row, uint32_t member, uint64_t seed) {
  const uint4 r = philox4x32_10(make_uint4(row, member, 0u, 0u), make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
  const double u1 = ((double)r.x + 1.0) * 2.3283064365386963e-10;  // 2^-32
  const double u2 = (double)r.y * 2.3283064365386963e-10;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Particle-filter resampling (pf.py:98-109) with jitter (pf.py:129-130):
// member j's row i becomes member idx[j]'s (src[j] = its state) plus
// sigma * N(0,1) from jitter_normal(CATHY row, j, seed) when sigma > 0.
// All members' rows are read before any is written (in place).
template <int NE>
__global__ __launch_bounds__(256) void k_pf_resample(int n, int nens, const double *const *__restrict__ src,
                                                     const int *__restrict__ iperm, double sigma, uint64_t seed,
                                                     double *const *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[NE];
#pragma unroll
  for (int j = 0; j < NE; ++j) v[j] = j < nens ? src[j][i] : 0.0;
#pragma unroll
  for (int j = 0; j < NE; ++j)
    if (j < nens) dst[j][i] = v[j];
  if (sigma > 0.0) {  // the thread's own rows, just written
    const uint32_t ri = (uint32_t)iperm[i];
    for (int j = 0; j < nens; ++j) {
      double *d = dst[j];
      d[i] = d[i] + sigma * jitter_normal(ri, (uint32_t)j, seed);
    }
  }
}
template __global__ void k_pf_resample<8>(int, int, const double *const *, const int *, double, uint64_t,
                                          double *const *);
template __global__ void k_pf_resample<16>(int, int, const double *const *, const int *, double, uint64_t,
                                           double *const *);
template __global__ void k_pf_resample<32>(int, int, const double *const *, const int *, double, uint64_t,
                                           double *const *);
template __global__ void k_pf_resample<64>(int, int, const double *const *, const int *, double, uint64_t,
                                           double *const *);

// vel3d.f:28-52 per element: kr_e = mean nodal kr (nodelt.f), then
//   UU = (sum psi_j b_j) (-kr_e VOLUR IVOL) Kx, VV likewise with c / Ky,
//   WW = ((sum psi_j d_j) (-kr_e VOLUR IVOL) - kr_e) Kz.
// vt: per element b[4], c[4], d[4], VOLUR, IVOL, Kx, Ky, Kz (stride 17).
__global__ void k_vel3d(int nt, const int4 *__restrict__ enode, const double *__restrict__ pnew,
                        const double *__restrict__ ckrw, const double *__restrict__ vt, double *__restrict__ uu,
                        double *__restrict__ vv, double *__restrict__ ww) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  const int4 v = enode[e];
  const int nd[4] = {v.x, v.y, v.z, v.w};
  const double *t = vt + (size_t)e * 17;
  double kr = 0.0;
  for (int j = 0; j < 4; ++j) kr = kr + ckrw[nd[j]];
  kr = kr * 0.25;
  double bb = 0.0, cc = 0.0, dd = 0.0;
  const double xyz = -kr * t[12] * t[13];
  for (int j = 0; j < 4; ++j) {
    const double p = pnew[nd[j]];
    bb = bb + p * t[j];
    cc = cc + p * t[4 + j];
    dd = dd + p * t[8 + j];
  }
  uu[e] = bb * xyz * t[14];
  vv[e] = cc * xyz * t[15];
  ww[e] = (dd * xyz - kr) * t[16];
}

// vnod3d.f:17-32: nodal means over the incident elements, summed in element order.
__global__ void k_vnod3d(int n, const int *__restrict__ inc_ptr, const int *__restrict__ inc,
                         const double *__restrict__ uu, const double *__restrict__ vv,
                         const double *__restrict__ ww, double *__restrict__ unod, double *__restrict__ vnod,
                         double *__restrict__ wnod) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double su = 0.0, sv = 0.0, sw = 0.0;
  const int i0 = inc_ptr[r], i1 = inc_ptr[r + 1];
  for (int i = i0; i < i1; ++i) {
    const int e = inc[i] >> 2;
    su = su + uu[e];
    sv = sv + vv[e];
    sw = sw + ww[e];
  }
  const double tp = (double)(i1 - i0);
  unod[r] = su / tp;
  vnod[r] = sv / tp;
  wnod[r] = sw / tp;
}


// ------------------------------------------- batched ensemble launches --
// One grid over the members of an ensemble: each member's groups run the
// single-member body on its own arrays (MemberDev), with the group numbers
// of its own launch.
#define CG_BATCH_HEAD                                      \
  const int g_ = xcd_group();                              \
  const int m_ = batch_member(start, nmem, g_);            \
  const MemberDev &d = md[m_];                             \
  const Blk bk{g_ - start[m_], start[m_ + 1] - start[m_]};

__device__ __forceinline__ double *slot_part(const MemberDev &d, int k) { return d.part + (size_t)k * kMaxRedBlocks; }
template <typename FT>
__device__ __forceinline__ FT *fac_of(const MemberDev &d);
template <>
__device__ __forceinline__ float *fac_of<float>(const MemberDev &d) { return d.F32; }
template <>
__device__ __forceinline__ double *fac_of<double>(const MemberDev &d) { return d.F64; }
__device__ __forceinline__ double *vec_of(const MemberDev &d, int v) {
  return v == kVecTnoti ? d.tnoti : (v == kVecPdiff ? d.pdiff : (v == kVecR ? d.r : (v == kVecP ? d.p : d.z)));
}

__global__ void kb_constitutive(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  constitutive_body(bk, d.n, d.ptnew, d.pnew, d.soil, d.sw, d.ckrw, d.et1, d.et2, d.swnew);
}
__global__ void kb_bc_rows(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  bc_rows_body(bk, d.n, d.row_surf, d.qnode, d.ifatm, d.atmact, d.atmold, d.tetaf, nullptr, d.isdir, d.bcflux);
}
template <int MR>
__global__ __launch_bounds__(kBlock) void kb_assemble_plan(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  assemble_plan_body<MR>(bk, d.a, d.P, d.ckrw, d.et1);
}
template __global__ void kb_assemble_plan<16>(const int *, int, const MemberDev *);
template __global__ void kb_assemble_plan<32>(const int *, int, const MemberDev *);

template <typename FT, int NW>
__global__ __launch_bounds__(256) void kb_ic_factor_col3t(const int *start, int nmem, const MemberDev *md, int c) {
  extern __shared__ double smem[];
  CG_BATCH_HEAD
  ic_factor_col3t_body<FT, NW>(bk, smem, d.col[c], d.sl, d.ft, d.A, d.isdir, fac_of<FT>(d), d.dinv, d.flags + 1);
}
template <typename FT, int NW, bool LAST>
__global__ __launch_bounds__(256) void kb_fwd_col3(const int *start, int nmem, const MemberDev *md, int c, int vin,
                                                   int vout, int use_done) {
  extern __shared__ double smem[];
  CG_BATCH_HEAD
  fwd_col3_body<FT, NW, LAST>(bk, smem, d.col[c], d.sl, fac_of<FT>(d), d.dinv, vec_of(d, vin), vec_of(d, vout),
                              use_done ? d.flags : nullptr);
}
template <typename FT, int NW>
__global__ __launch_bounds__(256) void kb_bwd_col3(const int *start, int nmem, const MemberDev *md, int c, int vio,
                                                   int use_done) {
  extern __shared__ double smem[];
  CG_BATCH_HEAD
  double *x = vec_of(d, vio);
  bwd_col3_body<FT, NW>(bk, smem, d.col[c], d.sl, fac_of<FT>(d), d.dinv, x, x, use_done ? d.flags : nullptr);
}
#define CG_INST_BATCH_SWEEP(FT, U)                                                                         \
  template __global__ void kb_fwd_col3<FT, U, false>(const int *, int, const MemberDev *, int, int, int, int); \
  template __global__ void kb_fwd_col3<FT, U, true>(const int *, int, const MemberDev *, int, int, int, int);  \
  template __global__ void kb_bwd_col3<FT, U>(const int *, int, const MemberDev *, int, int, int);
#define CG_INST_BATCH_FT(FT)                                                                   \
  template __global__ void kb_ic_factor_col3t<FT, 0>(const int *, int, const MemberDev *, int);  \
  template __global__ void kb_ic_factor_col3t<FT, 4>(const int *, int, const MemberDev *, int);  \
  template __global__ void kb_ic_factor_col3t<FT, 8>(const int *, int, const MemberDev *, int);  \
  template __global__ void kb_ic_factor_col3t<FT, 12>(const int *, int, const MemberDev *, int); \
  CG_INST_BATCH_SWEEP(FT, 0)                                                                   \
  CG_INST_BATCH_SWEEP(FT, 4)                                                                   \
  CG_INST_BATCH_SWEEP(FT, 8)                                                                   \
  CG_INST_BATCH_SWEEP(FT, 12)                                                                  \
  CG_INST_BATCH_SWEEP(FT, 16)
CG_INST_BATCH_FT(float)
CG_INST_BATCH_FT(double)
#undef CG_INST_BATCH_FT
#undef CG_INST_BATCH_SWEEP

__global__ void kb_residual0(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  residual0_body(bk, d.n, d.sl, d.A, d.pdiff, d.tnoti, d.isdir, d.r, slot_part(d, 0));
}
__global__ void kb_pcg_init(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  pcg_init_body(bk, d.g_red, slot_part(d, 0), d.scal, d.flags, d.tol, d.imax);
}
__global__ void kb_spmv_dot2(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  spmv_dot2_body(bk, d.n, d.sl, d.fts, d.A, d.p, d.r, d.Bv, slot_part(d, 1), slot_part(d, 2), d.flags);
}
__global__ void kb_pcg_alpha(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  pcg_alpha_body(bk, d.n, d.g_rows, slot_part(d, 1), slot_part(d, 2), d.Bv, d.p, d.isdir, d.r, d.pdiff,
                 slot_part(d, 3), d.scal, d.flags);
}
__global__ void kb_dot(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  dot_body(bk, d.n, d.Bv, d.z, slot_part(d, 4), d.flags);
}
__global__ void kb_pcg_beta(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  pcg_beta_body(bk, d.n, d.g_vec, slot_part(d, 4), slot_part(d, 3), d.z, d.p, d.scal, d.flags);
}
// out[3m..3m+2] = done flag, iterations, last relative residual of member m
__global__ void kb_poll(const int *start, int nmem, const MemberDev *md, double *out) {
  CG_BATCH_HEAD
  (void)bk;
  if (threadIdx.x == 0) {
    out[3 * m_] = (double)d.flags[0];
    out[3 * m_ + 1] = (double)d.scal->niter;
    out[3 * m_ + 2] = d.scal->err;
  }
}
__global__ void kb_update_pnew(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  update_pnew_body(bk, d.n, d.pdiff, d.row_surf, d.ifatm, d.pold, d.pnew);
}
__global__ void kb_bkpic(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  bkpic_body(bk, d.nnod, d.surf_row, d.sl, d.A, d.pdiff, d.xt5, d.ifatm, d.atmold, d.tetaf, d.atmact);
}
__device__ __forceinline__ NormPart norm_part_of(const MemberDev &d) {
  return NormPart{slot_part(d, 5), slot_part(d, 6), slot_part(d, 7), slot_part(d, 8), slot_part(d, 9),
                  slot_part(d, 10), d.ipart};
}
__global__ void kb_norms(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  norms_body(bk, d.n, d.refid, d.pnew, d.pold, d.tnoti, d.swnew, d.swtimep, d.ptimep, d.volnod, d.snodi, d.pnodi,
             norm_part_of(d));
}
__global__ void kb_flux(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  flux_body(bk, d.nnod, d.ifatm, d.atmact, d.qbot, d.nq, slot_part(d, 11));
}
__global__ void kb_norms_final(const int *start, int nmem, const MemberDev *md) {
  CG_BATCH_HEAD
  norms_final_body(bk, d.g_red, d.g_flux, norm_part_of(d), slot_part(d, 11), d.res);
}
__global__ void kb_switch_old(const int *start, int nmem, const MemberDev *md, const int *act) {
  CG_BATCH_HEAD
  if (!(act[m_] & 1)) return;
  switch_old_body(bk, d.nnod, d.surf_row, d.ifatm, d.atmact, d.atmpot, d.pnew, d.pmin);
}
__global__ void kb_next_iter(const int *start, int nmem, const MemberDev *md, const int *act) {
  CG_BATCH_HEAD
  if (!(act[m_] & 2)) return;
  next_iter_body(bk, d.n, d.tetaf, d.pnew, d.ptimep, d.pold, d.ptnew, d.ptold);
}
#undef CG_BATCH_HEAD

}  // namespace cg
