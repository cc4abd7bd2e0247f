// flow_kernels.h -- declarations of the Picard/PCG kernels.
//
// Matrix layout: SELL-64.  Rows (device numbering) are grouped in slices of
// 64; slice s stores its rows' entries column-major, entry j of row r at
//   so[r >> 6] + 64*j + (r & 63),
// so the 64 lanes of a wavefront reading entry j of 64 consecutive rows
// touch one contiguous 512-byte (FP64) / 256-byte (FP32) segment.  Slice
// width = longest row of the slice; padding slots hold the row's own id and
// a zero value.  Each row's entries are sorted by column; len[r] is its
// length and diag[r] the slot of its diagonal.  The IC(0) factor lives in
// the same layout: lower slots L_rj, upper slots the transpose L_jr,
// diagonal L_rr.
#pragma once

#include "kernels.h"

namespace cg {

constexpr int kAsmBlock = 128;  // rows per assembly work-group (LDS row buffers)
constexpr int kChains = 8;      // chains per work-group in the column kernels (one row per thread)
constexpr int kUnroll = 16;     // row slots handled by the unrolled gather loops

// Column ids come from one of two sources.  Explicit: col[slot].  Coded
// (extruded meshes in the column ordering): the rows of colour c lie in
// tiles of TW chains (TW = nc_c: level-major, the whole colour plane by
// plane; TW = 1: chain-major; else TW chains plane by plane): row = base_c
// + tile*TW*nplane + k*w + (t - tile*TW), w = the tile's width
// (colour c, node plane k, rank t; chain g = goff_c + t) and every slot's
// column is (c', k + dk, t'), where the code (c', dk, t') depends only on
// the chain g and the plane type pt (0 surface, 1 interior, 2 base):
// gen3d.f repeats the same prism split in every layer, so all interior
// planes share one pattern.  code[(pt * width + j) * nchain + g] packs
// t' | (dk + 1) << 27 | c' << 29; clen / cdiag give the row length and the
// diagonal slot.  The set-up builds the codes from the explicit pattern
// and checks that they reproduce it exactly (else it keeps explicit ids).
struct Sell {
  const int *so;    // slice offsets
  const int *col;   // column id per slot (explicit source)
  const int *len;   // row lengths (explicit source)
  const int *diag;  // slot of the diagonal (explicit source)
  const uint32_t *code;
  const uint8_t *clen, *cdiag;
  int coded, nchain, width, nplane, ncolr;
  int ctw;   // chains per row tile of every colour class (0: level-major, one tile per class)
  int cbase[4], cnc[4], cgoff[4];
};

__device__ __forceinline__ int sp(const int *so, int r, int j) { return so[r >> 6] + j * 64 + (r & 63); }

// Position of a row in the column structure (coded source only).
struct RowPos {
  int k, g, pt;
};

template <typename T>
__device__ __forceinline__ T pick4(int i, T a0, T a1, T a2, T a3) {
  return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}

// Plane k and rank t of the row at offset off of a colour class with nc
// chains of np planes, in tiles of tw chains.
__device__ __host__ __forceinline__ void tile_decode(int off, int nc, int tw, int np, int &k, int &t) {
  const int tile = off / (tw * np);
  const int rem = off - tile * tw * np;
  const int w = nc - tile * tw < tw ? nc - tile * tw : tw;
  k = rem / w;
  t = tile * tw + (rem - k * w);
}
// Offset of (rank t, plane k) in such a class.
__device__ __host__ __forceinline__ int tile_offset(int t, int k, int nc, int tw, int np) {
  const int tile = t / tw;
  const int w = nc - tile * tw < tw ? nc - tile * tw : tw;
  return tile * tw * np + k * w + (t - tile * tw);
}

__device__ __forceinline__ RowPos row_pos(const Sell &sl, int r) {
  RowPos p{0, 0, 0};
  if (!sl.coded) return p;
  int c = 0;
#pragma unroll
  for (int i = 1; i < 4; ++i)
    if (i < sl.ncolr && r >= sl.cbase[i]) c = i;
  const int b = pick4(c, sl.cbase[0], sl.cbase[1], sl.cbase[2], sl.cbase[3]);
  const int nc = pick4(c, sl.cnc[0], sl.cnc[1], sl.cnc[2], sl.cnc[3]);
  const int off = r - b;
  int t;
  if (sl.ctw) {
    tile_decode(off, nc, sl.ctw, sl.nplane, p.k, t);
  } else {
    p.k = off / nc;
    t = off - p.k * nc;
  }
  p.g = pick4(c, sl.cgoff[0], sl.cgoff[1], sl.cgoff[2], sl.cgoff[3]) + t;
  p.pt = p.k == 0 ? 0 : (p.k == sl.nplane - 1 ? 2 : 1);
  return p;
}

__device__ __forceinline__ RowPos chain_pos(const Sell &sl, int g, int k) {
  RowPos p;
  p.k = k;
  p.g = g;
  p.pt = k == 0 ? 0 : (k == sl.nplane - 1 ? 2 : 1);
  return p;
}

// Code word of slot j of the row at p (coded source).
__device__ __forceinline__ uint32_t code_word(const Sell &sl, const RowPos &p, int j) {
  return sl.code[(size_t)(p.pt * sl.width + j) * sl.nchain + p.g];
}

// Column of a coded slot from its word, for a row on plane k.  Written as
// selects, not branches: a loop can load every word of a row first and
// decode afterwards, so the loads are in flight together.
__device__ __forceinline__ int code_col(const Sell &sl, int k, uint32_t w) {
  const int c2 = (int)(w >> 29), dk = (int)((w >> 27) & 3u) - 1, t2 = (int)(w & 0x7ffffffu);
  int b = sl.cbase[0], nc = sl.cnc[0];
  b = c2 == 1 ? sl.cbase[1] : b;
  nc = c2 == 1 ? sl.cnc[1] : nc;
  b = c2 == 2 ? sl.cbase[2] : b;
  nc = c2 == 2 ? sl.cnc[2] : nc;
  b = c2 == 3 ? sl.cbase[3] : b;
  nc = c2 == 3 ? sl.cnc[3] : nc;
  return sl.ctw ? b + tile_offset(t2, k + dk, nc, sl.ctw, sl.nplane) : b + (k + dk) * nc + t2;
}

// column of slot j (SELL position q) of the row at p / r
__device__ __forceinline__ int sell_col(const Sell &sl, const RowPos &p, int q, int j) {
  if (!sl.coded) return sl.col[q];
  return code_col(sl, p.k, code_word(sl, p, j));
}

// Column ids of a straight-line window of a row's slots: entry j is slot
// min(j0 + j, jl) (SELL position of slot 0 = q0; jl must be a slot of the
// row).  Every load is issued unconditionally and before any is used, so
// they are all in flight together.
template <int U>
__device__ __forceinline__ void sell_cols_window(const Sell &sl, const RowPos &p, int q0, int j0, int jl,
                                                 int (&c)[U]) {
  if (sl.coded) {
    uint32_t w[U];
#pragma unroll
    for (int j = 0; j < U; ++j) w[j] = code_word(sl, p, min(j0 + j, jl));
#pragma unroll
    for (int j = 0; j < U; ++j) c[j] = code_col(sl, p.k, w[j]);
  } else {
#pragma unroll
    for (int j = 0; j < U; ++j) c[j] = sl.col[q0 + 64 * min(j0 + j, jl)];
  }
}
__device__ __forceinline__ int sell_len(const Sell &sl, const RowPos &p, int r) {
  return sl.coded ? (int)sl.clen[p.pt * sl.nchain + p.g] : sl.len[r];
}
__device__ __forceinline__ int sell_diag(const Sell &sl, const RowPos &p, int r) {
  return sl.coded ? (int)sl.cdiag[p.pt * sl.nchain + p.g] : sl.diag[r];
}

// One colour class of the column ordering: chains t < nc of len planes
// (k = node plane, 0 = surface), in row tiles of tw chains: tw = nc
// (level-major, row base + k*nc + t), tw = 1 (chain-major, base + t*len +
// k) or tw = kChains (each work-group's rows one contiguous block).
struct ColDesc {
  int base, nc, len, g0;  // g0: chain id of t = 0
  int nf, nb;             // most earlier- / later-colour slots of a row (gathered by the sweeps)
  int tw;                 // chains per row tile
};
// rows between planes k and k+1 of chain t (its tile's width)
__device__ __host__ __forceinline__ int col_w(const ColDesc &C, int t) {
  const int t1 = (t / C.tw) * C.tw;
  return C.nc - t1 < C.tw ? C.nc - t1 : C.tw;
}
__device__ __host__ __forceinline__ int col_row(const ColDesc &C, int t, int k) {
  return C.tw >= C.nc ? C.base + k * C.nc + t : C.base + tile_offset(t, k, C.nc, C.tw, C.len);
}
// A column work-group holds kChains chains, one row per thread (or a loop
// over them): element i of its LDS arrays is chain t0 + t_local, plane k.
// Tiles of several chains put them side by side (i = k*kChains + t_local,
// consecutive threads on consecutive chains, as the rows lie); chain-major
// classes one chain after the other (i = t_local*len + k).
__device__ __forceinline__ void chain_elem(const ColDesc &C, int t0, int i, int &k, int &t) {
  if (C.tw == 1) {
    const int tl = i / C.len;
    k = i - tl * C.len;
    t = t0 + tl;
  } else {
    k = i / kChains;
    t = t0 + (i - k * kChains);
  }
}
// LDS index of (local chain tl, plane k) and the step between planes
__device__ __forceinline__ int chain_lds(const ColDesc &C, int tl, int k) {
  return C.tw == 1 ? tl * C.len + k : k * kChains + tl;
}
__device__ __forceinline__ int chain_step(const ColDesc &C) { return C.tw == 1 ? 1 : kChains; }

// One level of the natural-order schedule: rows lvl_rows[off .. off+cnt).
struct ListLevel {
  int off, cnt;
};

struct PcgScal {
  double xlung, pb, alpha, err, tol;
  int niter, imax;
};

struct NormPart {
  double *pl2, *fl2, *dstore, *finf, *pinf, *pdif;
  int *pidx;
};

struct AsmArgs {
  int n;
  Sell sl;
  const int *inc_ptr, *inc;
  const uint32_t *incoff;
  const double *S, *gz, *volu, *ckrwe, *et1e, *m4d, *m4full, *lmass;
  const double *ptnew, *pnew, *ptimep, *swnew, *swtimep, *et2;
  const uint8_t *isdir;
  const double *bcflux;
  const double *qtr;  // root uptake QTRANIE per row (null: none)
  double tetaf, rdt;
  double *A, *tnoti, *xt5;
};

__global__ void k_constitutive(int n, const double *ptnew, const double *pnew, SoilSoA s, double *sw,
                               double *ckrw, double *et1, double *et2, double *swnew);
__global__ void k_sw_kr(int n, const double *psi_in, SoilSoA s, double *sw, double *ckrw);
__global__ void k_elem_avg(int nt, const int4 *enode, const double *ckrw, const double *et1, double *ckrwe,
                           double *et1e);
__global__ void k_bc_rows(int n, const int *row_surf, const double *qnode, const int *ifatm,
                          const double *atmact, const double *atmold, double tetaf, const uint8_t *bcfix,
                          uint8_t *isdir, double *bcflux);
// listed-node boundary conditions (fixed head, Neumann lists, seepage faces)
__global__ void k_set_rows(int m, const int *rows, const double *vals, double v0, double *x);
__global__ void k_add_rows(int m, const int *rows, const double *vals, double *x);
__global__ void k_mark_rows(int m, const int *rows, int bit, int set, uint8_t *bcfix);
__global__ void k_set_int(int m, const int *idx, int v, int *x);
__global__ void k_bkpic_rows(int m, const int *rows, const int *mask, Sell sl, const double *A,
                             const double *pdiff, const double *xt5, double tetaf, const double *qold,
                             double *qnew);
__global__ void k_list_flux(int m, const double *q, const int *mask, double *out2);
__global__ void k_extall(int m, const int *rows, int *sfex, const int *sfexit, double *sfq, double *pnew,
                         uint8_t *bcfix, double *out);
__global__ void k_sync_exit(int m, const int *rows, const int *sfex, uint8_t *bcfix);
__global__ void k_zero_exit(int m, const int *rows, const int *sfex, double *pnew);
// Feddes root-uptake parameters per vegetation type (SOILCHAR.H /FEDDES/)
struct VegTab {
  const double *pcana, *pcref, *pcwlt, *zroot, *pz, *omgc;
};
__global__ void k_etran(int nnod, int nstr, const int *perm, const double *zc, const double *pnew,
                        const double *atmpot, const int *vegc, VegTab vt, double scf, double *qtr);
template <int MR, bool LUMPED>
__global__ void k_assemble(AsmArgs a);

// Precomputed lumped-mass assembly plan.  Per row, one record per incident
// element (ascending element id), lane-interleaved per slice: record m of
// row r at off[r>>6] + 64 m + (r&63) holds the element id (-1 = padding),
// the row slots of its 4 nodes (4 bytes; 0xFFFFFFFF = padding, as no slot
// byte is 0xFF) and the element-matrix row S_e(k, 0..3) as a double4.  w2 / gw (SELL layout) give the lumped
// storage and gravity terms as weights of the nodal ET1 / kr of the row's
// neighbours.
struct AsmPlan {
  const int *off;
  const int *ie;
  const uint32_t *io;
  const double4 *s4;
  const double *w2, *gw;
};
template <int MR>
__global__ void k_assemble_plan(AsmArgs a, AsmPlan P, const double *ckrw_node, const double *et1_node);

// Static IC(0) update pattern: for the lower slot q of row r (column c),
// pp[q]..pp[q+1] index pairs (slot of row r, slot of row c) with equal
// columns below c -- the terms of sum_m L_rm L_cm.
struct PairList {
  const int *pp;    // per SELL slot
  const int2 *pr;   // (slot in row r, slot in row c)
};
template <typename FT>
__global__ void k_ic_factor_col3(ColDesc C, Sell sl, PairList pl, const int *tpos, const double *A,
                                 const uint8_t *isdir, FT *F, double *dinv, int *nfall);

// IC(0) tables of an extruded mesh in the column ordering.  A row's
// coupling pattern depends only on its chain g and its class: plane 0,
// plane 1, planes 2..len-3, plane len-2, plane len-1 (the first and last
// two planes see the surface / base rows, whose slot layouts differ).  Per
// (class, g), kFacWords 32-bit words; word w at w[(cls*kFacWords + w)*nchain + g]:
//   bytes [0, 120):   for lower slot j (1 <= j < 16) and own slot a < j,
//                     byte j(j-1)/2 + a = slot b of row c = col(r, j) with
//                     col(c, b) = col(r, a), or 0xFF: the update pairs
//                     L_ra L_cb of L_rj;
//   bytes [120, 136): byte 120 + a: the same for the chain predecessor;
//   bytes [136, 152): byte 136 + j: slot of r in row col(r, j), j < 16 (the
//                     transposed position of L_rj).
// The set-up checks every row against the explicit pattern.
constexpr int kFacWords = 40;
struct FacTab {
  const uint32_t *w;
  int nchain;
};
__host__ __device__ inline int fac_class(int k, int len) {
  return k == 0 ? 0 : (k == len - 1 ? 4 : (k == 1 ? 1 : (k == len - 2 ? 3 : 2)));
}
// Column ordering, one colour class, from the tables: every gather of a
// row is issued before any is used and the row's own entries stay in
// registers.  Needs len*kChains <= 256 and at most NW earlier-colour slots
// per row.
template <typename FT, int NW>
__global__ void k_ic_factor_col3t(ColDesc C, Sell sl, FacTab ft, const double *A, const uint8_t *isdir, FT *F,
                                  double *dinv, int *nfall);
// LAST: the colour has no later-colour neighbours, so its backward sweep
// runs in the same launch and y holds the final x for its rows.
template <typename FT, int NW, bool LAST>
__global__ void k_fwd_col3(ColDesc C, Sell sl, const FT *F, const double *dinv, const double *tin, double *y,
                           const int *done);
template <typename FT, int NW>
__global__ void k_bwd_col3(ColDesc C, Sell sl, const FT *F, const double *dinv, const double *xin, double *x,
                           const int *done);
// Gather window of the column sweeps: the smallest instantiated size that
// covers n slots (larger rows take the kernels' loop).
inline int sweep_window(int n) { return n <= 0 ? 0 : (n <= 4 ? 4 : (n <= 8 ? 8 : (n <= 12 ? 12 : 16))); }
template <typename FT>
__global__ void k_ic_factor_list(ListLevel L, const int *rows, Sell sl, PairList pl, const int *tpos,
                                 const double *A, const uint8_t *isdir, FT *F, double *dinv, int *nfall);
template <typename FT>
__global__ void k_fwd_list(ListLevel L, const int *rows, Sell sl, const FT *F, const double *dinv,
                           const double *tin, double *y, const int *done);
template <typename FT>
__global__ void k_bwd_list(ListLevel L, const int *rows, Sell sl, const FT *F, const double *dinv,
                           const double *xin, double *x, const int *done);

__global__ void k_residual0(int n, Sell sl, const double *A, const double *x, const double *b,
                            const uint8_t *isdir, double *r, double *part);
// ft.w != null (column ordering, chain tables valid): each row reads its
// diagonal and upper slots and takes a lower entry a_rj = a_jr from row j's
// upper slot (the assembled matrix is bitwise symmetric), so HBM serves
// about one triangle and the last-level cache the other.
__global__ void k_spmv_dot2(int n, Sell sl, FacTab ft, const double *A, const double *p, const double *r,
                            double *B, double *part_pr, double *part_pb, const int *done);
__global__ void k_pcg_alpha(int n, int nparts, const double *part_pr, const double *part_pb, const double *B,
                            const double *p, const uint8_t *isdir, double *r, double *x, double *part_rr,
                            PcgScal *sc, const int *done);
__global__ void k_dot(int n, const double *u, const double *v, double *part, const int *done);
__global__ void k_pcg_beta(int n, int nparts, const double *part_bz, const double *part_rr, const double *z,
                           double *p, PcgScal *sc, int *done);
__global__ void k_pcg_init(int nparts, const double *part, PcgScal *sc, int *done, double tol, int imax);
__global__ void k_update_pnew(int n, const double *pdiff, const int *row_surf, const int *ifatm,
                              const double *pold, double *pnew);
__global__ void k_bkpic(int nnod, const int *surf_row, Sell sl, const double *A, const double *pdiff,
                        const double *xt5, const int *ifatm, const double *atmold, double tetaf, double *atmact);
__global__ void k_norms(int n, const int *refid, const double *pnew, const double *pold, const double *tnoti,
                        const double *swnew, const double *swtimep, const double *ptimep, const double *volnod,
                        const double *snodi, const double *pnodi, NormPart part);
__global__ void k_flux(int nnod, const int *ifatm, const double *atmact, const double *qbot, int nq,
                       double *part6);
__global__ void k_norms_final(int nb, int nbf, NormPart part, const double *part6, double *res);
__global__ void k_switch_surf(int nnod, const int *surf_row, int *ifatm, double *atmact, const double *atmpot,
                              const double *qtr, const double *arenod, double deltat, double pondh_min, double pmin,
                              double *pnew, int *ponding, const double *pondnod, double *ovflnod);
__global__ void k_nod_cell_div(int nnod, const double *arenod, double *ovflnod);
__global__ void k_nod_cell(int ncell, const int *corner, const int *target, double dx, double dy,
                           const double *nodval, double *cellval);
__global__ void k_gather_pairs(int m, const int *dst, const int *src, const double *from, double *to);
__global__ void k_cell_coarse(int ncell, const int *src, const double *cellval, double *coarse);
__global__ void k_cell_nod(int nnod, const int *ptr, const int *list, const double *coarse, double *pondnod);
__global__ void k_pondupd(int nnod, const int *surf_row, const int *ifatm, double pondh_min, double dtr,
                          const double *pondnod, const double *arenod, const double *atmpot, double *atmact,
                          double *pnew, int *ponding);
__global__ void k_pond_reset(int nnod, const int *surf_row, const double *pnew, double *pondnod);
__global__ void k_copy_surf(int nnod, const int *surf_row, const double *pnew, double *pold);
__global__ void k_put_double(double *x, double v);
__global__ void k_adrstn(int nnod, const int *surf_row, int *ifatm, const double *atmpot, double *atmact,
                         double *pnew, double pmin);
__global__ void k_put_int(int *x, int v);
__global__ void k_switch_old(int nnod, const int *surf_row, int *ifatm, double *atmact, const double *atmpot,
                             double *pnew, double pmin);
__global__ void k_next_iter(int n, double tetaf, const double *pnew, const double *ptimep, double *pold,
                            double *ptnew, double *ptold);
__global__ void k_start_iter(int n, double tetaf, const double *pnew, const double *ptimep, double *pold,
                             double *ptnew, double *ptold);
__global__ void k_atm_update(int nnod, double rate, const double *rate_node, const double *arenod,
                             const int *ifatm, double scf, int set_act, double *atmpot, double *atmact);
__global__ void k_free_drainage(int nnod, const int *bot_row, const double *arenod, const double *ckrw,
                                const double *kznod, double *qbot, double *qnode);
__global__ void k_store(int n, const double *sw, const double *volnod, const double *pnodi, double *part);
__global__ void k_sum_final(int nb, const double *part, double *out);
__global__ void k_gather_coef(int nslots, const int *map, const double *coef, double *A);
__global__ void k_check_codes(int n, Sell sl, int *bad);
__global__ void k_sumsq_free(int n, const double *u, const uint8_t *isdir, double *part);
__global__ void k_copy(size_t n, const double *src, double *dst);
__global__ void k_fill(size_t n, double v, double *dst);

// ------------------------------------------ batched ensemble launches --
// Everything one member's Picard iteration and IC(0)-PCG read or write,
// held in device memory per member (an array over the ensemble).  A batched
// kernel finds its member from the launch's group table and runs the
// single-member body on that member's arrays, so each member computes what
// its own launches would, bit for bit.
struct MemberDev {
  int n, nnod, nq;             // rows, surface nodes, base nodes with free drainage (0: none)
  int g_rows, g_red, g_flux;   // groups of the member's own row / reduction / flux launches
  int g_vec;                   // groups of its PCG vector launches (alpha, dot, beta)
  double tetaf, rdt, tol, pmin;
  int imax, ncol;
  Sell sl;
  SoilSoA soil;
  AsmArgs a;
  AsmPlan P;
  FacTab ft, fts;              // chain tables; fts: the SpMV's transposed reads (w null: off)
  ColDesc col[4];
  float *F32;
  double *F64;
  double *pnew, *pold, *ptnew, *ptold, *ptimep, *pdiff, *sw, *ckrw, *et1, *et2, *swnew, *swtimep;
  double *tnoti, *xt5, *bcflux, *qnode, *qbot, *atmact, *atmold, *atmpot;
  double *A, *dinv, *r, *p, *Bv, *z, *part, *res;
  uint8_t *isdir;
  int *ifatm, *flags, *ipart;
  const int *row_surf, *surf_row, *refid;
  const double *volnod, *snodi, *pnodi;
  PcgScal *scal;
};
// vectors a batched sweep reads / writes
enum SweepVec { kVecTnoti = 0, kVecPdiff, kVecR, kVecP, kVecZ };

// Batched launches: `start` (nmem + 1 group offsets, see batch_member)
// belongs to the launch shape of the kernel; `act` (per member, bits
// 1 switch_old, 2 next iteration) selects members for the last two.
__global__ void kb_constitutive(const int *start, int nmem, const MemberDev *md);
__global__ void kb_bc_rows(const int *start, int nmem, const MemberDev *md);
template <int MR>
__global__ void kb_assemble_plan(const int *start, int nmem, const MemberDev *md);
template <typename FT, int NW>
__global__ void kb_ic_factor_col3t(const int *start, int nmem, const MemberDev *md, int c);
template <typename FT, int NW, bool LAST>
__global__ void kb_fwd_col3(const int *start, int nmem, const MemberDev *md, int c, int vin, int vout,
                            int use_done);
template <typename FT, int NW>
__global__ void kb_bwd_col3(const int *start, int nmem, const MemberDev *md, int c, int vio, int use_done);
__global__ void kb_residual0(const int *start, int nmem, const MemberDev *md);
__global__ void kb_pcg_init(const int *start, int nmem, const MemberDev *md);
__global__ void kb_spmv_dot2(const int *start, int nmem, const MemberDev *md);
__global__ void kb_pcg_alpha(const int *start, int nmem, const MemberDev *md);
__global__ void kb_dot(const int *start, int nmem, const MemberDev *md);
__global__ void kb_pcg_beta(const int *start, int nmem, const MemberDev *md);
__global__ void kb_poll(const int *start, int nmem, const MemberDev *md, double *out);
__global__ void kb_update_pnew(const int *start, int nmem, const MemberDev *md);
__global__ void kb_bkpic(const int *start, int nmem, const MemberDev *md);
__global__ void kb_norms(const int *start, int nmem, const MemberDev *md);
__global__ void kb_flux(const int *start, int nmem, const MemberDev *md);
__global__ void kb_norms_final(const int *start, int nmem, const MemberDev *md);
__global__ void kb_switch_old(const int *start, int nmem, const MemberDev *md, const int *act);
__global__ void kb_next_iter(const int *start, int nmem, const MemberDev *md, const int *act);

// ------------------------------------------- Newton (newton_kernels.hip) --
struct NewtonArgs {
  const double *etai, *ckrw, *dckrw, *detai, *gz;
};
struct BicgScal {
  double xlung, roj, rojm1, alfj, alfjm1, etaj, etaj_next, err, tol;
  int niter, imax;
};
__global__ void k_constitutive_newton(int n, const double *ptnew, SoilSoA s, double *sw, double *ckrw,
                                      double *etai, double *dckrw, double *detai);
template <int MR>
__global__ void k_assemble_newton(AsmArgs a, AsmPlan P, NewtonArgs nw);
template <typename FT>
__global__ void k_ilu_factor_col3(ColDesc C, Sell sl, PairList pl, const int *tpos, const double *A,
                                  const uint8_t *isdir, FT *F, double *dinv, int *nfall);
template <typename FT>
__global__ void k_ilu_factor_list(ListLevel L, const int *rows, Sell sl, PairList pl, const int *tpos,
                                  const double *A, const uint8_t *isdir, FT *F, double *dinv, int *nfall);
template <typename FT>
__global__ void k_mul_upper(int n, Sell sl, const FT *F, const double *x, double *y);
__global__ void k_spmv(int n, Sell sl, const double *A, const double *x, double *y, const int *done);
__global__ void k_bicg_init(int nxl, const double *part_xl, int nrr, const double *part_rr, BicgScal *sc,
                            int *done, double tol, int imax);
__global__ void k_bicg_p(int n, const double *r, const double *v, double *p, const BicgScal *sc, const int *done);
__global__ void k_bicg_s(int n, int nparts, const double *part_sig, const double *r, const double *v, double *s,
                         BicgScal *sc, const int *done);
__global__ void k_bicg_st(int n, const double *s, const double *t, double *part_st, double *part_tt,
                          const int *done);
__global__ void k_bicg_update(int n, int nparts, const double *part_st, const double *part_tt, const double *p,
                              const double *s, const double *t, const double *rt, const uint8_t *isdir, double *y,
                              double *r, double *part_rr, double *part_rtr, BicgScal *sc, const int *done);
__global__ void k_bicg_check(int n, int nparts, const double *part_rr, const double *part_rtr, BicgScal *sc,
                             int *done);
__global__ void k_sw_only(int n, const double *psi, SoilSoA s, double *sw);
constexpr int kEnkfMax = 64;  // members per analysis launch
__global__ void k_enkf_update(int n, int nens, const double *const *X, const double *W, double alpha,
                              double *const *Xa);
// the localised analysis: states X/Xa (device order, one pointer per
// member), S and COV^-1 D (m x nens, row-major), device row -> CATHY row,
// and either the dense localisation (CATHY rows x m) or node / observation
// (x, y) positions with the Gaspari-Cohn radius
struct LocArgs {
  int n, nens, m;
  const double *const *X;
  double *const *Xa;
  const double *s, *invd;
  const int *iperm;
  const double *loc, *node_xy, *obs_xy;
  const int *row_zone, *obs_zone;  // zone localisation: L(i,j) = [zone of row i == zone of observation j]
  double radius, alpha;
};
template <int NE>
__global__ void k_enkf_localized(LocArgs a);
constexpr int kGramRows = 32;    // rows staged per work-group round
constexpr int kGramGroups = 512;  // work-groups (one partial Gram matrix each)
template <int NE>
__global__ void k_enkf_gram(int n, int nens, const double *const *X, double *part);
__global__ void k_enkf_gram_sum(int nb, int nn, const double *part, double *g);
template <int NE>
__global__ void k_pf_resample(int n, int nens, const double *const *src, const int *iperm, double sigma,
                              uint64_t seed, double *const *dst);
__global__ void k_vel3d(int nt, const int4 *enode, const double *pnew, const double *ckrw, const double *vt,
                        double *uu, double *vv, double *ww);
__global__ void k_vnod3d(int n, const int *inc_ptr, const int *inc, const double *uu, const double *vv,
                         const double *ww, double *unod, double *vnod, double *wnod);

}  // namespace cg
