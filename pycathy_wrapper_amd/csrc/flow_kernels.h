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
constexpr int kChains = 16;     // chains per work-group in the column kernels

struct Sell {
  const int *so;    // slice offsets
  const int *col;   // column id per slot
  const int *len;   // row lengths
  const int *diag;  // slot of the diagonal
};

__device__ __forceinline__ int sp(const int *so, int r, int j) { return so[r >> 6] + j * 64 + (r & 63); }

// One colour class of the column ordering: chains t < nc, chain t = rows
// base + t + k*nc for k < len (k = node plane, 0 = surface).
struct ColDesc {
  int base, nc, len;
};

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
  double tetaf, rdt;
  double *A, *tnoti, *xt5;
};

__global__ void k_constitutive(int n, const double *ptnew, const double *pnew, SoilSoA s, double *sw,
                               double *ckrw, double *et1, double *et2, double *swnew);
__global__ void k_sw_kr(int n, const double *psi_in, SoilSoA s, double *sw, double *ckrw);
__global__ void k_elem_avg(int nt, const int4 *enode, const double *ckrw, const double *et1, double *ckrwe,
                           double *et1e);
__global__ void k_bc_rows(int n, const int *row_surf, const double *qnode, const int *ifatm,
                          const double *atmact, const double *atmold, double tetaf, uint8_t *isdir,
                          double *bcflux);
template <int MR, bool LUMPED>
__global__ void k_assemble(AsmArgs a);

// Precomputed lumped-mass assembly plan.  Per row, one record per incident
// element (ascending element id), lane-interleaved per slice: record m of
// row r at off[r>>6] + 64 m + (r&63) holds the element id (-1 = padding),
// the row slots of its 4 nodes (4 bytes) and the element-matrix row
// S_e(k, 0..3) as a double4.  w2 / gw (SELL layout) give the lumped
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
template <typename FT>
__global__ void k_fwd_col3(ColDesc C, Sell sl, const FT *F, const double *dinv, const double *tin, double *y,
                           const int *done);
template <typename FT>
__global__ void k_bwd_col3(ColDesc C, Sell sl, const FT *F, const double *dinv, double *x, const int *done);
template <typename FT>
__global__ void k_ic_factor_list(ListLevel L, const int *rows, Sell sl, PairList pl, const int *tpos,
                                 const double *A, const uint8_t *isdir, FT *F, double *dinv, int *nfall);
template <typename FT>
__global__ void k_fwd_list(ListLevel L, const int *rows, Sell sl, const FT *F, const double *dinv,
                           const double *tin, double *y, const int *done);
template <typename FT>
__global__ void k_bwd_list(ListLevel L, const int *rows, Sell sl, const FT *F, const double *dinv, double *x,
                           const int *done);

__global__ void k_residual0(int n, Sell sl, const double *A, const double *x, const double *b,
                            const uint8_t *isdir, double *r, double *part);
__global__ void k_spmv_dot2(int n, Sell sl, const double *A, const double *p, const double *r, double *B,
                            double *part_pr, double *part_pb, const int *done);
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
__global__ void k_sumsq_free(int n, const double *u, const uint8_t *isdir, double *part);
__global__ void k_copy(size_t n, const double *src, double *dst);
__global__ void k_fill(size_t n, double v, double *dst);

}  // namespace cg
