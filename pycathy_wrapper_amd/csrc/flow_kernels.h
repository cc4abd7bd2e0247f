// flow_kernels.h -- declarations of the Picard/PCG kernels.
#pragma once

#include "kernels.h"

namespace cg {

constexpr int kAsmBlock = 128;  // rows per assembly work-group (LDS row buffers)

// One level of a triangular-sweep / factorisation schedule: `nchains`
// independent chains; chain t is the rows base + t + k*stride, k < len
// (strided form) or the single row lvl_rows[rows_off + t] (list form).
struct LevelDesc {
  int base, nchains, stride, len, rows_off;
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
  const int *rowptr, *col, *diagpos, *inc_ptr, *inc;
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
template <bool LUMPED>
__global__ void k_assemble(AsmArgs a);
template <typename FT>
__global__ void k_ic_factor(LevelDesc L, const int *lvl_rows, const int *rowptr, const int *col,
                            const int *diagpos, const int *tpos, const double *A, const uint8_t *isdir, FT *F,
                            double *dinv, int *nfall);
template <typename FT>
__global__ void k_sweep_fwd(LevelDesc L, const int *lvl_rows, const int *rowptr, const int *col,
                            const int *diagpos, const FT *F, const double *dinv, const double *tin, double *y,
                            const int *done);
template <typename FT>
__global__ void k_sweep_bwd(LevelDesc L, const int *lvl_rows, const int *rowptr, const int *col,
                            const int *diagpos, const FT *F, const double *dinv, double *x, const int *done);
__global__ void k_residual0(int n, const int *rowptr, const int *col, const double *A, const double *x,
                            const double *b, const uint8_t *isdir, double *r, double *part);
__global__ void k_spmv_dot2(int n, const int *rowptr, const int *col, const double *A, const double *p,
                            const double *r, double *B, double *part_pr, double *part_pb, const int *done);
__global__ void k_pcg_alpha(int n, int nparts, const double *part_pr, const double *part_pb, const double *B,
                            const double *p, const uint8_t *isdir, double *r, double *x, double *part_rr,
                            PcgScal *sc, const int *done);
__global__ void k_dot(int n, const double *u, const double *v, double *part, const int *done);
__global__ void k_pcg_beta(int n, int nparts, const double *part_bz, const double *part_rr, const double *z,
                           double *p, PcgScal *sc, int *done);
__global__ void k_pcg_init(int nparts, const double *part, PcgScal *sc, int *done, double tol, int imax);
__global__ void k_update_pnew(int n, const double *pdiff, const int *row_surf, const int *ifatm,
                              const double *pold, double *pnew);
__global__ void k_bkpic(int nnod, const int *surf_row, const int *rowptr, const int *col, const double *A,
                        const double *pdiff, const double *xt5, const int *ifatm, const double *atmold,
                        double tetaf, double *atmact);
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
__global__ void k_copy(size_t n, const double *src, double *dst);
__global__ void k_copy_int(int n, const int *src, int *dst);
__global__ void k_fill(size_t n, double v, double *dst);

}  // namespace cg
