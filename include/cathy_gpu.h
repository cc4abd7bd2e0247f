/*
 * cathy_gpu.h -- C ABI of the MI355X (gfx950) CATHY subsurface flow step.
 *
 * Every entry point takes its arguments BY REFERENCE, with Fortran's
 * conventions for the arrays the CATHY host already holds (1-based node
 * and entry ids, column-major 2-D arrays, REAL*8 / INTEGER*4), so that the
 * Fortran host binds them with a plain ISO_C_BINDING interface block and
 * passes its own arrays unchanged.  Return value: 0 on success, > 0 for a
 * solver outcome the host handles (linear-solver failure), < 0 for errors.
 * All calls are synchronous with respect to the host.
 *
 * Reference interfaces replaced (paths relative to the vendored CATHY
 * source, pyCATHY/tests/weil_exemple/my_cathy_prj/src/):
 *   cathy_gpu_symslv        SYMSLV   solscal-extended.f:4669-4699
 *   cathy_gpu_flow3d_*      FLOW3D   flow3d.f:8-294 (Picard: picard.f:8;
 *                           Newton: newton.f with NSYSLV ISOLV=0)
 *                           with the per-time-step host work of
 *                           cathy_main.f:2882-3860 (NEUMANN free drainage,
 *                           ATMNXT/ATMBAK, SWITCH_OLD, BKSTEP, CHVELO,
 *                           STORCAL, TIMUPD) available as one call.
 *   cathy_gpu_route_*       ROUTE    route.f:4-234 (MC mc.f, SERBATOIO
 *                           serbatoio.f, CALCOLI calcoli.f) followed by
 *                           ALTEZZE altezze.f:4-107, per routing substep
 */
#ifndef CATHY_GPU_H
#define CATHY_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CATHY_GPU_OK 0
#define CATHY_GPU_LSFAIL 1    /* PCG reached ITMXCG (picard.f:172-177) */
#define CATHY_GPU_EINVAL (-1) /* bad argument / inconsistent sizes */
#define CATHY_GPU_EDEVICE (-2) /* HIP runtime error */
#define CATHY_GPU_ENOMEM (-3)
#define CATHY_GPU_EHANDLE (-4) /* unknown or destroyed handle */

/* preconditioner orderings (cathy_gpu_flow3d_create, *ordering) */
#define CATHY_GPU_ORDER_NATURAL 0 /* the reference's IC(0): exact INCLDP order */
#define CATHY_GPU_ORDER_COLUMN3 1 /* IC(0) in 3-colour column order (fast) */
#define CATHY_GPU_ORDER_COLUMN3C 2 /* the same classes, each soil column's rows contiguous
                                    (chain-major; same results to rounding) */
#define CATHY_GPU_ORDER_COLUMN3T 3 /* the same classes, rows in tiles of 8 soil columns,
                                    each tile plane by plane (same results to
                                    rounding; COLUMN3 is the fastest) */

/* Human-readable text of the last error on this thread. */
const char *cathy_gpu_last_error(void);

/* Device 0 name and compute-unit count; returns <0 without a GPU. */
int cathy_gpu_device_info(char *name, int *name_len, int *cu_count);

/* ------------------------------------------------------------------ *
 * SYMSLV (solscal-extended.f:4669-4699): solve COEF1 * PNEW = TNOTI by
 * IC(0)-preconditioned CG as INCLDP (:2142-2193) + PRODDP (:4169-4204) +
 * GRADDP (:1260-1386) + RESSYM (:4389-4438) do: natural-order FP64
 * factor with the breakdown fallback, x0 = M^-1 b, stop when
 * ||r_free|| / ||b_free|| <= TOLCG or NITER >= ITMXCG.
 * TOPOL(N+1) / JA(NTERM) / COEF1(NTERM): the reference's 1-based upper
 * CSR with the diagonal first in each row (strpic.f).  NODDIR(1:NUMDIR):
 * Dirichlet rows.  The reference's 1e-9*RMAX penalty on their diagonal
 * decouples them to ~1e-90 relative; here they are masked out of the
 * factor and the products, and PNEW = 0 on them.  FP64 throughout.
 * Outputs: PNEW, NITER, ERR (GRADDP's last relative residual), RES
 * (RESSYM's RMS residual: over ||b_free|| unless that is <= RMIN).
 * The pattern set-up is kept per calling thread and reused while N,
 * NTERM, TOPOL and JA are unchanged (the Picard loop's case).
 * Returns CATHY_GPU_LSFAIL when NITER >= ITMXCG.
 * ------------------------------------------------------------------ */
int cathy_gpu_symslv(const int *n, const int *nterm, const int *numdir,
                     int *niter, const int *itmxcg, const double *tolcg,
                     const double *rmin, const int *noddir, const int *ja,
                     const int *topol, double *pnew, const double *tnoti,
                     const double *coef1, double *err, double *res);

/* Frees the calling thread's SYMSLV set-up (no reference counterpart). */
int cathy_gpu_symslv_release(void);

/* ------------------------------------------------------------------ *
 * FLOW3D family.  Set-up takes the mesh and soil arrays the host builds
 * once (GRDSYS / TPNODI / CHPARM outputs, cathy_main.f:2559-2704):
 *   TETRA(5,NT)  1-based sorted vertices + zone      (gen3d.f, riord.f)
 *   BI,CI,DI(4,NT) P1 coefficients /6, VOLU(NT), IVOL(NT)   (volbas.f)
 *   VOLNOD(N), ARENOD(NNOD)                          (volbas.f, area2d.f)
 *   PERMX/Y/Z(MAXSTR,NZONE), ELSTOR/POROS only through SNODI/PNODI(N)
 *   VGN, VGRMC, VGPSAT(N) node vG parameters            (tpnodi.f)
 *   KZNOD(N)                                  (cathy_main.f:2687-2694)
 * with CATHY's node numbering (node = layer*NNOD + surface id).
 * LUMP as in input/parm.  ORDERING: CATHY_GPU_ORDER_*.
 * FP64 state; IC(0) factor stored in FP32.
 * ------------------------------------------------------------------ */
int cathy_gpu_flow3d_create(int64_t *handle, const int *n, const int *nt,
                            const int *ntri, const int *nnod,
                            const int *nstr, const int *nzone,
                            const int *maxstr, const int *tetra,
                            const double *bi, const double *ci,
                            const double *di, const double *volu,
                            const int *ivol, const double *volnod,
                            const double *arenod, const double *permx,
                            const double *permy, const double *permz,
                            const double *snodi, const double *pnodi,
                            const double *vgn, const double *vgrmc,
                            const double *vgpsat, const double *kznod,
                            const int *lump, const int *ordering);

int cathy_gpu_flow3d_destroy(int64_t *handle);

/* input/parm and soil-file scalars (datin.f:80-122, :421).
 * IOPT 1 = Picard + IC(0)-PCG, 2 = Newton + ILU(0)-BiCGSTAB (LUMP=1).
 * iprm: [0] IOPT [1] L2NORM [2] ITUNS [3] ITUNS1 [4] ITUNS2 [5] ITMXCG
 *       [6] IETO [7] free drainage at the base (NNEU<0) [8] HSPATM
 * dprm: [0] TETAF [1] TOLUNS [2] TOLSWI [3] ERNLMX [4] TOLCG [5] DELTAT
 *       [6] DTMIN [7] DTMAX [8] TMAX [9] DTMAGA [10] DTMAGM [11] DTREDS
 *       [12] DTREDM [13] PMIN [14] PONDH_MIN [15] SCF */
int cathy_gpu_flow3d_set_params(int64_t *handle, const int *iprm,
                                const double *dprm);

/* Atmospheric series (the atmbc file read by ATMONE/ATMNXT): NATM
 * records at ATMTIM(NATM), values ATMVAL(NATM) if HSPATM=1 else
 * ATMVAL(NNOD,NATM) (m/s, +ve rain).  The record width is the one of the
 * HSPATM in force at this call, so set_params comes first; after a
 * set_params that changes HSPATM, init and advance refuse the series until
 * it is set again. */
int cathy_gpu_flow3d_set_forcing(int64_t *handle, const int *natm,
                                 const double *atmtim, const double *atmval);

/* Initial condition: PTIMEP(N) (INDP / ICV*) then the INITAL sequence
 * (INIT1, ATMONE, MBINIT, CHVELO, STORCAL, NEUMANN) on the device. */
int cathy_gpu_flow3d_init(int64_t *handle, const double *ptimep);

/* One accepted time step of cathy_main.f's loop (back-steps included)
 * with FLOW3D's Picard iterations on the device.  stats(32) receives:
 *  [0] NSTEP [1] DELTAT [2] TIME [3] ITER [4] NITERT [5] back-steps
 *  [6] STORE1 [7] STORE2 [8] DSTORE [9] CDSTOR [10] VIN [11] CVIN
 *  [12] VOUT [13] CVOUT [14] VIN+VOUT [15] VTOT [16] ERRAS [17] ERREL
 *  [18] CERRAS [19] CAERAS [20] PINF [21] ITRTOT [22] ITLIN [23] KLSFAI
 *  [24] PL2 [25] FINF [26] FL2 [27] IKMAX (1-based) [28] wall-clock ms
 *  of the step (host clock) [29] Picard and [30] PCG iterations of the
 *  step, back-stepped attempts included [31] PONDING at the step's end
 *  (coupled surface, cathy_gpu_flow3d_set_surface)
 * Returns 0 while the run continues, 1 at TMAX, 2 on NOBACK, 3 when a
 * coupled run's step ends with a ponded surface node (the next step would
 * route it first, which the device does not do; the state is the step's
 * end), <0 on error. */
int cathy_gpu_flow3d_advance(int64_t *handle, double *stats);

/* State read-back / overwrite (CATHY node order): PNEW, SW, CKRW (N),
 * IFATM (NNOD).  Any pointer may be NULL. */
int cathy_gpu_flow3d_get_state(int64_t *handle, double *pnew, double *sw,
                               double *ckrw, int *ifatm);
int cathy_gpu_flow3d_set_pnew(int64_t *handle, const double *pnew);

/* ---------------------------------------------------------------------
 * Learned Picard initial iterate (no counterpart in the reference).
 *
 * A per-column 1-D convolutional network (bf16 weights and activations,
 * fp32 accumulation on MFMA) corrects the extrapolated increment of each
 * soil column.  Per level it sees (standardised): psi_n, theta_n, log10 Ks,
 * alpha, n, rainfall rate, log10 dt, depth below the surface and
 * E = dt_n (psi_n - psi_{n-1}) / dt_{n-1}; each attempt after the first
 * step starts its Picard loop from
 *   psi_n + beta E + s_col * out * clamp(y, -10, 10),
 *   s_col = rms_col(E) + floor
 * instead of psi_n.  The Picard loop, its tolerance and its convergence
 * test are unchanged.
 *
 * Network entries (host buffers, synchronous):
 *   create   levels = NSTR+1 per column; *nparam receives 40385
 *   set/get  params (fp32, per layer W[k][ci][co] then b[co]; layers
 *            16->64, 64->64 x3, 64->1, kernel 3) and scale[36] = feature
 *            means[16], standard deviations[16], out, floor (m), beta,
 *            band (m; >= 0: nodes with psi_n or the guess above -band
 *            keep psi_n; < 0: off);
 *            either pointer may be NULL
 *   forward  x: bf16 bits [ncol][Lp][16] (Lp = levels rounded up to 16,
 *            padded levels zero) -> y: fp32 [ncol][Lp]
 *   gradient mean squared error against t [ncol][Lp] over live levels:
 *            loss and its gradient [nparam]
 *   load     a training set of nsamp columns (x as forward, t as gradient)
 *   train    nsteps Adam steps (beta 0.9/0.999, eps 1e-8) on batches of
 *            the loaded set; idx [nsteps][batch] sample numbers; loss[nsteps]
 * A network may be attached to several flow handles, also ones advanced
 * from different threads: its calls and evaluations are serialised.
 * ------------------------------------------------------------------- */
int cathy_gpu_guess_create(int64_t *net, const int *levels, int *nparam);
int cathy_gpu_guess_destroy(int64_t *net);
int cathy_gpu_guess_set(int64_t *net, const float *params, const float *scale);
int cathy_gpu_guess_get(int64_t *net, float *params, float *scale);
int cathy_gpu_guess_forward(int64_t *net, const int *ncol, const uint16_t *x,
                            float *y);
int cathy_gpu_guess_gradient(int64_t *net, const int *ncol, const uint16_t *x,
                             const float *t, float *loss, float *grad);
int cathy_gpu_guess_load(int64_t *net, const int64_t *nsamp,
                         const uint16_t *x, const float *t);
int cathy_gpu_guess_train(int64_t *net, const int *nsteps, const int *batch,
                          const int *idx, const float *lr, float *loss);

/* Replay of step sizes (no counterpart in the reference; for comparing
 * runs step by step): after init and before the first advance, dts[0] is
 * step 1's DELTAT and dts[k] the first attempt's DELTAT of step k+1 in
 * place of TIMUPD's rule (back-steps still reduce it; after nd entries the
 * rule resumes). */
int cathy_gpu_flow3d_set_dt_schedule(int64_t *handle, const int *nd,
                                     const double *dts);

/* Attach a network to a flow run (*net = 0 detaches; net NULL leaves it)
 * and/or set the static column features stat[N][5] in CATHY node order:
 * log10 Ks, alpha (1/VGPSAT), VGN, depth below the surface, porosity. */
int cathy_gpu_flow3d_guess(int64_t *handle, const int64_t *net,
                           const float *stat);
/* Recording: with *on != 0 every attempt keeps its raw features and every
 * accepted step its increment PNEW - PTIMEP; guess_sample returns those of
 * the last accepted step: feat [NNOD][Lp][16], target [NNOD][Lp]. */
int cathy_gpu_flow3d_guess_record(int64_t *handle, const int *on);
int cathy_gpu_flow3d_guess_sample(int64_t *handle, float *feat,
                                  float *target);

/* EnKF analysis of the member states in place (pyCATHY DA/enkf.py:16-224,
 * enkf_analysis; state inflation as in enkf_analysis_localized_with_
 * inflation, enkf.py:316-318; no localisation).  handles: nens flow
 * handles on one mesh and ordering (2 <= nens <= 64); w: the nens x nens
 * weight W(k,j) at w[j*nens + k] such that P_xo COV^-1 D = X' W, i.e.
 * W = S^T COV^-1 D / (nens-1) from the observation-space quantities (the
 * Sakov form replaces COV^-1 D by D / diag(R)); *inflation: 1 = none.
 * PNEW of every member becomes x_j + sum_k (x_k - xbar) W(k,j) (then
 * inflated about the analysis mean), and each run continues from it as
 * from an accepted step (cathy_gpu_flow3d_adopt_state). */
int cathy_gpu_enkf_update(const int *nens, int64_t *handles, const double *w,
                          const double *inflation);
/* The same analysis without allocating device memory.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  W and
 * the member state pointer tables are staged in one fixed device array per
 * GPU (kScratchBytes = 33,792 bytes: 64 x 64 FP64 weights + 2 x 64
 * pointers), held until the process ends; NENS above 64 is refused.  Calls
 * that overlap on one GPU share that array and race: serialise them.
 * Results are those of cathy_gpu_enkf_update (the same kernel). */
int cathy_gpu_enkf_update_noalloc(const int *nens, int64_t *handles,
                                  const double *w, const double *inflation);
/* Localised EnKF analysis of the member states in place (pyCATHY DA/enkf.py:
 * 225-343, enkf_analysis_localized_with_inflation: P_xo[states] * L before
 * the update, :307-313; state inflation :318-320).  handles: nens flow
 * handles on one mesh and ordering (2 <= nens <= 64); *nobs = m
 * observations; s: S = HX - mean(HX), s[j*nens + k]; invd: COV^-1 D (or
 * D / diag(R) in the Sakov form), same layout.  The localisation is
 * either loc, the dense L(i, j) at loc[i*m + j] over the N state rows in
 * CATHY order (the reference's L argument), or, with loc NULL, the
 * Gaspari-Cohn weights of DA/localisation.py:136-188 formed on the device
 * from node_xy (N x 2, CATHY order), obs_xy (m x 2) and *radius.  Member
 * x_k of each row becomes x_k + sum_j (sum_l x'_l S(j,l))/(nens-1) L(i,j)
 * invD(j,k), then inflation about the analysis mean (*inflation, 1 =
 * none); each run continues from it (cathy_gpu_flow3d_adopt_state).  The
 * augmented state's parameters are the caller's (unlocalised, as in the
 * reference). */
int cathy_gpu_enkf_update_localized(const int *nens, int64_t *handles,
                                    const int *nobs, const double *s,
                                    const double *invd, const double *loc,
                                    const double *node_xy,
                                    const double *obs_xy,
                                    const double *radius,
                                    const double *inflation);
/* Domain-localised analysis (pyCATHY DA/cathy_DA.py:671-803 with
 * localisation 'zones' / 'veg_map' / 'nodes', DA/localisation.py:12-48):
 * one analysis per zone, each over the state rows of its zone with the
 * observations in it, all in one launch.  s and invd as for
 * cathy_gpu_enkf_update_localized, each observation's rows formed from its
 * own zone's observations (COV of that zone alone); row_zone: the zone of
 * each of the N state rows (CATHY order), obs_zone: the zone of each of
 * the m observations.  Row i moves by the observations j with obs_zone[j]
 * == row_zone[i] only; rows of a zone without observations keep their
 * states. */
int cathy_gpu_enkf_update_zones(const int *nens, int64_t *handles,
                                const int *nobs, const double *s,
                                const double *invd, const int *row_zone,
                                const int *obs_zone, const double *inflation);
/* Gram matrix of the member state perturbations, gram[i*nens + j] =
 * sum over the N state rows of (x_i - xbar)(x_j - xbar), xbar the row's
 * member mean: X'^T X', the N_ens x N_ens matrix through which the dual
 * analysis's second step (pyCATHY DA/enkf.py:345-361, the states as
 * observations of themselves) reduces to member-space algebra.  Fixed
 * summation order (deterministic); states are not changed. */
int cathy_gpu_enkf_gram(const int *nens, int64_t *handles, double *gram);
/* Particle-filter resampling of the member states in place (pyCATHY
 * DA/pf.py:98-109 ensemble[:, indices], then the state jitter of
 * :129-130).  idx: nens 0-based member indices (systematic_resample's
 * output; the identity to jitter only); member j's PNEW becomes member
 * idx[j]'s, plus *jitter_std times a standard normal drawn per (CATHY row,
 * member j) from Philox4x32-10 under *seed with the Box-Muller transform
 * (0: exact copies).  Each run continues from its new state. */
int cathy_gpu_pf_resample(const int *nens, int64_t *handles, const int *idx,
                          const double *jitter_std, const int64_t *seed);
/* Continue a run from its current PNEW (after set_pnew or an analysis):
 * PTIMEP = PNEW, Sw and kr at it, storage baseline reset. */
int cathy_gpu_flow3d_adopt_state(int64_t *handle);

/* VEL3D (vel3d.f) + VNOD3D (vnod3d.f), cathy_main.f:3243-3246: Darcy
 * velocities UU/VV/WW(NT) per element and their nodal means UNOD/VNOD/
 * WNOD(N) at the current PNEW and kr.  TETRA, PERMX/Y/Z(MAXSTR,NZONE),
 * BI/CI/DI(4,NT), VOLUR(NT), IVOL(NT) are read on the first call only and
 * kept on the device.  Any output pointer may be NULL. */
int cathy_gpu_flow3d_vel3d(int64_t *handle, const int *tetra,
                           const double *permx, const double *permy,
                           const double *permz, const double *bi,
                           const double *ci, const double *di,
                           const double *volur, const int *ivol, double *uu,
                           double *vv, double *ww, double *unod, double *vnod,
                           double *wnod);

/* IC(0) factor precision.  The reference factors in REAL*8 (INCLDP,
 * solscal-extended.f:2142-2193); the natural ordering does so always.  In
 * the column ordering the factor is FP32 unless ON != 0 selects FP64 (same
 * elimination order, twice the factor's bytes).  Call between steps; the
 * next solve refactors. */
int cathy_gpu_flow3d_set_factor_fp64(int64_t *handle, const int *on);

/* PCG SpMV variant.  No counterpart in the reference (its products read
 * the upper CSR once, solscal-extended.f:1326-1335).  ON != 0: each row
 * reads its diagonal and upper entries and takes the lower ones from the
 * transposed slots (the assembled matrix is bitwise symmetric), so HBM
 * streams about one triangle; results are bitwise the same.  Faster for one
 * member alone on the GPU, slower when several members share it; default 0.
 * ACTIVE (may be NULL) receives 1 if the variant applies to this mesh
 * (column ordering, per-chain tables valid, rows of at most 16 entries). */
int cathy_gpu_flow3d_set_spmv_transposed(int64_t *handle, const int *on,
                                         int *active);

/* PCG launch batches (no counterpart in the reference, whose PCG loop is
 * a host loop, solscal-extended.f:1325-1377).  The device enqueues PCG
 * iterations in batches and polls its convergence flag between them; the
 * first batch covers FIRST_PCT % of the previous solve's iterations (8 before
 * any), later ones FOLLOWUP iterations each.  Launches past convergence exit
 * at once, so the results do not depend on these; defaults 100 and 2. */
int cathy_gpu_flow3d_set_pcg_batches(int64_t *handle, const int *first_pct,
                                     const int *followup);

/* Non-atmospheric, non-seepage-face boundary series; the run starts over
 * at the next init.
 * KIND 0: fixed-head (Dirichlet) nodes with their heads PRESC (m), the
 * nansfdirbc series; KIND 1: Neumann nodes with their fluxes Q (m^3/s,
 * > 0 into the domain), the nansfneubc series.  Record k (NREC of them,
 * TIMES not decreasing, the first at 0) lists COUNTS(k) 1-based NODES and
 * their VALUES, expanded as RDNDBC/READBC expand a record (surface nodes
 * replicated down the layers); for KIND 1 under free drainage only the
 * entries after the NNOD base nodes.  Replaces, per time step:
 *   CALL BCONE / BCNXT / BCBAK('...',IIN8|IIN9,...,NP|NQ,...,PTIM|QTIM,
 *        PINP|QINP,CONTP|CONTQ,PRESC|Q,ANP|ANQ,ACONTP|ACONTQ)
 *     (bcone.f:1-4, bcnxt.f:5-8, bcbak.f:4-5; cathy_main.f:2892-2903,
 *      bkstep.f:73-88) and NEUMANN's listed part (neumann.f:25-31),
 * and feeds BCPIC (bcpic.f:33-39 fixed heads, :71-74 Neumann fluxes),
 * SHLPIC (shlpic.f:30-35), BKPIC (bkpic.f:38-41) and FLUXMB's NDIN/NDOUT
 * and NNIN/NNOUT (fluxmb.f:30-46); listed surface nodes leave the
 * atmospheric boundary (IFATM = -1, atmnxt.f:80-102).  A node may appear
 * once per fixed-head record.  Picard scheme (IOPT = 1) only. */
int cathy_gpu_flow3d_set_bc(int64_t *handle, const int *kind, const int *nrec,
                            const double *times, const int *counts,
                            const int *nodes, const double *values);

/* Seepage faces (the sfbc file, time-invariant): NSF faces, face i holding
 * COUNTS(i) 1-based NODES listed top to bottom; ISFCVG = 1 adds the exit
 * points' convergence to the Picard test (inital.f:218-224, SFCHEK).  A
 * node with head >= 1e-8 m at the start is an exit node held at 0
 * (sfinit.f:22-45); exit nodes are Dirichlet rows of BCPIC
 * (bcpic.f:41-53) reset to 0 by SHLPIC (shlpic.f:36-45), their flux SFQ is
 * back-calculated by BKPIC (bkpic.f:42-49) and leaves through VSFFLW
 * (fluxmb.f:53-62, masbal.f:69-73); after every Picard iteration EXTALL /
 * EXTCVG (extall.f:22-47, extcvg.f:20-26) move the exit points, and
 * back-steps restore them (bkstep.f:56-67).  Replaces, with DUPUIT = 0 and
 * ISFONE = 0, the seepage-face arguments NSF, NSFNUM, NSFNOD, SFEX, SFEXIT,
 * SFEXP, SFQ, SFQP, LHSSF, SFCHEK, KSFZER of FLOW3D (flow3d.f:8-35).  The
 * run starts over at the next init; Picard scheme only. */
int cathy_gpu_flow3d_set_seepage(int64_t *handle, const int *nsf,
                                 const int *counts, const int *nodes,
                                 const int *isfcvg);

/* The seepage faces' SFEX (1 = exit node) and SFQ (m^3/s) by node
 * position (face by face, top to bottom); null outputs are skipped.
 * Returns the number of seepage nodes. */
int cathy_gpu_flow3d_get_seepage(int64_t *handle, int *sfex, double *sfq);

/* Root water uptake (QTRANIE) of the Feddes model, recomputed at the start
 * of every time step.  VEG_TYPE(NNOD) 1-based types of the surface nodes;
 * per type PCANA, PCREF, PCWLT, ZROOT (> 0), PZ, OMGC (the soil file's
 * vegetation lines, datin.f:425-428); Z(N) node elevations.  Every root
 * zone must end above its column's base.  Replaces
 *   CALL ETRAN(N,NNOD,NSTR,ATMPOT,Z,PNEW,PNODI,VEG_TYPE,QTRANIE)
 *     (etran.f:1; cathy_main.f:2929)
 * whose QTRANIE BCPIC subtracts from every row (bcpic.f:84-86).  Without
 * it QTRANIE is zero, which is what ETRAN gives for ATMPOT >= 0.  The run
 * starts over at the next init; Picard scheme only. */
/* The coupled surface of ISIMGR = 2 (SURF = 1) or the subsurface-only
 * run (0, the default).  With SURF = 1 the atmospheric switching is the
 * coupled one: CONVER calls SWITCH instead of SWITCH_OLD (conver.f:61-71,
 * switch.f) and step starts and back-steps call ADRSTN (cathy_main.f:
 * 2942-2946, bkstep.f:135-139); PONDH_MIN (set_params) is the parm file's,
 * which ATMONE and SWITCH test.  Replaces, for a surface that never ponds,
 *   CALL SWITCH(NNOD,IFATM,PONDING,TIME,DELTAT,PONDH_MIN,ARENOD,PONDNOD,
 *               ATMPOT,ATMACT,QTRANIE,PNEW,OVFLNOD)        (conver.f:67)
 *   CALL ADRSTN(NNOD,IFATM,ATMPOT,ATMACT,PNEW)    (cathy_main.f:2945)
 * The routing (SURF_FLOWTRA, PONDUPD) is not run: PONDNOD stays zero, and
 * advance returns 3 after a step that ends with PONDING set.  The run
 * starts over at the next init. */
int cathy_gpu_flow3d_set_surface(int64_t *handle, const int *surf);

/* The routing block of a coupled run (SURF = 1 with ponding): links a
 * routing handle (cathy_gpu_route_create: the project's drainage network,
 * no reservoirs, ALTEZZE cell types 0) to the flow handle.  Each attempt
 * that starts with PONDING set then runs, on the device, what
 * cathy_main.f:2993-3058 runs before FLOW3D:
 *   CALL SURF_FLOWTRA(NCELL,NNOD,NROW,NCOL,NTRI,DOSTEP,NUMRES,...,
 *                     TIME,DELTAT,ARENOD,OVFLNOD,PONDNOD,...)
 *     (surf_flowtra.f:9-16: NOD_CELL, TRANSFER_F3D_SURF, ROUTE + ALTEZZE
 *      substeps of CU <= 1, TRANSFER_SURF_F3D, CELL_NOD)
 *   CALL PONDUPD(NNOD,PONDING,IFATM,PONDH_MIN,DELTAT,PONDNOD,ARENOD,
 *                ATMPOT,ATMACT,PNEW)                    (pondupd.f:5)
 * with the routing state saved for back-steps and advanced at each
 * accepted step (cathy_main.f:3145-3157, 3770-3782; bkstep.f:148-157), and
 * SWITCH takes the ponded heads into account; advance no longer stops at
 * a ponded step.  DATIN's arrays (datin.f:6-20), 1-based: CELL(5,
 * NCELL_COARSE), INDCEL and INDCELWL (INDEX / INDEX_WITH_LAKES) and
 * DEM_MAP with element (I, J) at (I-1) + (J-1)*NROW (I the DEM row from
 * the north), TRIANG(4, NTRI).  DOSTEP = 1, no lakes.  The run starts
 * over at the next init. */
int cathy_gpu_flow3d_set_routing(int64_t *handle, const int64_t *route,
                                 const int *nrow, const int *ncol,
                                 const double *delta_x, const double *delta_y,
                                 const int *ncell_coarse, const int *cell,
                                 const int *indcel, const int *indcelwl,
                                 const double *dem_map, const int *ntri,
                                 const int *triang);

/* Initial ponding heads PONDNOD(NNOD) of a coupled run (IPOND = 1 or 2;
 * datin.f:394-403 reads them and sets PTIMEP = PONDNOD at ponded surface
 * nodes, which the caller's initial head carries).  With them the run starts
 * with PONDING set (inital.f:201-207), so its first step routes; they need
 * set_surface and set_routing.  Applied at every init; NULL clears them. */
int cathy_gpu_flow3d_set_initial_ponding(int64_t *handle, const double *pondnod);

/* The routing block's state: PONDNOD(NNOD) and OVFLNOD(NNOD) by surface
 * node, H_WATER_KKP1_SN(NROW*NCOL) by I_BASIN - 1 (null outputs skipped);
 * info[2]: routing blocks run, PONDING at the last step's end. */
int cathy_gpu_flow3d_routing_state(int64_t *handle, double *pondnod,
                                   double *ovflnod, double *h_water,
                                   int *info);

int cathy_gpu_flow3d_set_vegetation(int64_t *handle, const int *nveg,
                                    const int *veg_type, const double *pcana,
                                    const double *pcref, const double *pcwlt,
                                    const double *zroot, const double *pz,
                                    const double *omgc, const double *z);

/* Measurement only (no counterpart in the reference): the PCG SpMV on the
 * run's current matrix and search direction, NREP launches each after a
 * FLUSH_MB copy that evicts the caches; OUT2 = mean ms of those launches
 * and of a second, warm launch right after each.  The run's state is kept
 * (the SpMV's outputs are PCG scratch). */
int cathy_gpu_flow3d_bench_spmv(int64_t *handle, const int *nrep,
                                const double *flush_mb, double *out2);

/* Counters of the last advance: Picard its, PCG its, IC(0) breakdown
 * fallbacks, kernel time of the linear solver (ms), total device ms. */
int cathy_gpu_flow3d_counters(int64_t *handle, double *out6);

/* ------------------------------------------------------------------ *
 * Surface routing substep (ISIMGR = 2): ROUTE + ALTEZZE.
 * Replaces, per routing substep of SURF_ROUTE / cathy_main.f:3060-3068:
 *   CALL ROUTE(NROW,NCELL,TIPO_R,RESERVR,CELLS_R,N_HA,TIME,DELTAT)
 *     (route.f:4-5; MC mc.f:4-5)
 *   CALL ALTEZZE(NCELL,TIPO_R,DELTA_X,DELTA_Y,Q_IN_KK_SN,...,ELEVATION)
 *     (altezze.f:4-9)
 * with SERBATOIO / CALCOLI for reservoirs (serbatoio.f:6-7, calcoli.f:4-6)
 * once cathy_gpu_route_set_reservoirs is called.  The reference reads
 * the network from COMMON /RIVERNETWORK/ and the flows from /SURFWATER/;
 * here they are arguments.
 * ------------------------------------------------------------------ */
/* The drainage network: QOI_SN(NCELL) (1-based I_BASIN = (I-1)*NROW + J,
 * routing order), and the DTM_* grids of RIVERNETWORK.H, element (I, J)
 * at (I-1) + (J-1)*LDC (LDC = COLMAX >= NCOL): P_OUTFLOW_1/2, W_1/2,
 * LOCAL_SLOPE_1/2, EPL_1/2, KSS1_SF_1/2, WS1_SF_1/2, B1_SF, Y1_SF, NRC.
 * Every routed branch (W /= 0; the last cell's cardinal branch excepted)
 * must drain to a grid cell listed later in QOI_SN, or to one QOI_SN does
 * not list (a no-data cell: its inflow then enters only ALTEZZE), else -1.
 * The downstream I_BASIN is formed as route.f:86-88 forms it, so a branch
 * past the grid's edge names the grid cell that index falls on. */
int cathy_gpu_route_create(int64_t *handle, const int *nrow, const int *ncol,
                           const int *ldc, const int *ncell, const int *qoi_sn,
                           const int *p_outflow_1, const int *p_outflow_2,
                           const double *w_1, const double *w_2,
                           const double *local_slope_1, const double *local_slope_2,
                           const double *epl_1, const double *epl_2,
                           const double *kss1_sf_1, const double *kss1_sf_2,
                           const double *ws1_sf_1, const double *ws1_sf_2,
                           const double *b1_sf, const double *y1_sf, const double *nrc);
/* Reservoirs (optional; without this call every TIPO_R is taken as 0):
 * TIPO_R, RESERVR and ELEVATION (NROW*NCOL, by I_BASIN), NUMRES <= MAXRES,
 * N_HA(NUMRES) in 2..30, HRES / ARES (MAXRES, 30), H_SOGLIA, H_FONDO
 * (NUMRES) as in SURFWATER.H and cathy_main.f:2312-2313.  Refused (-1) where
 * the reference stops: a routing cell with TIPO_R >= 1000 (route.f:138-141).
 * A reservoir's new level is the last of its outlet branches in QOI_SN order,
 * as the reference's sequential writes leave it. */
int cathy_gpu_route_set_reservoirs(int64_t *handle, const int *tipo_r, const int *reservr,
                                   const int *numres, const int *maxres, const int *n_ha,
                                   const double *hres, const double *ares,
                                   const double *h_soglia, const double *h_fondo,
                                   const double *elevation);
/* ALTEZZE's cell-type array (its second argument, altezze.f:11,33-105):
 * CELTYPE(NROW*NCOL) by I_BASIN; 0 takes the volume balance, 1..999 the
 * pool depth of reservoir NUM_R(c) = RESERVR(c), >= 1000 the lake depth of
 * the reservoir of cell CELTYPE-1000.  Default, and what the reference's
 * time loop passes (cathy_main.f:3062, surf_flowtra.f:108, where CELTYPE is
 * never assigned): 0 everywhere.  Non-zero types need set_reservoirs first
 * (set_reservoirs clears them). */
int cathy_gpu_route_set_celtype(int64_t *handle, const int *celtype);
/* info(4): levels of the network, most feeders of one cell, NCELL,
 * NROW*NCOL. */
int cathy_gpu_route_info(int64_t *handle, int *info);
/* One substep for NMEMBER members sharing the network; every array is
 * (NROW*NCOL, NMEMBER) indexed by I_BASIN (SURFWATER.H names).
 * In: SURFACE_WATER_SN, Q_IN_KK_SN, Q_OUT_KK_SN_1/2, VOLUME_KK_SN.
 * In/out: Q_IN_KKP1_SN (receives the upstream outflows), Q_OUT_KKP1_SN_1/2
 * (a branch with W = 0 keeps its value), VOLUME_KKP1_SN (cells of non-zero
 * ALTEZZE type keep theirs).  Out: H_WATER_KKP1_SN.  cu_ak(2, NMEMBER): CU_MAX
 * (out), AK_MAX (in/out: kept when no branch is routed).  With reservoirs:
 * H_POOL_KK_VEC (in) and H_POOL_KKP1_VEC (in/out), (NUMRES, NMEMBER); else
 * both may be NULL.  Returns -1 where CALCOLI stops (a pool level at or
 * below its table).  Synchronous; calls on one handle from several threads
 * run one after another. */
int cathy_gpu_route_step(int64_t *handle, const int *nmember, const double *deltat,
                         const double *delta_x, const double *delta_y,
                         const double *surface_water_sn, const double *q_in_kk_sn,
                         const double *q_out_kk_sn_1, const double *q_out_kk_sn_2,
                         const double *volume_kk_sn, double *q_in_kkp1_sn,
                         double *q_out_kkp1_sn_1, double *q_out_kkp1_sn_2,
                         double *volume_kkp1_sn, double *h_water_kkp1_sn, double *cu_ak,
                         const double *h_pool_kk_vec, double *h_pool_kkp1_vec);
int cathy_gpu_route_destroy(int64_t *handle);

/* ------------------------------------------------------------------ *
 * Lockstep ensemble.  Replaces the per-member ./cathy processes of
 * pyCATHY's data-assimilation forecast (pyCATHY/DA/cathy_DA.py:1373-1390:
 * multiprocessing.Pool(cpu_count()-5).map over subprocess_run_multi,
 * cathy_tools.py:76-98) by NMEMBER flow handles advanced together on one
 * GPU.  Each round, every member still iterating takes one Picard
 * iteration, and every kernel of it is ONE launch over all members; the PCG
 * keeps a convergence mask per member (a converged member's groups exit at
 * once), and each member decides on its own when to iterate, back-step or
 * end its step, so members of uneven mesh size and iteration counts share
 * the launches without waiting for each other's steps.  Each member's
 * results are those of cathy_gpu_flow3d_advance alone, bit for bit.
 * Members: flow handles after flow3d_init, in the column ordering, with
 * IOPT = 1, LUMP = 1 and one IC(0) precision.  While an ensemble advances,
 * its members must not be used through other calls.
 * ------------------------------------------------------------------ */
int cathy_gpu_ensemble_create(int64_t *ens, const int *nmember,
                              const int64_t *members);
/* Advance every member by NSTEPS accepted steps (or to the end of its run).
 * RC(NMEMBER): 0 after NSTEPS steps, 1 at TMAX, 2 on NOBACK; TAKEN(NMEMBER):
 * steps taken; STATS(32, NSTEPS, NMEMBER) (may be NULL): each step's
 * statistics as cathy_gpu_flow3d_advance returns them. */
int cathy_gpu_ensemble_advance(int64_t *ens, const int *nsteps, int *rc,
                               int *taken, double *stats);
/* HIP graphs: with ON != 0 each round's fixed launch sequences (the
 * Picard iteration up to the PCG loop, a follow-up batch of PCG
 * iterations, the post-solve balance) are captured once and replayed until
 * the set of members still iterating changes.  Results are the same bits;
 * default 0. */
int cathy_gpu_ensemble_set_graphs(int64_t *ens, const int *on);
/* out5: members, Picard rounds, PCG polls, PCG launch batches and graph
 * captures of the last advance. */
int cathy_gpu_ensemble_info(int64_t *ens, int64_t *out5);
int cathy_gpu_ensemble_destroy(int64_t *ens);

/* ------------------------------------------------------------------ *
 * Instrumentation (no counterpart in the reference; used by bench.py).
 * ------------------------------------------------------------------ */
/* Enable/disable HIP-event timing of the PCG SpMV kernel (k_spmv_dot2) on
 * the handle's stream and read the accumulated time / launch count. */
int cathy_gpu_flow3d_profile(int64_t *handle, const int *enable,
                             double *spmv_ms, int64_t *launches);
/* Algorithmic bytes of one SpMV launch: 12*nnz + 4*(N+1) + 16*N. */
int cathy_gpu_flow3d_spmv_bytes(int64_t *handle, double *bytes);
/* Matrix layout: out6 = N, nnz of the full pattern, SELL-64 slots
 * (padding included), longest row, 1 if column ids are decoded from the
 * per-chain codes of the extruded mesh (else read per slot), and the
 * number of slots whose code failed the set-up check (0). */
int cathy_gpu_flow3d_layout(int64_t *handle, int64_t *out6);
/* Copy kernel over NBYTES (read NBYTES, write NBYTES), REPS launches;
 * average ms per launch.  Calibrates HBM counters. */
int cathy_gpu_calib_copy(const int64_t *nbytes, const int *reps,
                         double *ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif
