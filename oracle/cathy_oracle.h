/*
 * cathy_oracle.h -- CPU restatement of the CATHY subsurface flow step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (pycathy_wrapper_amd) never links or calls it.
 *
 * Every routine restates one reference Fortran routine of the vendored CATHY
 * processor (pyCATHY/tests/weil_exemple/my_cathy_prj/src/, cited below as
 * SRC/<file>.f:<line>).  Conventions of the restatement:
 *   - all indices are 0-based (node ids, CSR offsets, column ids);
 *   - TETRA is int[nt][5]: 4 node ids sorted ascending (RIORD) + zone id
 *     (1-based, as in the reference's TETRA(5,*));
 *   - per-(layer, zone) soil tables are double[nstr][nzone] row-major;
 *   - the symmetric matrix is the reference's upper CSR: row k holds the
 *     diagonal first, then its columns > k in ascending order (TOPOL, JA);
 *   - the van Genuchten node tables live in one array `vg` of VG_NFIELD*n
 *     doubles, field f of node i at vg[f*n + i] (the /VG/ common block of
 *     SRC/SOILCHAR.H with NMAX replaced by n).
 * Loop order and the order of floating-point operations follow the
 * reference so that, built without FMA contraction, results match the
 * flang-built reference bit-for-bit where libm agrees.
 */
#ifndef CATHY_ORACLE_H
#define CATHY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VG_PNOT = 0, VG_N, VG_M, VG_RMC, VG_PSAT, VG_N1, VG_M52, VG_NR, VG_MM1,
  VG_MR, VG_PSN, VG_NFIELD
};

/* ---------------------------------------------------------------- mesh -- */
/* SRC/triangoli.f:1-159 + SRC/tpnodi2d.f (surface z = mean of the DEM value
 * of incident triangles) + SRC/riord.f (vertex sort).  dem/zone are
 * nrow*ncol row-major (row 0 = north).  Cells whose DEM value is exactly 0
 * are inactive.  Returns 0, fills nnod/ntri; x,y,z sized (nrow+1)*(ncol+1),
 * triang sized 2*nrow*ncol*4 (3 sorted vertex ids + zone). */
int orc_triangoli(int nrow, int ncol, double dx, double dy, double west,
                  double south, const double *dem, const int *zone,
                  double factor, int *nnod, int *ntri, double *x, double *y,
                  double *z, int *triang);

/* SRC/riord.f: sort the first `nod` ids of each of nelt records of width
 * nod+1. */
void orc_riord(int nelt, int *eltcon, int nod);

/* SRC/gen3d.f:13-130 (IVERT 0..4) followed by RIORD(NT,TETRA,4)
 * (SRC/grdsys.f, IOPT=1 and IOPT=2 alike: the sort is harmless for the full
 * pattern).  x,y,z hold the surface in [0,nnod) on entry and all n =
 * nnod*(nstr+1) nodes on exit.  tetra sized 3*ntri*nstr*5. */
void orc_gen3d(int nnod, int nstr, int ntri, int ivert, const int *triang,
               const double *zratio, double base, const double *depth,
               double *x, double *y, double *z, int *tetra, int sort_tetra);

/* SRC/volbas.f:10-54 with SRC/volums.f and SRC/basis6.f (P1 basis
 * coefficients already divided by 6, signed volume -> |V|, IVOL=sign). */
int orc_volbas(int n, int nt, const int *tetra, const double *x,
               const double *y, const double *z, double *ai, double *bi,
               double *ci, double *di, double *volnod, double *volu,
               double *volur, int *ivol);

/* SRC/area2d.f + SRC/areas.f: nodal surface area (1/3 of each incident
 * triangle). */
int orc_area2d(int nnod, int ntri, const int *triang, const double *x,
               const double *y, double *arenod);

/* Symmetric upper pattern of SRC/strpic.f:8-104 (diag first, sorted
 * columns).  topol sized n+1, ja sized cap.  Returns NTERM or -1 if cap is
 * too small. */
int orc_strpic(int n, int nt, const int *tetra, int *topol, int *ja, int cap);

/* Full pattern of SRC/strnew.f (all columns of a row sorted ascending). */
int orc_strnew(int n, int nt, const int *tetra, int *topol, int *ja, int cap);

/* SRC/tetpic.f:10-31: tetja[e*16 + k*4 + l] = position of (TETRA(k),
 * TETRA(l)) in row TETRA(k), for k <= l.  Entries k > l are set to -1. */
int orc_tetpic(int nt, const int *tetra, const int *topol, const int *ja,
               int *tetja);
/* SRC/tetnew.f: the same lookup for all 16 (k,l) in the full pattern. */
int orc_tetnew(int nt, const int *tetra, const int *topol, const int *ja,
               int *tetja);

/* SRC/locmas.f: 4x4 local mass factors (LUMP=0 consistent, 1 lumped). */
void orc_locmas(double *lmass, int lump);

/* ------------------------------------------------------ soil params -- */
/* SRC/tpnodi.f:9-70 (IPEAT=0): node averages of per-(layer, zone) porosity,
 * storage and vG parameters over incident tets; PEL per element. */
int orc_tpnodi(int n, int nt, int ntri, const int *tetra, int nzone,
               const double *poros, const double *elstor,
               const double *vgncell, const double *vgrmccell,
               const double *vgpsatcell, double *pnodi, double *snodi,
               double *vg, double *pel, int *tp);

/* SRC/chparm.f:11-34 (IVGHU=0). */
void orc_chparm(int n, const double *pnodi, double *vg);

/* SRC/eltnod.f: node value = mean over incident elements. */
void orc_eltnod(int n, int nt, const int *tp, const int *tetra,
                const double *velt, double *vnod);
/* SRC/nodelt.f: element value = mean of its 4 nodes. */
void orc_nodelt(int nt, const int *tetra, const double *vnod, double *velt);

/* ---------------------------------------------- van Genuchten (a2) -- */
double orc_fvgse(double psi, const double *vg, int n, int i);   /* fvgse.f */
double orc_fvgkr(double psi, double se, const double *vg, int n, int i);
double orc_fvgdse(double psi, const double *vg, int n, int i); /* fvgdse.f */

/* SRC/chpic0.f:23-36 (IVGHU=0). */
void orc_chpic0(int n, const double *ptnew, const double *pnew,
                const double *ptimep, double *swnew, double *swtimep,
                const double *snodi, const double *pnodi, const double *vg,
                double *sw, double *ckrw, double *etai, double *et1,
                double *et2);
/* SRC/chvelo.f (IVGHU=0): Sw and kr at PNEW. */
void orc_chvelo(int n, const double *pnew, const double *pnodi,
                const double *vg, double *sw, double *ckrw);
/* SRC/chmass.f (IVGHU=0): ETAI at PNEW. */
void orc_chmass(int n, const double *pnew, const double *snodi,
                const double *pnodi, const double *vg, double *etai);

/* ------------------------------------------------- assembly (a3-a7) -- */
/* SRC/asspic.f:9-54.  perm* are [nstr][nzone].  Accumulates (caller
 * zeroes coef1/2/4 as SRC/picard.f:87-90 does). */
void orc_asspic(int nt, int ntri, int nzone, const int *tetra,
                const int *tetja, const double *lmass, double *coef1,
                double *coef2, double *coef4, const double *et1e,
                const double *pel, const double *permx, const double *permy,
                const double *permz, const double *ckrwe,
                const double *volu, const double *volur, const double *bi,
                const double *ci, const double *di);
/* SRC/rhspic.f:10-41. */
void orc_rhspic(int n, const int *ja, const int *topol, double deltat,
                const double *ptnew, const double *pnew,
                const double *ptimep, const double *swnew,
                const double *swtimep, const double *coef1,
                const double *coef2, const double *coef4, double *tnoti);
/* SRC/cfmatp.f:9-29 (coef5 is scratch, zeroed by the caller). */
void orc_cfmatp(int n, int nterm, double tetaf, double deltat, double *coef1,
                const double *coef2, const double *coef4, double *coef5,
                const int *topol, const double *et2);
/* SRC/rhsgrv.f:9-33. */
void orc_rhsgrv(int nt, int ntri, int nzone, const int *tetra,
                double *tnoti, const double *di, const double *permz,
                const int *ivol, const double *ckrwe);
/* SRC/bcpic.f:10-89 without seepage faces (NSF=0).  noddir must hold the
 * NP fixed-head nodes in [0,np) on entry (SRC/cathy_main.f:2983); returns
 * NUMDIR. */
int orc_bcpic(int np, int nq, const int *contp, const int *contq,
              double *lhsp, const double *q, int nnod, int n,
              double *lhsatm, const int *ifatm, const double *atmact,
              const double *atmold, const int *topol, double *coef1,
              double *tnoti, double tetaf, double rmax, int *noddir,
              const double *qtranie);
/* SRC/shlpic.f:16-60 without seepage faces. */
void orc_shlpic(int np, const int *topol, const int *contp,
                const double *presc, const double *lhsp, double *coef1,
                double *pnew, const double *pold, int nnod, const int *ifatm,
                const double *lhsatm);

/* ------------------------------------------------ IC(0)-PCG (a8-a11) -- */
/* SRC/solscal-extended.f:2142-2193 INCLDP.  Returns the number of
 * breakdown fallbacks taken (A<=0 -> previous diagonal squared). */
int orc_incldp(const int *ja, const int *topol, const double *coefr,
               double *lriga, int n, int nterm);
/* :4169-4204 PRODDP: x = (U^T U)^-1 b. */
void orc_proddp(double *x, const double *b, int n, const int *topol,
                const int *ja, const double *coef2);
/* :1260-1386 GRADDP (PCG).  scr,res,b,r,p are n-vectors.  Returns NITER;
 * *err receives the last relative residual. */
int orc_graddp(int n, double *poten, double *scr, double *res, double *bv,
               double *r, double *p, const int *topol, const int *ja,
               const double *coef1, const double *coef2,
               const double *tnoti, double tol, int imax, int np,
               const int *contp, double *err);
/* :4389-4438 RESSYM. */
double orc_ressym(const int *topol, const int *ja, const double *coef1,
                  const double *tnoti, const double *pnew, double *resv,
                  int n, int numdir, const int *noddir, double rmin);
/* :4669-4699 SYMSLV.  coef2 (nterm) and scr (5n) are scratch.  Returns
 * NITER; err/res receive GRADDP's and RESSYM's residuals. */
int orc_symslv(int n, int nterm, int numdir, int itmxcg, double tolcg,
               double rmin, const int *noddir, const int *ja,
               const int *topol, double *pnew, const double *tnoti,
               const double *coef1, double *coef2, double *scr, double *err,
               double *res);

/* ------------------------------------------- mass balance (a13-a14) -- */
/* SRC/bkpic.f:9-56 (no seepage faces). scr (n) receives A*dpsi - XT5. */
void orc_bkpic(int n, int np, int nnod, double tetaf, const int *topol,
               const int *ja, const int *contp, const int *ifatm,
               const double *pdiff, double *qpnew, const double *qpold,
               double *atmact, const double *atmold, const double *coef1,
               const double *xt5, double *scr);
/* SRC/stormb.f:10-29. */
double orc_stormb(const double *pnew, const double *ptimep,
                  const double *swnew, const double *swtimep, int n,
                  const double *volnod, const double *snodi,
                  const double *pnodi);
/* SRC/norms.f:9-41.  Returns IKMAX (0-based). */
int orc_norms(int n, double *pikmax, double *pinf, double *pl2, double *finf,
              double *fl2, const double *pnew, const double *pold,
              const double *tnoti);
/* SRC/etran.f:15-80: Feddes root uptake QTRANIE(n); vegtype[nnod] 0-based
 * into the per-type tables.  Returns -1 if a root zone reaches the base. */
int orc_etran(int n, int nnod, int nstr, const double *atmpot,
              const double *z, const double *psi, const int *vegtype,
              const double *pcana, const double *pcref, const double *pcwlt,
              const double *zroot, const double *pz, const double *omgc,
              double scf, double *qtranie);
/* SRC/switch_old.f:12-53. */
void orc_switch_old(int nnod, int *ifatm, double *atmact,
                    const double *atmpot, double *pnew, double pmin);
/* SRC/switch.f:8-423: the coupled (SURF) switching CONVER calls instead
 * of SWITCH_OLD.  PL = PONDNOD + (ATMPOT - (ATMACT - QTRANIE))*DELTAT/
 * ARENOD decides the IFATM 1 / 2 cases, PNEW the IFATM 0 ones.  Arrays
 * are per surface node.  Returns PONDING. */
int orc_switch(int nnod, int *ifatm, double deltat, double pondh_min,
               double pmin, const double *arenod, const double *pondnod,
               const double *atmpot, double *atmact, const double *qtranie,
               double *pnew, double *ovflnod);
/* SRC/adrstn.f:13-33: an 'air dry' surface node under rainfall (or whose
 * actual flux exceeds the potential) goes back to the Neumann condition
 * at PNEW = PMIN. */
void orc_adrstn(int nnod, int *ifatm, const double *atmpot, double *atmact,
                double *pnew, double pmin);
/* SRC/timupd.f:8-41.  Returns DTGMIN. */
int orc_timupd(double *timep, double *time, double *deltat, double dtmin,
               double dtmax, double tmax, double dtmaga, double dtmagm,
               double dtreds, double dtredm, int iter, int ituns1,
               int ituns2);
/* SRC/icvdwt.f:8-31: hydrostatic head below a water table at depth wt. */
void orc_icvdwt(int nnod, int nstr, double wtposition, double *ptimep,
                const double *z);
/* SRC/neumann.f:5-37 free-drainage branch (NNEU<0): bottom flux
 * -A*kr*Kz at the nnod nodes listed in acontq. */
void orc_neumann_free(int nnod, const int *acontq, const double *arenod,
                      const double *ckrw, const double *kznod, double *q);

/* ------------------------- Newton linearisation, IOPT=2 (a16) -- */
/* SRC/fvgdds.f, SRC/fvgdkr.f. */
double orc_fvgdds(double psi, const double *vg, int n, int i);
double orc_fvgdkr(double psi, const double *vg, int n, int i);
/* SRC/chnew0.f (IVGHU=0): Sw, kr, ETAI and their psi-derivatives. */
void orc_chnew0(int n, const double *ptnew, const double *snodi,
                const double *pnodi, const double *vg, double *sw,
                double *ckrw, double *etai, double *dckrw, double *detai);
/* SRC/assnew.f on the full pattern; tetja[e*16+k*4+l] = position of
 * (TETRA(k), TETRA(l)).  Accumulates (caller zeroes coef1/2/3). */
void orc_assnew(int nt, int ntri, int nzone, const int *tetra,
                const int *tetja, const double *lmass, double *coef1,
                double *coef2, double *coef3, const double *permx,
                const double *permy, const double *permz, const double *ckrwe,
                const double *etae, const double *ptnew, const double *ptimep,
                const double *pnew, const double *dckrw, const double *detai,
                double tetaf, double deltat, const double *volu,
                const double *volur, const int *ivol, const double *bi,
                const double *ci, const double *di);
/* SRC/rhsnew.f, SRC/cfmatn.f. */
void orc_rhsnew(int n, const int *ja, const int *topol, double deltat,
                const double *ptnew, const double *pnew, const double *ptimep,
                const double *coef1, const double *coef2, double *tnoti);
void orc_cfmatn(int nterm, double tetaf, double deltat, double *coef1,
                const double *coef2, const double *coef3);
/* SRC/bcnew.f, SRC/shlnew.f, SRC/bknew.f (NP=0, NSF=0). */
int orc_bcnew(int nq, const int *contq, const double *q, int nnod,
              double *lhsatm, const int *ifatm, const double *atmact,
              const double *atmold, const int *ja, const int *topol,
              double *coef1, double *tnoti, double tetaf, double rmax,
              int *noddir);
void orc_shlnew(const int *topol, const int *ja, double *coef1, double *pnew,
                const double *pold, int nnod, const int *ifatm,
                const double *lhsatm);
void orc_bknew(int nnod, double tetaf, const int *topol, const int *ja,
               const int *ifatm, const double *pdiff, double *atmact,
               const double *atmold, const double *coef1, const double *xt5);
/* SRC/topia.f. */
void orc_topia(int n, const int *topol, int *ia);
/* SRC/solscal-extended.f:2194-2270 INCLU (Crout ILU(0)); returns the
 * number of zero pivots replaced. */
int orc_inclu(const int *ia, const int *ja, const int *topol,
              const double *coefr, double *coefc, double *lriga,
              double *lcolon, int n, int nterm);
/* :4205 PRODL, :4228 PRODU, :4251 PRU, :4107 PROD1. */
void orc_prodl(double *z, const double *y, const int *ia, const int *ja,
               const double *coef2, int nterm);
void orc_produ(double *x, const double *h, int n, const int *ia,
               const int *ja, const double *coef2, const int *topol);
void orc_pru(double *y, const double *x, const int *ia, const int *ja,
             const double *coef2, int nterm);
void orc_prod1(double *x, const double *b, int n, const int *ia,
               const int *ja, const double *coef2, int nterm,
               const int *topol);
/* :4332 RESNSY. */
double orc_resnsy(const int *topol, const int *ja, const double *coef1,
                  const double *tnoti, const double *pnew, double *resv,
                  int n, int numdir, const int *noddir, double rmin);
/* :1010-1128 GCSTAS (scr: 10 n). */
int orc_gcstas(int n, double *poten, const double *tnoti, double *scr,
               const int *topol, const int *ia, const int *ja,
               const double *coef1, const double *coef2, int nterm, int imax,
               double tol, int np, const int *contp, double *err_out);
/* :3063-3174 NSYSLV with ISOLV=0 (scr: max(nterm, 10 n)); coef2 receives
 * the ILU factor, coef3 is overwritten.  Returns NITER. */
int orc_nsyslv(int n, int nterm, int numdir, const int *noddir, int itmxcg,
               double tolcg, double rmin, const int *ia, const int *ja,
               const int *topol, const double *coef1, double *coef2,
               double *coef3, double *scr, double *pnew, const double *tnoti,
               double *err, double *res, double *resini);

/* ------------------------------------------------ Darcy velocities -- */
/* SRC/vel3d.f: UU/VV/WW per element (perm* are [nstr][nzone]). */
void orc_vel3d(int nt, int ntri, int nzone, const int *tetra, const double *pnew,
               const double *permx, const double *permy, const double *permz,
               double *uu, double *vv, double *ww, const double *bi,
               const double *ci, const double *di, const double *ckrwe,
               const double *volur, const int *ivol);
/* SRC/vnod3d.f: nodal means over the tp[i] incident elements. */
void orc_vnod3d(int n, int nt, const int *tp, const int *tetra, const double *uu,
                const double *vv, const double *ww, double *unod, double *vnod,
                double *wnod);

/* --------------------------------------------- whole run (a1-a18) -- */
typedef struct orc_params {
  /* input/parm (SRC/datin.f:80-122) */
  double tetaf;
  int lump, iopt, l2norm;
  double toluns, tolswi, ernlmx;
  int ituns, ituns1, ituns2;
  int itmxcg;
  double tolcg;
  double deltat, dtmin, dtmax, tmax;
  double dtmaga, dtmagm, dtreds, dtredm;
  /* soil file head (SRC/datin.f:421) and atmospheric options */
  double pmin, pondh_min, scf;
  int ieto;
  /* free drainage at the base (nansfneubc NNEU<0) */
  int free_drainage;
} orc_params;

typedef struct orc_run orc_run;

/* Mesh + soil + initial head + forcing.  Arrays are copied.  atm_times
 * (n_atm) / atm_vals (n_atm * (hspatm ? 1 : nnod)) hold the atmbc series
 * (m/s, +ve rain).  ptimep0 is the initial pressure head (n). */
orc_run *orc_run_create(int nnod, int nstr, int ntri, const int *tetra,
                        const double *x, const double *y, const double *z,
                        const double *arenod, int nzone,
                        const double *permx, const double *permy,
                        const double *permz, const double *elstor,
                        const double *poros, const double *vgn,
                        const double *vgrmc, const double *vgpsat,
                        const double *ptimep0, int n_atm, int hspatm,
                        const double *atm_times, const double *atm_vals,
                        const orc_params *prm);
void orc_run_free(orc_run *r);

/* Advance one accepted time step (back-steps included).  Returns 0 while
 * the run continues, 1 when TMAX was reached, 2 on NOBACK termination.
 * out_stats (if not NULL) receives 24 doubles:
 *  [0] NSTEP [1] DELTAT [2] TIME [3] ITER [4] NITERT (PCG its of the step)
 *  [5] back-steps [6] STORE1 [7] STORE2 [8] DSTORE [9] CDSTOR [10] VIN
 *  [11] CVIN [12] VOUT [13] CVOUT [14] VIN+VOUT [15] VTOT [16] ERRAS
 *  [17] ERREL [18] CERRAS [19] CAERAS [20] PINF [21] total Picard its
 *  [22] total PCG its [23] linear-solver failures */
int orc_run_step(orc_run *r, double *out_stats);

/* State access (n doubles / nnod ints). */
void orc_run_get_psi(const orc_run *r, double *psi);
void orc_run_get_sw(const orc_run *r, double *sw);
void orc_run_get_ifatm(const orc_run *r, int *ifatm);
int orc_run_n(const orc_run *r);
int orc_run_nterm(const orc_run *r);
/* Pattern / geometry views for tests (pointers stay owned by the run). */
const int *orc_run_topol(const orc_run *r);
const int *orc_run_ja(const orc_run *r);

/* ---- surface routing (ISIMGR = 2): one ROUTE + ALTEZZE substep ----------
 * Grids of the drainage network are the reference's DTM_*(COLMAX, ROWMAX)
 * arrays, element (I, J) at (I-1) + (J-1)*ldc, I = 1..ncol, J = 1..nrow;
 * cell arrays are indexed by I_BASIN - 1 = (I-1)*nrow + J - 1.  qoi holds
 * the 1-based I_BASIN of the ncell cells in routing order (QOI_SN).  Only
 * cells without reservoirs (TIPO_R = 0) are restated; the network is taken
 * as given.  Q_OUT_KKP1_SN_1/2 of a branch with zero weight keep their input
 * values, as in the reference; q_in_kkp1 enters with the values the caller
 * holds and receives the upstream outflows.  cu_ak[0] = CU_MAX (reset to 0),
 * cu_ak[1] = AK_MAX (kept if no branch is routed). */
typedef struct {
  int nrow, ncol, ldc, ncell;
  const int *qoi, *pout1, *pout2;
  const double *w1, *w2, *slope1, *slope2, *epl1, *epl2, *kss1, *kss2, *ws1, *ws2, *b1, *y1, *nrc;
  /* reservoirs (NULL tipo_r: none).  tipo_r / reservr / elevation per cell
   * (I_BASIN - 1): TIPO_R, RESERVR (1-based reservoir), ELEVATION.  Per
   * reservoir r (0-based): n_ha[r] table entries hres/ares[r + i*maxres]
   * (HRES(MAXRES,30), ARES), h_soglia[r], h_fondo[r]. */
  const int *tipo_r, *reservr, *n_ha;
  int maxres;
  const double *hres, *ares, *h_soglia, *h_fondo, *elevation;
} orc_network;

/* SERBATOIO (SRC/serbatoio.f:6-119) with CALCOLI (SRC/calcoli.f): RK4
 * level-pool step of reservoir r.  Returns 0, or -1 where CALCOLI stops
 * (level below its height table). */
int orc_serbatoio(const orc_network *net, int r, double q_in_kk, double q_in_kkp1, double h_pool_kk,
                  double delta_t, double surface_water, double *q_out_kkp1, double *h_pool_kkp1);

void orc_mc(double local_slope, double epl, double ks, double w, double b1, double y1, double deltat,
            double q_in_kk, double q_in_kkp1, double q_out_kk, double q_overland, double *cu, double *ak,
            double *q_out_kkp1);
/* Returns 0, or -1 where the reference stops (a reservoir level outside its
 * table, or a routed cell with TIPO_R >= 1000).  h_pool_kk / h_pool_kkp1:
 * H_POOL_KK_VEC / H_POOL_KKP1_VEC (only with reservoirs). */
int orc_route(const orc_network *net, double deltat, const double *surface_water, const double *q_in_kk,
              const double *q_out_kk1, const double *q_out_kk2, double *q_in_kkp1, double *q_out_kkp1_1,
              double *q_out_kkp1_2, double *cu_ak, const double *h_pool_kk, double *h_pool_kkp1);
/* ALTEZZE (SRC/altezze.f:11-107) over ncell cells with its type array
 * celtype (NULL: 0 everywhere, what the reference's time loop passes,
 * cathy_main.f:3062): cells of type 1..999 take the depth of reservoir
 * NUM_R(c) = net->reservr[c] from H_POOL_KKP1_VEC, type >= 1000 the lake
 * depth of the reservoir of cell type-1000; both keep volume_kkp1. */
void orc_altezze(const orc_network *net, const int *celtype, int ncell, double dx, double dy, double deltat,
                 const double *q_in_kk, const double *q_in_kkp1, const double *q_out_kk1, const double *q_out_kk2,
                 const double *q_out_kkp1_1, const double *q_out_kkp1_2, const double *volume_kk,
                 const double *surface_water, double *volume_kkp1, double *h_water_kkp1,
                 const double *h_pool_kkp1);

#ifdef __cplusplus
}
#endif
#endif
