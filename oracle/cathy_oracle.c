/*
 * cathy_oracle.c -- CPU restatement of the CATHY subsurface Picard step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see the header).
 * Each function cites the reference routine it restates as SRC/<file>:<line>
 * where SRC = pyCATHY/tests/weil_exemple/my_cathy_prj/src.  Built with
 * -ffp-contract=off so the arithmetic order matches the reference's.
 */
#include "cathy_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define VGF(vg, n, f, i) ((vg)[(size_t)(f) * (size_t)(n) + (size_t)(i)])

/* ===================================================================== */
/*                                 mesh                                  */
/* ===================================================================== */

void orc_riord(int nelt, int *eltcon, int nod) {
  /* SRC/riord.f: exchange sort of the first nod ids of each record. */
  for (int e = 0; e < nelt; ++e) {
    int *c = eltcon + (size_t)e * (nod + 1);
    for (int k = 0; k < nod - 1; ++k)
      for (int j = k + 1; j < nod; ++j)
        if (c[k] > c[j]) {
          int t = c[k];
          c[k] = c[j];
          c[j] = t;
        }
  }
}

int orc_triangoli(int nrow, int ncol, double dx, double dy, double west,
                  double south, const double *dem, const int *zone,
                  double factor, int *nnod_out, int *ntri_out, double *x,
                  double *y, double *z, int *triang) {
  /* SRC/triangoli.f:25-150 with DOSTEP=1: a node exists at each corner of
   * an active cell; nodes numbered row by row (north first); each cell
   * (i,j) gives triangles (i,j)-(i+1,j)-(i+1,j+1) and
   * (i,j)-(i+1,j+1)-(i,j+1). */
  int nr1 = nrow + 1, nc1 = ncol + 1;
  int *nodi = (int *)calloc((size_t)nr1 * nc1, sizeof(int));
  double *eltria = (double *)malloc(sizeof(double) * 2 * (size_t)nrow * ncol);
  if (!nodi || !eltria) return -1;
  for (int i = 0; i < nrow; ++i)
    for (int j = 0; j < ncol; ++j)
      if (dem[(size_t)i * ncol + j] != 0.0) {
        nodi[(size_t)i * nc1 + j] = 1;
        nodi[(size_t)(i + 1) * nc1 + j] = 1;
        nodi[(size_t)(i + 1) * nc1 + j + 1] = 1;
        nodi[(size_t)i * nc1 + j + 1] = 1;
      }
  int inod = 0;
  for (int i = 0; i < nr1; ++i)
    for (int j = 0; j < nc1; ++j)
      if (nodi[(size_t)i * nc1 + j] == 1) {
        x[inod] = west + j * dx;
        y[inod] = south + (nrow - i) * dy; /* (nrow-i+1) with 1-based i */
        nodi[(size_t)i * nc1 + j] = inod + 1; /* keep 1-based marker */
        ++inod;
      }
  int itria = 0;
  for (int i = 0; i < nrow; ++i)
    for (int j = 0; j < ncol; ++j) {
      double d = dem[(size_t)i * ncol + j];
      if (d == 0.0) continue;
      double el = d * factor / 1.0; /* one non-null cell per triangle */
      int a = nodi[(size_t)i * nc1 + j] - 1;
      int b = nodi[(size_t)(i + 1) * nc1 + j] - 1;
      int c = nodi[(size_t)(i + 1) * nc1 + j + 1] - 1;
      int dd = nodi[(size_t)i * nc1 + j + 1] - 1;
      int zn = zone ? zone[(size_t)i * ncol + j] : 1;
      int *t = triang + (size_t)itria * 4;
      t[0] = a; t[1] = b; t[2] = c; t[3] = zn;
      eltria[itria++] = el;
      t = triang + (size_t)itria * 4;
      t[0] = a; t[1] = c; t[2] = dd; t[3] = zn;
      eltria[itria++] = el;
    }
  /* SRC/tpnodi2d.f: surface z = mean DEM value of incident triangles. */
  int *tp2d = (int *)calloc((size_t)inod, sizeof(int));
  for (int k = 0; k < inod; ++k) z[k] = 0.0;
  for (int k = 0; k < itria; ++k)
    for (int ii = 0; ii < 3; ++ii) {
      int nd = triang[(size_t)k * 4 + ii];
      z[nd] = z[nd] + eltria[k];
      tp2d[nd] += 1;
    }
  int bad = 0;
  for (int k = 0; k < inod; ++k) {
    if (tp2d[k] == 0) bad = 1;
    else z[k] = z[k] / tp2d[k];
  }
  /* SRC/grdsys.f: RIORD(NTRI,TRIANG,3) before GEN3D. */
  orc_riord(itria, triang, 3);
  *nnod_out = inod;
  *ntri_out = itria;
  free(tp2d);
  free(nodi);
  free(eltria);
  return bad ? -2 : 0;
}

void orc_gen3d(int nnod, int nstr, int ntri, int ivert, const int *triang,
               const double *zratio, double base, const double *depth,
               double *x, double *y, double *z, int *tetra, int sort_tetra) {
  /* SRC/gen3d.f:24-45: three tets per prism (layer J between node planes
   * J-1 and J). */
  for (int J = 1; J <= nstr; ++J)
    for (int I = 0; I < ntri; ++I) {
      const int *tr = triang + (size_t)I * 4;
      size_t nt0 = 3 * ((size_t)(J - 1) * ntri + I);
      int *t1 = tetra + nt0 * 5, *t2 = t1 + 5, *t3 = t2 + 5;
      int lo = (J - 1) * nnod, hi = J * nnod;
      t1[0] = lo + tr[0]; t1[1] = lo + tr[1]; t1[2] = lo + tr[2];
      t1[3] = hi + tr[0]; t1[4] = tr[3];
      t2[0] = t1[3]; t2[1] = hi + tr[1]; t2[2] = hi + tr[2];
      t2[3] = t1[2]; t2[4] = tr[3];
      t3[0] = t1[1]; t3[1] = t1[2]; t3[2] = t2[1];
      t3[3] = t1[3]; t3[4] = tr[3];
    }
  /* SRC/gen3d.f:53-75: vertical coordinates. */
  double zmin = 1.7e100;
  for (int i = 0; i < nnod; ++i)
    if (z[i] < zmin) zmin = z[i];
  for (int i = 0; i < nnod; ++i) {
    double zthick = (z[i] - zmin) + base;
    double zrsum = 0.0;
    for (int J = 1; J <= nstr; ++J) {
      size_t kk = (size_t)J * nnod + i;
      x[kk] = x[i];
      y[kk] = y[i];
      zrsum = zrsum + zratio[J - 1];
      if (ivert == 0) z[kk] = z[i] - zrsum * base;
      else if (ivert == 1) z[kk] = z[i] - zrsum * zthick;
      else if (ivert == 2) z[kk] = zmin - zrsum * base;
      else if (ivert == 3) z[kk] = z[i] - zrsum * depth[i];
      else if (ivert == 4) {
        z[kk] = z[i] - zrsum * base;
        if (J == nstr) z[kk] = zmin - base;
      }
    }
  }
  if (sort_tetra) orc_riord(3 * ntri * nstr, tetra, 4);
}

/* IP4 of SRC/cathy_main.f:2491 (column-major DATA), 0-based. */
static const int IP4[4][4] = {
    /* IP4(r, nn) stored as IP4c[nn][r] */
    {0, 1, 2, 3}, {1, 2, 3, 0}, {2, 3, 0, 1}, {3, 0, 1, 2}};

/* Build the 3x3 cyclic table of SRC/volums.f:22-33 / SRC/basis6.f:26-37
 * for local vertex nn: column 0 = the other three vertices, columns 1 and 2
 * their cyclic shifts. */
static void ip3_for(int nn, int ip3[3][3]) {
  for (int ii = 1; ii < 4; ++ii) ip3[ii - 1][0] = IP4[nn][ii];
  for (int lc = 1; lc < 3; ++lc)
    for (int lr = 0; lr < 3; ++lr)
      ip3[lr][lc] = (lr == 2) ? ip3[0][lc - 1] : ip3[lr + 1][lc - 1];
}

int orc_volbas(int n, int nt, const int *tetra, const double *x,
               const double *y, const double *z, double *ai, double *bi,
               double *ci, double *di, double *volnod, double *volu,
               double *volur, int *ivol) {
  static const double amen4[4] = {-1.0, 1.0, -1.0, 1.0};
  static const double amen5[5] = {-1.0, 1.0, -1.0, 1.0, -1.0};
  for (int i = 0; i < n; ++i) volnod[i] = 0.0;
  int bad = 0;
  for (int e = 0; e < nt; ++e) {
    const int *t = tetra + (size_t)e * 5;
    int ip3[3][3];
    /* SRC/volums.f: signed volume by cofactor expansion. */
    double vol = 0.0;
    for (int nn = 0; nn < 4; ++nn) {
      ip3_for(nn, ip3);
      double a2 = 0.0, a3 = 0.0;
      for (int ii = 0; ii < 3; ++ii) {
        int I = t[ip3[ii][0]], J = t[ip3[ii][1]], M = t[ip3[ii][2]];
        a3 = y[I] * z[J] + a3;
        a2 = y[I] * z[M] + a2;
      }
      vol = vol + x[t[nn]] * amen4[nn] * (a3 - a2) / 6.0;
    }
    /* SRC/basis6.f: cofactors / 6. */
    for (int nn = 0; nn < 4; ++nn) {
      ip3_for(nn, ip3);
      double a3 = 0.0, a2;
      for (int ii = 0; ii < 3; ++ii) {
        int I = t[ip3[ii][0]], J = t[ip3[ii][1]], M = t[ip3[ii][2]];
        a3 = a3 + x[M] * (y[I] * z[J] - z[I] * y[J]);
      }
      ai[(size_t)e * 4 + nn] = amen5[nn + 1] * a3 / 6.0;
      a2 = 0.0; a3 = 0.0;
      for (int ii = 0; ii < 3; ++ii) {
        int I = t[ip3[ii][0]], J = t[ip3[ii][1]], M = t[ip3[ii][2]];
        a3 = y[I] * z[J] + a3;
        a2 = y[I] * z[M] + a2;
      }
      bi[(size_t)e * 4 + nn] = amen5[nn] * (a3 - a2) / 6.0;
      a2 = 0.0; a3 = 0.0;
      for (int ii = 0; ii < 3; ++ii) {
        int I = t[ip3[ii][0]], J = t[ip3[ii][1]], M = t[ip3[ii][2]];
        a3 = x[I] * z[J] + a3;
        a2 = x[I] * z[M] + a2;
      }
      ci[(size_t)e * 4 + nn] = amen5[nn + 1] * (a3 - a2) / 6.0;
      a2 = 0.0; a3 = 0.0;
      for (int ii = 0; ii < 3; ++ii) {
        int I = t[ip3[ii][0]], J = t[ip3[ii][1]], M = t[ip3[ii][2]];
        a3 = x[I] * y[J] + a3;
        a2 = x[I] * y[M] + a2;
      }
      di[(size_t)e * 4 + nn] = amen5[nn] * (a3 - a2) / 6.0;
    }
    if (vol == 0.0) bad = 1;
    double vol4 = fabs(vol) * 0.25;
    for (int k = 0; k < 4; ++k) volnod[t[k]] = volnod[t[k]] + vol4;
    ivol[e] = 1;
    if (vol < 0.0) {
      ivol[e] = -1;
      vol = -vol;
    }
    volu[e] = vol;
    volur[e] = 1.0 / vol;
  }
  return bad ? -1 : 0;
}

int orc_area2d(int nnod, int ntri, const int *triang, const double *x,
               const double *y, double *arenod) {
  /* SRC/area2d.f + SRC/areas.f with IP3 = {1,2,3; 2,3,1; 3,1,2}. */
  static const int ip3[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
  const double r3 = 1.0 / 3.0;
  for (int i = 0; i < nnod; ++i) arenod[i] = 0.0;
  int bad = 0;
  for (int e = 0; e < ntri; ++e) {
    const int *t = triang + (size_t)e * 4;
    double a2 = 0.0, a3 = 0.0;
    for (int ii = 0; ii < 3; ++ii) {
      int I = t[ip3[ii][0]], J = t[ip3[ii][1]], M = t[ip3[ii][2]];
      a3 = x[I] * y[J] + a3;
      a2 = x[I] * y[M] + a2;
    }
    double are = 0.5 * (a3 - a2);
    if (are == 0.0) bad = 1;
    double are3 = fabs(are) * r3;
    for (int k = 0; k < 3; ++k) arenod[t[k]] = arenod[t[k]] + are3;
  }
  return bad ? -1 : 0;
}

/* Row-wise sorted insertion of tet connectivity.  upper=1 reproduces the
 * STRPIC pattern (diag first, then columns > row), upper=0 the STRNEW
 * full pattern (all columns sorted, diag included). */
static int build_pattern(int n, int nt, const int *tetra, int *topol, int *ja,
                         int cap, int upper) {
  const int W = 32; /* slots per row while building (N1) */
  int *slot = (int *)malloc(sizeof(int) * (size_t)n * W);
  int *cnt = (int *)calloc((size_t)n, sizeof(int));
  if (!slot || !cnt) return -1;
  if (upper)
    for (int r = 0; r < n; ++r) {
      slot[(size_t)r * W] = r;
      cnt[r] = 1;
    }
  int ok = 1;
  for (int e = 0; e < nt && ok; ++e) {
    int v[4];
    for (int k = 0; k < 4; ++k) v[k] = tetra[(size_t)e * 5 + k];
    /* ascending order of the element's nodes (STRPIC/STRNEW selection) */
    for (int k = 0; k < 3; ++k)
      for (int j = k + 1; j < 4; ++j)
        if (v[k] > v[j]) { int t = v[k]; v[k] = v[j]; v[j] = t; }
    for (int a = 0; a < 4; ++a)
      for (int b = upper ? a + 1 : 0; b < 4; ++b) {
        int r = v[a], c = v[b];
        int *s = slot + (size_t)r * W;
        int m = cnt[r], pos = upper ? 1 : 0;
        while (pos < m && s[pos] < c) ++pos;
        if (pos < m && s[pos] == c) continue;
        if (m >= W) { ok = 0; break; }
        memmove(s + pos + 1, s + pos, sizeof(int) * (size_t)(m - pos));
        s[pos] = c;
        cnt[r] = m + 1;
      }
  }
  int nterm = -1;
  if (ok) {
    size_t tot = 0;
    for (int r = 0; r < n; ++r) tot += (size_t)cnt[r];
    if (tot <= (size_t)cap) {
      size_t m = 0;
      for (int r = 0; r < n; ++r) {
        topol[r] = (int)m;
        for (int k = 0; k < cnt[r]; ++k) ja[m++] = slot[(size_t)r * W + k];
      }
      topol[n] = (int)m;
      nterm = (int)m;
    }
  }
  free(slot);
  free(cnt);
  return nterm;
}

int orc_strpic(int n, int nt, const int *tetra, int *topol, int *ja,
               int cap) {
  /* SRC/strpic.f:8-104. */
  return build_pattern(n, nt, tetra, topol, ja, cap, 1);
}

int orc_strnew(int n, int nt, const int *tetra, int *topol, int *ja,
               int cap) {
  /* SRC/strnew.f. */
  return build_pattern(n, nt, tetra, topol, ja, cap, 0);
}

static int find_in_row(const int *topol, const int *ja, int r, int c) {
  for (int m = topol[r]; m < topol[r + 1]; ++m)
    if (ja[m] == c) return m;
  return -1;
}

int orc_tetpic(int nt, const int *tetra, const int *topol, const int *ja,
               int *tetja) {
  /* SRC/tetpic.f:19-30 (requires sorted TETRA). */
  int bad = 0;
  for (int e = 0; e < nt; ++e)
    for (int k = 0; k < 4; ++k)
      for (int l = 0; l < 4; ++l) {
        int v = -1;
        if (l >= k) {
          v = find_in_row(topol, ja, tetra[(size_t)e * 5 + k],
                          tetra[(size_t)e * 5 + l]);
          if (v < 0) bad = 1;
        }
        tetja[(size_t)e * 16 + k * 4 + l] = v;
      }
  return bad ? -1 : 0;
}

int orc_tetnew(int nt, const int *tetra, const int *topol, const int *ja,
               int *tetja) {
  /* SRC/tetnew.f. */
  int bad = 0;
  for (int e = 0; e < nt; ++e)
    for (int k = 0; k < 4; ++k)
      for (int l = 0; l < 4; ++l) {
        int v = find_in_row(topol, ja, tetra[(size_t)e * 5 + k],
                            tetra[(size_t)e * 5 + l]);
        if (v < 0) bad = 1;
        tetja[(size_t)e * 16 + k * 4 + l] = v;
      }
  return bad ? -1 : 0;
}

void orc_locmas(double *lmass, int lump) {
  /* SRC/locmas.f:13-48. */
  const double r4 = 1.0 / 4.0, r10 = 1.0 / 10.0, r20 = 1.0 / 20.0;
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 4; ++i) {
      double v;
      if (lump == 0) v = (k == i) ? r10 : r20;
      else v = (k == i) ? r4 : 0.0;
      lmass[k * 4 + i] = v;
    }
}

/* ===================================================================== */
/*                              soil params                              */
/* ===================================================================== */

int orc_tpnodi(int n, int nt, int ntri, const int *tetra, int nzone,
               const double *poros, const double *elstor,
               const double *vgncell, const double *vgrmccell,
               const double *vgpsatcell, double *pnodi, double *snodi,
               double *vg, double *pel, int *tp) {
  /* SRC/tpnodi.f:28-70 with INIT1 zeroing (SRC/init1.f). */
  for (int i = 0; i < n; ++i) {
    pnodi[i] = 0.0;
    snodi[i] = 0.0;
    tp[i] = 0;
    VGF(vg, n, VG_N, i) = 0.0;
    VGF(vg, n, VG_RMC, i) = 0.0;
    VGF(vg, n, VG_PSAT, i) = 0.0;
  }
  for (int e = 0; e < nt; ++e) {
    int mt = tetra[(size_t)e * 5 + 4] - 1;
    int istr = e / (3 * ntri);
    size_t ix = (size_t)istr * nzone + mt;
    for (int ii = 0; ii < 4; ++ii) {
      int nd = tetra[(size_t)e * 5 + ii];
      pnodi[nd] = pnodi[nd] + poros[ix];
      snodi[nd] = snodi[nd] + elstor[ix];
      VGF(vg, n, VG_N, nd) = VGF(vg, n, VG_N, nd) + vgncell[ix];
      VGF(vg, n, VG_RMC, nd) = VGF(vg, n, VG_RMC, nd) + vgrmccell[ix];
      VGF(vg, n, VG_PSAT, nd) = VGF(vg, n, VG_PSAT, nd) + vgpsatcell[ix];
      tp[nd] += 1;
    }
  }
  int bad = 0;
  for (int k = 0; k < n; ++k) {
    if (tp[k] == 0) { bad = 1; continue; }
    pnodi[k] = pnodi[k] / tp[k];
    snodi[k] = snodi[k] / tp[k];
    VGF(vg, n, VG_N, k) = VGF(vg, n, VG_N, k) / tp[k];
    VGF(vg, n, VG_PSAT, k) = VGF(vg, n, VG_PSAT, k) / tp[k];
    VGF(vg, n, VG_RMC, k) = VGF(vg, n, VG_RMC, k) / tp[k];
  }
  for (int e = 0; e < nt; ++e) {
    int mt = tetra[(size_t)e * 5 + 4] - 1;
    int istr = e / (3 * ntri);
    pel[e] = poros[(size_t)istr * nzone + mt];
  }
  return bad ? -1 : 0;
}

void orc_chparm(int n, const double *pnodi, double *vg) {
  /* SRC/chparm.f:24-41 (IVGHU=0). */
  for (int i = 0; i < n; ++i) {
    double vgn = VGF(vg, n, VG_N, i);
    double vgm = (vgn - 1.0) / vgn;
    VGF(vg, n, VG_M, i) = vgm;
    VGF(vg, n, VG_N1, i) = vgn - 1.0;
    VGF(vg, n, VG_NR, i) = 1.0 / vgn;
    VGF(vg, n, VG_MM1, i) = vgm - 1.0;
    VGF(vg, n, VG_PSN, i) = pow(fabs(VGF(vg, n, VG_PSAT, i)), vgn);
    VGF(vg, n, VG_MR, i) = 1.0 / vgm;
    VGF(vg, n, VG_M52, i) = 2.5 * vgm;
  }
  for (int i = 0; i < n; ++i)
    VGF(vg, n, VG_PNOT, i) = (pnodi[i] - VGF(vg, n, VG_RMC, i)) / pnodi[i];
}

void orc_eltnod(int n, int nt, const int *tp, const int *tetra,
                const double *velt, double *vnod) {
  /* SRC/eltnod.f. */
  for (int i = 0; i < n; ++i) vnod[i] = 0.0;
  for (int e = 0; e < nt; ++e)
    for (int ii = 0; ii < 4; ++ii) {
      int nd = tetra[(size_t)e * 5 + ii];
      vnod[nd] = vnod[nd] + velt[e];
    }
  for (int i = 0; i < n; ++i) vnod[i] = vnod[i] / tp[i];
}

void orc_nodelt(int nt, const int *tetra, const double *vnod, double *velt) {
  /* SRC/nodelt.f:9-29. */
  const double r4 = 1.0 / 4.0;
  for (int e = 0; e < nt; ++e) {
    double v = 0.0;
    for (int j = 0; j < 4; ++j) v = v + vnod[tetra[(size_t)e * 5 + j]];
    velt[e] = v * r4;
  }
}

/* ===================================================================== */
/*                          van Genuchten (a2)                           */
/* ===================================================================== */

double orc_fvgse(double psi, const double *vg, int n, int i) {
  /* SRC/fvgse.f:18-23. */
  if (psi < -1.0e-14) {
    double beta = pow(fabs(psi / VGF(vg, n, VG_PSAT, i)), VGF(vg, n, VG_N, i));
    return pow(fabs(1.0 / (beta + 1.0)), VGF(vg, n, VG_M, i));
  }
  return 1.0;
}

double orc_fvgkr(double psi, double se, const double *vg, int n, int i) {
  /* SRC/fvgkr.f:18-24. */
  if (psi < -1.0e-14) {
    double omega = pow(fabs(se), VGF(vg, n, VG_MR, i));
    double v1 = 1.0 - pow(fabs(1.0 - omega), VGF(vg, n, VG_M, i));
    return sqrt(se) * v1 * v1;
  }
  return 1.0;
}

double orc_fvgdse(double psi, const double *vg, int n, int i) {
  /* SRC/fvgdse.f:18-26. */
  if (psi < -1.0e-14) {
    double beta = pow(fabs(psi / VGF(vg, n, VG_PSAT, i)), VGF(vg, n, VG_N, i));
    double b1 = beta + 1.0;
    double b1r = 1.0 / b1;
    double vgn1 = VGF(vg, n, VG_N1, i);
    return vgn1 * (pow(fabs(psi), vgn1) / VGF(vg, n, VG_PSN, i)) *
           pow(fabs(b1), VGF(vg, n, VG_NR, i)) * b1r * b1r;
  }
  return 0.0;
}

void orc_chpic0(int n, const double *ptnew, const double *pnew,
                const double *ptimep, double *swnew, double *swtimep,
                const double *snodi, const double *pnodi, const double *vg,
                double *sw, double *ckrw, double *etai, double *et1,
                double *et2) {
  /* SRC/chpic0.f:23-36. */
  for (int i = 0; i < n; ++i) {
    double pnot = VGF(vg, n, VG_PNOT, i), rmc = VGF(vg, n, VG_RMC, i);
    swnew[i] = pnot * orc_fvgse(pnew[i], vg, n, i) + rmc / pnodi[i];
    swtimep[i] = pnot * orc_fvgse(ptimep[i], vg, n, i) + rmc / pnodi[i];
  }
  for (int i = 0; i < n; ++i) {
    double pnot = VGF(vg, n, VG_PNOT, i), rmc = VGF(vg, n, VG_RMC, i);
    double psi = ptnew[i];
    double se = orc_fvgse(psi, vg, n, i);
    sw[i] = pnot * se + rmc / pnodi[i];
    etai[i] = sw[i] * snodi[i] + pnodi[i] * pnot * orc_fvgdse(psi, vg, n, i);
    et1[i] = sw[i] * snodi[i];
    et2[i] = pnot * orc_fvgdse(psi, vg, n, i);
    ckrw[i] = orc_fvgkr(psi, se, vg, n, i);
  }
}

void orc_chvelo(int n, const double *pnew, const double *pnodi,
                const double *vg, double *sw, double *ckrw) {
  /* SRC/chvelo.f (IVGHU=0 branch). */
  for (int i = 0; i < n; ++i) {
    double psi = pnew[i];
    double se = orc_fvgse(psi, vg, n, i);
    sw[i] = VGF(vg, n, VG_PNOT, i) * se + VGF(vg, n, VG_RMC, i) / pnodi[i];
    ckrw[i] = orc_fvgkr(psi, se, vg, n, i);
  }
}

void orc_chmass(int n, const double *pnew, const double *snodi,
                const double *pnodi, const double *vg, double *etai) {
  /* SRC/chmass.f (IVGHU=0 branch). */
  for (int i = 0; i < n; ++i) {
    double psi = pnew[i];
    double se = orc_fvgse(psi, vg, n, i);
    double swi = VGF(vg, n, VG_PNOT, i) * se + VGF(vg, n, VG_RMC, i) / pnodi[i];
    etai[i] = swi * snodi[i] +
              pnodi[i] * VGF(vg, n, VG_PNOT, i) * orc_fvgdse(psi, vg, n, i);
  }
}

/* ===================================================================== */
/*                           assembly (a3-a7)                            */
/* ===================================================================== */

void orc_asspic(int nt, int ntri, int nzone, const int *tetra,
                const int *tetja, const double *lmass, double *coef1,
                double *coef2, double *coef4, const double *et1e,
                const double *pel, const double *permx, const double *permy,
                const double *permz, const double *ckrwe,
                const double *volu, const double *volur, const double *bi,
                const double *ci, const double *di) {
  /* SRC/asspic.f:26-51. */
  for (int e = 0; e < nt; ++e) {
    int istr = e / (3 * ntri);
    int mt = tetra[(size_t)e * 5 + 4] - 1;
    size_t ix = (size_t)istr * nzone + mt;
    double vk = volur[e] * ckrwe[e];
    double ve1 = volu[e] * et1e[e];
    double ve = volu[e] * pel[e];
    double pvk1 = permx[ix] * vk, pvk2 = permy[ix] * vk, pvk3 = permz[ix] * vk;
    const double *b = bi + (size_t)e * 4, *c = ci + (size_t)e * 4,
                 *d = di + (size_t)e * 4;
    for (int k = 0; k < 4; ++k) {
      double pvkb = pvk1 * b[k], pvkc = pvk2 * c[k], pvkd = pvk3 * d[k];
      for (int l = k; l < 4; ++l) {
        int ind = tetja[(size_t)e * 16 + k * 4 + l];
        coef1[ind] = coef1[ind] + pvkb * b[l] + pvkc * c[l] + pvkd * d[l];
        coef2[ind] = coef2[ind] + ve1 * lmass[k * 4 + l];
        coef4[ind] = coef4[ind] + ve * lmass[k * 4 + l];
      }
    }
  }
}

void orc_rhspic(int n, const int *ja, const int *topol, double deltat,
                const double *ptnew, const double *pnew,
                const double *ptimep, const double *swnew,
                const double *swtimep, const double *coef1,
                const double *coef2, const double *coef4, double *tnoti) {
  /* SRC/rhspic.f:24-37. */
  for (int k = 0; k < n; ++k) tnoti[k] = 0.0;
  double rdt = 1.0 / deltat;
  for (int k = 0; k < n; ++k) {
    int i = topol[k], j = topol[k + 1];
    for (int m = i; m < j; ++m) {
      int c = ja[m];
      double coef = coef2[m] * rdt, coeff = coef4[m] * rdt;
      tnoti[k] = tnoti[k] - coef1[m] * ptnew[c] - coef * (pnew[c] - ptimep[c]) -
                 coeff * (swnew[c] - swtimep[c]);
      if (m != i)
        tnoti[c] = tnoti[c] - coef1[m] * ptnew[k] -
                   coef * (pnew[k] - ptimep[k]) -
                   coeff * (swnew[k] - swtimep[k]);
    }
  }
}

void orc_cfmatp(int n, int nterm, double tetaf, double deltat, double *coef1,
                const double *coef2, const double *coef4, double *coef5,
                const int *topol, const double *et2) {
  /* SRC/cfmatp.f:21-26. */
  double rdt = 1.0 / deltat;
  for (int k = 0; k < n; ++k) coef5[topol[k]] = coef4[topol[k]] * et2[k];
  for (int k = 0; k < nterm; ++k)
    coef1[k] = tetaf * coef1[k] + coef2[k] * rdt + coef5[k] * rdt;
}

void orc_rhsgrv(int nt, int ntri, int nzone, const int *tetra,
                double *tnoti, const double *di, const double *permz,
                const int *ivol, const double *ckrwe) {
  /* SRC/rhsgrv.f:22-31. */
  for (int e = 0; e < nt; ++e) {
    int istr = e / (3 * ntri);
    int mt = tetra[(size_t)e * 5 + 4] - 1;
    double t3 = ckrwe[e] * permz[(size_t)istr * nzone + mt];
    for (int in = 0; in < 4; ++in) {
      int nd = tetra[(size_t)e * 5 + in];
      double t1 = t3 * di[(size_t)e * 4 + in] * (double)ivol[e];
      tnoti[nd] = tnoti[nd] - t1;
    }
  }
}

int orc_bcpic(int np, int nq, const int *contp, const int *contq,
              double *lhsp, const double *q, int nnod, int n,
              double *lhsatm, const int *ifatm, const double *atmact,
              const double *atmold, const int *topol, double *coef1,
              double *tnoti, double tetaf, double rmax, int *noddir,
              const double *qtranie) {
  /* SRC/bcpic.f:34-87 (NSF=0). */
  for (int k = 0; k < np; ++k) {
    int j = contp[k], ind = topol[j];
    tnoti[j] = 0.0;
    lhsp[k] = coef1[ind];
    coef1[ind] = 1.0e-9 * rmax;
  }
  int numdir = np;
  for (int k = 0; k < nnod; ++k)
    if (ifatm[k] == 1 || ifatm[k] == 2) {
      noddir[numdir++] = k;
      int ind = topol[k];
      tnoti[k] = 0.0;
      lhsatm[k] = coef1[ind];
      coef1[ind] = 1.0e-9 * rmax;
    }
  for (int k = 0; k < nq; ++k) tnoti[contq[k]] = tnoti[contq[k]] + q[k];
  for (int k = 0; k < nnod; ++k)
    if (ifatm[k] == 0)
      tnoti[k] = tnoti[k] + (tetaf * atmact[k] + (1.0 - tetaf) * atmold[k]);
  for (int k = 0; k < n; ++k) tnoti[k] = tnoti[k] - qtranie[k];
  return numdir;
}

void orc_shlpic(int np, const int *topol, const int *contp,
                const double *presc, const double *lhsp, double *coef1,
                double *pnew, const double *pold, int nnod, const int *ifatm,
                const double *lhsatm) {
  /* SRC/shlpic.f:30-58 (NSF=0). */
  for (int i = 0; i < np; ++i) {
    int j = contp[i];
    pnew[j] = presc[i];
    coef1[topol[j]] = lhsp[i];
  }
  for (int i = 0; i < nnod; ++i)
    if (ifatm[i] == 1 || ifatm[i] == 2) {
      pnew[i] = pold[i];
      coef1[topol[i]] = lhsatm[i];
    }
}

/* ===================================================================== */
/*                            IC(0)-PCG (a8-a11)                         */
/* ===================================================================== */

int orc_incldp(const int *ja, const int *topol, const double *coefr,
               double *lriga, int n, int nterm) {
  /* SRC/solscal-extended.f:2152-2190 (right-looking row IC(0)). */
  int nfall = 0;
  for (int k = 0; k < nterm; ++k) lriga[k] = 0.0;
  for (int kk = 0; kk < n; ++kk) {
    int k = topol[kk];
    double a = coefr[k] - lriga[k];
    if (a <= 0.0) {
      /* :2157-2160 breakdown fallback: previous diagonal squared */
      a = (kk > 0) ? lriga[topol[kk - 1]] * lriga[topol[kk - 1]] : 1.0;
      ++nfall;
    }
    lriga[k] = sqrt(a);
    if (kk == n - 1) continue;
    int i = topol[kk] + 1, j = topol[kk + 1] - 1;
    if (j < i) continue;
    for (int k1 = i; k1 <= j; ++k1)
      lriga[k1] = (coefr[k1] - lriga[k1]) / lriga[k];
    for (int k2 = i; k2 <= j; ++k2) {
      int row = ja[k2];
      int j1 = topol[row];
      lriga[j1] = lriga[j1] + lriga[k2] * lriga[k2];
      /* merge the rest of row kk (columns > ja[k2]) into row ja[k2] */
      int i1 = k2 + 1, jj = j1 + 1, end = topol[row + 1];
      while (i1 <= j && jj < end) {
        if (ja[jj] < ja[i1]) ++jj;
        else if (ja[jj] > ja[i1]) ++i1;
        else {
          lriga[jj] = lriga[jj] + lriga[k2] * lriga[i1];
          ++i1;
        }
      }
    }
  }
  return nfall;
}

void orc_proddp(double *x, const double *b, int n, const int *topol,
                const int *ja, const double *coef2) {
  /* SRC/solscal-extended.f:4175-4203: U^T y = b (column sweep), U x = y. */
  for (int k = 0; k < n; ++k) x[k] = 0.0;
  for (int k = 0; k < n; ++k) {
    int i = topol[k], j = topol[k + 1];
    for (int m = i; m < j; ++m) {
      if (m == i) x[k] = (b[k] - x[k]) / coef2[m];
      else x[ja[m]] = x[ja[m]] + coef2[m] * x[k];
    }
  }
  double a = 0.0;
  for (int kk = n - 1; kk >= 0; --kk) {
    int i = topol[kk], j = topol[kk + 1] - 1;
    for (int m = i; m <= j; ++m) {
      int mm = j - m + i;
      if (mm == i) {
        x[kk] = (x[kk] - a) / coef2[mm];
        a = 0.0;
      } else {
        a = a + coef2[mm] * x[ja[mm]];
      }
    }
  }
}

/* symmetric upper-CSR product y (+)= A x exactly as GRADDP's loops do */
int orc_graddp(int n, double *poten, double *scr, double *res, double *bv,
               double *r, double *p, const int *topol, const int *ja,
               const double *coef1, const double *coef2,
               const double *tnoti, double tol, int imax, int np,
               const int *contp, double *err) {
  /* SRC/solscal-extended.f:1275-1385. */
  int niter = 1;
  for (int k = 0; k < n; ++k) {
    scr[k] = tnoti[k];
    r[k] = tnoti[k];
  }
  for (int i = 0; i < np; ++i) scr[contp[i]] = 0.0;
  double xlung = 0.0;
  for (int i = 0; i < n; ++i) xlung = xlung + scr[i] * scr[i];
  for (int i = 0; i < np; ++i) scr[contp[i]] = tnoti[contp[i]];
  for (int k = 0; k < n; ++k) {
    int i = topol[k], j = topol[k + 1];
    for (int m = i; m < j; ++m) {
      r[k] = r[k] - coef1[m] * poten[ja[m]];
      res[k] = r[k];
      if (m != i) {
        r[ja[m]] = r[ja[m]] - coef1[m] * poten[k];
        res[ja[m]] = r[ja[m]];
      }
    }
  }
  orc_proddp(p, r, n, topol, ja, coef2);
  for (;;) {
    for (int k = 0; k < n; ++k) bv[k] = 0.0;
    for (int k = 0; k < n; ++k) {
      int i = topol[k], j = topol[k + 1];
      for (int m = i; m < j; ++m) {
        bv[k] = bv[k] + coef1[m] * p[ja[m]];
        if (m != i) bv[ja[m]] = bv[ja[m]] + coef1[m] * p[k];
      }
    }
    double a = 0.0, bb = 0.0;
    for (int k = 0; k < n; ++k) {
      a = a + p[k] * r[k];
      bb = bb + p[k] * bv[k];
    }
    double alfa = a / bb;
    for (int k = 0; k < n; ++k) {
      r[k] = r[k] - alfa * bv[k];
      poten[k] = poten[k] + alfa * p[k];
    }
    orc_proddp(scr, r, n, topol, ja, coef2);
    double c = 0.0;
    for (int k = 0; k < n; ++k) c = c + bv[k] * scr[k];
    double beta = -c / bb;
    for (int k = 0; k < n; ++k) p[k] = scr[k] + beta * p[k];
    for (int k = 0; k < n; ++k) res[k] = r[k];
    for (int k = 0; k < np; ++k) res[contp[k]] = 0.0;
    a = 0.0;
    for (int k = 0; k < n; ++k) a = a + res[k] * res[k];
    *err = (xlung > 0.0) ? sqrt(a / xlung) : sqrt(a / n);
    if (*err > tol && niter < imax) {
      ++niter;
      continue;
    }
    break;
  }
  return niter;
}

double orc_ressym(const int *topol, const int *ja, const double *coef1,
                  const double *tnoti, const double *pnew, double *resv,
                  int n, int numdir, const int *noddir, double rmin) {
  /* SRC/solscal-extended.f:4400-4437. */
  for (int k = 0; k < n; ++k) resv[k] = tnoti[k];
  for (int k = 0; k < numdir; ++k) resv[noddir[k]] = 0.0;
  double xlung = 0.0;
  for (int k = 0; k < n; ++k) xlung = xlung + resv[k] * resv[k];
  for (int k = 0; k < n; ++k) {
    int m = topol[k], mm = topol[k + 1];
    for (int i = m; i < mm; ++i) {
      if (i > m) resv[ja[i]] = resv[ja[i]] - coef1[i] * pnew[k];
      resv[k] = resv[k] - coef1[i] * pnew[ja[i]];
    }
  }
  for (int k = 0; k < numdir; ++k) resv[noddir[k]] = 0.0;
  double res = 0.0;
  for (int k = 0; k < n; ++k) res = res + resv[k] * resv[k];
  return (xlung <= rmin) ? sqrt(res / n) : sqrt(res / xlung);
}

int orc_symslv(int n, int nterm, int numdir, int itmxcg, double tolcg,
               double rmin, const int *noddir, const int *ja,
               const int *topol, double *pnew, const double *tnoti,
               const double *coef1, double *coef2, double *scr, double *err,
               double *res) {
  /* SRC/solscal-extended.f:4679-4697. */
  orc_incldp(ja, topol, coef1, coef2, n, nterm);
  orc_proddp(pnew, tnoti, n, topol, ja, coef2);
  int niter = orc_graddp(n, pnew, scr, scr + n, scr + 2 * (size_t)n,
                         scr + 3 * (size_t)n, scr + 4 * (size_t)n, topol, ja,
                         coef1, coef2, tnoti, tolcg, itmxcg, numdir, noddir,
                         err);
  *res = orc_ressym(topol, ja, coef1, tnoti, pnew, scr, n, numdir, noddir,
                    rmin);
  return niter;
}

/* ===================================================================== */
/*                         mass balance (a13-a14)                        */
/* ===================================================================== */

void orc_bkpic(int n, int np, int nnod, double tetaf, const int *topol,
               const int *ja, const int *contp, const int *ifatm,
               const double *pdiff, double *qpnew, const double *qpold,
               double *atmact, const double *atmold, const double *coef1,
               const double *xt5, double *scr) {
  /* SRC/bkpic.f:28-54 (NSF=0). */
  double tetafr = 1.0 / tetaf;
  for (int k = 0; k < n; ++k) scr[k] = 0.0;
  for (int k = 0; k < n; ++k) {
    int i = topol[k], j = topol[k + 1];
    for (int m = i; m < j; ++m) {
      int jm = ja[m];
      scr[k] = scr[k] + coef1[m] * pdiff[jm];
      if (m != i) scr[jm] = scr[jm] + coef1[m] * pdiff[k];
    }
    scr[k] = scr[k] - xt5[k];
  }
  for (int i = 0; i < np; ++i) {
    int k = contp[i];
    qpnew[i] = (scr[k] - (1.0 - tetaf) * qpold[i]) * tetafr;
  }
  for (int k = 0; k < nnod; ++k)
    if (ifatm[k] == 1 || ifatm[k] == 2)
      atmact[k] = (scr[k] - (1.0 - tetaf) * atmold[k]) * tetafr;
}

double orc_stormb(const double *pnew, const double *ptimep,
                  const double *swnew, const double *swtimep, int n,
                  const double *volnod, const double *snodi,
                  const double *pnodi) {
  /* SRC/stormb.f:20-25. */
  double dstore = 0.0;
  for (int k = 0; k < n; ++k)
    dstore = dstore +
             volnod[k] * (snodi[k] * (swnew[k] + swtimep[k]) * 0.5 *
                              (pnew[k] - ptimep[k]) +
                          pnodi[k] * (swnew[k] - swtimep[k]));
  return dstore;
}

int orc_norms(int n, double *pikmax, double *pinf, double *pl2, double *finf,
              double *fl2, const double *pnew, const double *pold,
              const double *tnoti) {
  /* SRC/norms.f:23-40 (ties keep the last index). */
  int ikmax = 0;
  double spl2 = 0.0, sfl2 = 0.0;
  *pinf = 0.0;
  *finf = 0.0;
  *pikmax = 0.0;
  for (int k = 0; k < n; ++k) {
    double diff = pnew[k] - pold[k];
    double difabs = fabs(diff);
    spl2 = spl2 + diff * diff;
    sfl2 = sfl2 + tnoti[k] * tnoti[k];
    if (difabs >= *pinf) {
      *pinf = difabs;
      *pikmax = diff;
      ikmax = k;
    }
    double fabs_ = fabs(tnoti[k]);
    if (fabs_ >= *finf) *finf = fabs_;
  }
  *pl2 = sqrt(spl2);
  *fl2 = sqrt(sfl2);
  return ikmax;
}

int orc_etran(int n, int nnod, int nstr, const double *atmpot,
              const double *z, const double *psi, const int *vegtype,
              const double *pcana, const double *pcref, const double *pcwlt,
              const double *zroot, const double *pz, const double *omgc,
              double scf, double *qtranie) {
  /* SRC/etran.f:15-80.  Returns 0, or -1 if a root zone reaches the base
   * of its column (the reference would read past the column there). */
  double *btrani = (double *)calloc((size_t)n, sizeof(double));
  double *btran = (double *)calloc((size_t)nnod, sizeof(double));
  double *etp = (double *)calloc((size_t)nnod, sizeof(double));
  double *omg = (double *)calloc((size_t)nnod, sizeof(double));
  int rc = 0;
  for (int k = 0; k < n; ++k) qtranie[k] = 0.0;
  for (int i = 0; i < nnod && rc == 0; ++i) {
    int v = vegtype[i];
    int j = 1;
    double zsurf = z[i], depth = 0.0;
    btran[i] = 0.0;
    omg[i] = 0.0;
    etp[i] = (atmpot[i] < 0.0) ? -1.0 * scf * atmpot[i] : 0.0;
    while (depth <= zroot[v]) {
      if (j > nstr) { rc = -1; break; }
      int k = (j - 1) * nnod + i;
      double s1 = pcana[v], s2 = pcana[v] + 1.0e-03;
      double dz = (j == 1) ? (zsurf - z[k + nnod]) / 2.0
                           : (z[k - nnod] - z[k + nnod]) / 2.0;
      double sh2o = psi[k];
      double gx1 = (sh2o - pcwlt[v]) / (pcref[v] - pcwlt[v]);
      gx1 = fmin(1.0, fmax(0.0, gx1));
      double gx2 = 1.0 - (sh2o - s1) / (s2 - s1);
      gx2 = fmin(1.0, fmax(0.0, gx2));
      double gx = fmin(gx1, gx2);
      double beta = (1.0 - depth / zroot[v]) *
                    exp(-1.0 * pz[v] * depth / zroot[v]);
      btrani[k] = fmax(0.0, beta * dz * gx);
      btran[i] = btran[i] + beta * dz;
      omg[i] = omg[i] + gx * beta * dz;
      j = j + 1;
      depth = zsurf - z[(j - 1) * nnod + i];
    }
    btran[i] = fmax(0.0, btran[i]);
    omg[i] = omg[i] / btran[i];
  }
  if (rc == 0)
    for (int i = 0; i < nnod; ++i)
      for (int j = 1; j <= nstr + 1; ++j) {
        int k = (j - 1) * nnod + i, v = vegtype[i];
        qtranie[k] = etp[i] * btrani[k] / btran[i] / fmax(omg[i], omgc[v]);
      }
  free(btrani); free(btran); free(etp); free(omg);
  return rc;
}

void orc_switch_old(int nnod, int *ifatm, double *atmact,
                    const double *atmpot, double *pnew, double pmin) {
  /* SRC/switch_old.f:22-52. */
  for (int i = 0; i < nnod; ++i) {
    if (ifatm[i] == -1) continue;
    if (ifatm[i] == 1 && pnew[i] >= 0.0) {
      if (atmpot[i] < 0.0 || atmact[i] > atmpot[i]) {
        ifatm[i] = 0;
        atmact[i] = atmpot[i];
        continue;
      }
    }
    if (ifatm[i] == 1 && pnew[i] <= pmin) {
      if (atmpot[i] > 0.0 || atmact[i] < atmpot[i]) {
        ifatm[i] = 0;
        atmact[i] = atmpot[i];
        continue;
      }
    }
    if (ifatm[i] == 0 && pnew[i] >= 0.0 && atmpot[i] >= 0.0) {
      ifatm[i] = 1;
      pnew[i] = 0.0;
      continue;
    }
    if (ifatm[i] == 0 && pnew[i] <= pmin && atmpot[i] < 0.0) {
      ifatm[i] = 1;
      pnew[i] = pmin;
      continue;
    }
  }
}

/* The band of a head h against PONDH_MIN (ph) and PMIN, tested in the
 * order SWITCH tests them: 1 h >= ph, 2 0 <= h < ph, 3 pmin < h < 0,
 * 4 h <= pmin; 0 none (NaN). */
static int sw_band(double h, double ph, double pmin) {
  if (h >= ph) return 1;
  if (h >= 0.0 && h < ph) return 2;
  if (h > pmin && h < 0.0) return 3;
  if (h <= pmin) return 4;
  return 0;
}

int orc_switch(int nnod, int *ifatm, double deltat, double pondh_min,
               double pmin, const double *arenod, const double *pondnod,
               const double *atmpot, double *atmact, const double *qtranie,
               double *pnew, double *ovflnod) {
  /* SRC/switch.f:27-423. */
  int ponding = 0;
  for (int i = 0; i < nnod; ++i) {
    const int fa = ifatm[i];
    if (fa == -1) {
      ovflnod[i] = 0.0;
      continue;
    }
    double dif = atmpot[i] - (atmact[i] - qtranie[i]);
    if (fabs(dif) < 1.0e-14) dif = 0.0;
    const double pl = pondnod[i] + (dif * deltat / arenod[i]);
    const double drain = -pondnod[i] * arenod[i] / deltat;
    const int rain = atmpot[i] >= 0.0, infil = atmact[i] >= 0.0;
    if (fa == 2 || fa == 1) {
      /* guards of cases G and H (IFATM 1, evaporation) */
      if (fa == 1 && !rain && pnew[i] <= pmin) {
        if (!infil && atmact[i] < atmpot[i]) { /* H */
          ifatm[i] = 0;
          atmact[i] = atmpot[i];
          pnew[i] = pmin;
          ovflnod[i] = 0.0;
        }
        continue;
      }
      /* cases B and F (rainfall, exfiltration) end at PL < 0 with IFATM 1;
       * cases C and D (IFATM 2, evaporation) zero PNEW in bands 2 and 3 */
      const int exf_rain = rain && !infil;
      const int zero_p = (fa == 2 && !rain);
      int b = sw_band(pl, pondh_min, pmin);
      if (exf_rain && b >= 3) b = 5;
      switch (b) {
        case 1:
          ponding = 1;
          ifatm[i] = 2;
          pnew[i] = pl;
          ovflnod[i] = dif;
          break;
        case 2:
          ifatm[i] = 1;
          if (zero_p) pnew[i] = 0.0;
          ovflnod[i] = dif;
          break;
        case 3:
          ifatm[i] = 0;
          atmact[i] = atmpot[i];
          if (zero_p) pnew[i] = 0.0;
          ovflnod[i] = drain;
          break;
        case 4:
          ifatm[i] = 0;
          atmact[i] = atmpot[i];
          ovflnod[i] = drain;
          break;
        case 5:
          ifatm[i] = 1;
          ovflnod[i] = dif;
          break;
        default:
          break;
      }
      continue;
    }
    if (fa == 0) { /* cases I and J: the same tests on PNEW */
      const double over = (pnew[i] - pondnod[i]) * arenod[i] / deltat;
      switch (sw_band(pnew[i], pondh_min, pmin)) {
        case 1:
          ponding = 1;
          ifatm[i] = 2;
          ovflnod[i] = over;
          break;
        case 2:
          ifatm[i] = 1;
          ovflnod[i] = over;
          break;
        case 3:
          ifatm[i] = 0;
          ovflnod[i] = drain;
          break;
        case 4:
          ifatm[i] = 1;
          pnew[i] = pmin;
          ovflnod[i] = drain;
          break;
        default:
          break;
      }
    }
  }
  return ponding;
}

void orc_adrstn(int nnod, int *ifatm, const double *atmpot, double *atmact,
                double *pnew, double pmin) {
  /* SRC/adrstn.f:22-31. */
  for (int i = 0; i < nnod; ++i)
    if (pnew[i] <= pmin && (atmpot[i] > 0.0 || atmact[i] < atmpot[i])) {
      ifatm[i] = 0;
      atmact[i] = atmpot[i];
      pnew[i] = pmin;
    }
}

int orc_timupd(double *timep, double *time, double *deltat, double dtmin,
               double dtmax, double tmax, double dtmaga, double dtmagm,
               double dtreds, double dtredm, int iter, int ituns1,
               int ituns2) {
  /* SRC/timupd.f:21-40. */
  *timep = *time;
  if (iter < ituns1) {
    *deltat = *deltat * dtmagm + dtmaga;
    if (*deltat > dtmax) *deltat = dtmax;
  }
  if (iter >= ituns2) {
    *deltat = *deltat * dtredm - dtreds;
    if (*deltat < dtmin) *deltat = dtmin;
  }
  if ((*time + *deltat) >= tmax) {
    *deltat = tmax - *time;
    *time = tmax;
  } else {
    if ((*time + 2 * *deltat) > tmax) *deltat = (tmax - *time) / 2;
    *time = *time + *deltat;
  }
  return (*deltat <= dtmin) ? 0 : 1;
}

void orc_icvdwt(int nnod, int nstr, double wtposition, double *ptimep,
                const double *z) {
  /* SRC/icvdwt.f:20-25. */
  for (int i = 0; i < nnod; ++i)
    for (int k = 0; k <= nstr; ++k) {
      size_t kk = (size_t)k * nnod + i;
      ptimep[kk] = z[i] - z[kk] - wtposition;
    }
}

void orc_neumann_free(int nnod, const int *acontq, const double *arenod,
                      const double *ckrw, const double *kznod, double *q) {
  /* SRC/neumann.f:19-24. */
  for (int i = 0; i < nnod; ++i) {
    int nn = acontq[i];
    q[i] = -1.0 * arenod[i] * ckrw[nn] * kznod[nn];
  }
}

/* ===================================================================== */
/*                            whole run driver                           */
/* ===================================================================== */
/*              Newton linearisation (IOPT=2) and BiCGSTAB (a16)         */
/* ===================================================================== */

double orc_fvgdds(double psi, const double *vg, int n, int i) {
  /* SRC/fvgdds.f: second derivative of Se. */
  if (psi < -1.0e-14) {
    double vgn = VGF(vg, n, VG_N, i);
    double beta = pow(fabs(psi / VGF(vg, n, VG_PSAT, i)), vgn);
    double b1 = beta + 1.0;
    double b1r = 1.0 / b1;
    return VGF(vg, n, VG_N1, i) * (beta / psi) * (1.0 / psi) *
           ((1.0 + vgn * (beta - 1.0)) / pow(fabs(b1), VGF(vg, n, VG_M, i))) *
           b1r * b1r;
  }
  return 0.0;
}

double orc_fvgdkr(double psi, const double *vg, int n, int i) {
  /* SRC/fvgdkr.f: d kr / d psi. */
  if (psi < -1.0e-14) {
    double psat = VGF(vg, n, VG_PSAT, i), m = VGF(vg, n, VG_M, i);
    double beta = pow(fabs(psi / psat), VGF(vg, n, VG_N, i));
    double b1 = beta + 1.0;
    double v1 = pow(fabs(b1), m) - pow(fabs(beta), m);
    double v2 = psat / psi;
    double v3 = VGF(vg, n, VG_N1, i) * beta * v2 *
                pow(fabs(1.0 / b1), VGF(vg, n, VG_M52, i)) / psat;
    double v4 = v2 * ((2.5 / b1) * beta - 2.0) -
                0.5 * pow(fabs(b1), VGF(vg, n, VG_MM1, i));
    return v3 * v1 * v4;
  }
  return 0.0;
}

void orc_chnew0(int n, const double *ptnew, const double *snodi,
                const double *pnodi, const double *vg, double *sw,
                double *ckrw, double *etai, double *dckrw, double *detai) {
  /* SRC/chnew0.f (IVGHU=0 branch). */
  for (int i = 0; i < n; ++i) {
    double psi = ptnew[i];
    double pnot = VGF(vg, n, VG_PNOT, i);
    double se = orc_fvgse(psi, vg, n, i);
    double dsedp = orc_fvgdse(psi, vg, n, i);
    double dswdp = pnot * dsedp;
    sw[i] = pnot * se + VGF(vg, n, VG_RMC, i) / pnodi[i];
    etai[i] = sw[i] * snodi[i] + pnodi[i] * dswdp;
    detai[i] = dswdp * snodi[i] + pnodi[i] * pnot * orc_fvgdds(psi, vg, n, i);
    ckrw[i] = orc_fvgkr(psi, se, vg, n, i);
    dckrw[i] = orc_fvgdkr(psi, vg, n, i);
  }
}

void orc_assnew(int nt, int ntri, int nzone, const int *tetra,
                const int *tetja, const double *lmass, double *coef1,
                double *coef2, double *coef3, const double *permx,
                const double *permy, const double *permz, const double *ckrwe,
                const double *etae, const double *ptnew, const double *ptimep,
                const double *pnew, const double *dckrw, const double *detai,
                double tetaf, double deltat, const double *volu,
                const double *volur, const int *ivol, const double *bi,
                const double *ci, const double *di) {
  /* SRC/assnew.f: stiffness (COEF1), storage (COEF2) and the Newton
   * derivative terms (COEF3) on the full pattern; tetja[e*16+k*4+l] is the
   * position of (TETRA(k), TETRA(l)) in row TETRA(k). */
  double rdt = 1.0 / deltat;
  for (int e = 0; e < nt; ++e) {
    int istr = e / (3 * ntri);
    int mt = tetra[(size_t)e * 5 + 4] - 1;
    size_t ix = (size_t)istr * nzone + mt;
    const int *t = tetra + (size_t)e * 5;
    const int *tj = tetja + (size_t)e * 16;
    const double *b = bi + (size_t)e * 4, *c = ci + (size_t)e * 4,
                 *d = di + (size_t)e * 4;
    double ve = volu[e] * etae[e];
    double pv1 = permx[ix] * volur[e], pv2 = permy[ix] * volur[e],
           pv3 = permz[ix] * volur[e];
    double pvk1 = pv1 * ckrwe[e], pvk2 = pv2 * ckrwe[e], pvk3 = pv3 * ckrwe[e];
    double tpi = tetaf * permz[ix] * (double)ivol[e];
    double tvd = tetaf * volu[e] * rdt;
    for (int k = 0; k < 4; ++k) {
      int knod = t[k];
      double pvb = pv1 * b[k], pvc = pv2 * c[k], pvd = pv3 * d[k];
      double pvkb = pvk1 * b[k], pvkc = pvk2 * c[k], pvkd = pvk3 * d[k];
      double sum = 0.0, sum1 = 0.0;
      for (int m = 0; m < 4; ++m) {
        int mnod = t[m];
        sum = sum + (pvb * b[m] + pvc * c[m] + pvd * d[m]) * ptnew[mnod];
        sum1 = sum1 + lmass[k * 4 + m] * (pnew[mnod] - ptimep[mnod]);
      }
      double tsumtd = tetaf * sum + tpi * d[k];
      double sum1tv = sum1 * tvd;
      for (int l = 0; l < 4; ++l) {
        int lnod = t[l];
        int ind = tj[k * 4 + l];
        if (lnod >= knod) {
          coef1[ind] = coef1[ind] + pvkb * b[l] + pvkc * c[l] + pvkd * d[l];
          coef2[ind] = coef2[ind] + ve * lmass[k * 4 + l];
        }
        coef3[ind] = coef3[ind] + dckrw[lnod] * tsumtd + detai[lnod] * sum1tv;
      }
    }
    for (int k = 0; k < 4; ++k)
      for (int l = 0; l < 4; ++l)
        if (t[l] < t[k]) {
          coef1[tj[k * 4 + l]] = coef1[tj[l * 4 + k]];
          coef2[tj[k * 4 + l]] = coef2[tj[l * 4 + k]];
        }
  }
}

void orc_rhsnew(int n, const int *ja, const int *topol, double deltat,
                const double *ptnew, const double *pnew, const double *ptimep,
                const double *coef1, const double *coef2, double *tnoti) {
  /* SRC/rhsnew.f: row sums over the full pattern. */
  double rdt = 1.0 / deltat;
  for (int k = 0; k < n; ++k) {
    tnoti[k] = 0.0;
    for (int m = topol[k]; m < topol[k + 1]; ++m) {
      int c = ja[m];
      tnoti[k] = tnoti[k] - coef1[m] * ptnew[c] -
                 coef2[m] * (pnew[c] - ptimep[c]) * rdt;
    }
  }
}

void orc_cfmatn(int nterm, double tetaf, double deltat, double *coef1,
                const double *coef2, const double *coef3) {
  /* SRC/cfmatn.f. */
  double rdt = 1.0 / deltat;
  for (int k = 0; k < nterm; ++k)
    coef1[k] = tetaf * coef1[k] + coef2[k] * rdt + coef3[k];
}

static int diag_pos(const int *topol, const int *ja, int j) {
  int ind = topol[j];
  while (ja[ind] != j) ++ind;
  return ind;
}

int orc_bcnew(int nq, const int *contq, const double *q, int nnod,
              double *lhsatm, const int *ifatm, const double *atmact,
              const double *atmold, const int *ja, const int *topol,
              double *coef1, double *tnoti, double tetaf, double rmax,
              int *noddir) {
  /* SRC/bcnew.f with NP=0, NSF=0: penalty diagonals on the atmospheric
   * Dirichlet nodes, Neumann and atmospheric fluxes. */
  int numdir = 0;
  for (int k = 0; k < nnod; ++k)
    if (ifatm[k] == 1 || ifatm[k] == 2) {
      noddir[numdir++] = k;
      tnoti[k] = 0.0;
      int ind = diag_pos(topol, ja, k);
      lhsatm[k] = coef1[ind];
      coef1[ind] = 1.0e-9 * rmax;
    }
  for (int k = 0; k < nq; ++k) tnoti[contq[k]] = tnoti[contq[k]] + q[k];
  for (int k = 0; k < nnod; ++k)
    if (ifatm[k] == 0)
      tnoti[k] = tnoti[k] + (tetaf * atmact[k] + (1.0 - tetaf) * atmold[k]);
  return numdir;
}

void orc_shlnew(const int *topol, const int *ja, double *coef1, double *pnew,
                const double *pold, int nnod, const int *ifatm,
                const double *lhsatm) {
  /* SRC/shlnew.f with NP=0, NSF=0. */
  for (int i = 0; i < nnod; ++i)
    if (ifatm[i] == 1 || ifatm[i] == 2) {
      pnew[i] = pold[i];
      coef1[diag_pos(topol, ja, i)] = lhsatm[i];
    }
}

void orc_bknew(int nnod, double tetaf, const int *topol, const int *ja,
               const int *ifatm, const double *pdiff, double *atmact,
               const double *atmold, const double *coef1, const double *xt5) {
  /* SRC/bknew.f with NP=0, NSF=0: fluxes at the atmospheric Dirichlet
   * nodes from the (restored) Newton matrix. */
  double tetafr = 1.0 / tetaf;
  for (int k = 0; k < nnod; ++k)
    if (ifatm[k] == 1 || ifatm[k] == 2) {
      double v = 0.0;
      for (int m = topol[k]; m < topol[k + 1]; ++m) v = v + coef1[m] * pdiff[ja[m]];
      v = v - xt5[k];
      atmact[k] = (v - (1.0 - tetaf) * atmold[k]) * tetafr;
    }
}

void orc_topia(int n, const int *topol, int *ia) {
  /* SRC/topia.f: row index of every entry. */
  for (int k = 0; k < n; ++k)
    for (int m = topol[k]; m < topol[k + 1]; ++m) ia[m] = k;
}

int orc_inclu(const int *ia, const int *ja, const int *topol,
              const double *coefr, double *coefc, double *lriga,
              double *lcolon, int n, int nterm) {
  /* SRC/solscal-extended.f:2194-2270 INCLU: Crout ILU(0) on the full
   * (structurally symmetric, column-sorted) pattern.  On return LRIGA
   * holds, row by row, L (lower slots and the diagonal) and the unit upper
   * factor U (upper slots); LCOLON holds the transposes.  COEFC receives
   * A^T in A's slots.  Row r is computed after its transposes were copied
   * in from the rows above: L(c,r) = A(c,r) - sum_m L(c,m) U(m,r) into
   * LCOLON(r,c) and U(r,c) = (A(r,c) - sum_m L(r,m) U(m,c)) / L(r,r) into
   * LRIGA(r,c), sums over common columns m < r in ascending order.  A zero
   * pivot takes the previous row's.  Returns the number of zero pivots. */
  (void)n;
  int nfall = 0;
  for (int k = 0; k < nterm; ++k) {
    int i = topol[ja[k]];
    while (ja[i] < ia[k]) ++i;
    coefc[i] = coefr[k];
  }
  int k = 0;
  do { /* first row: column of L and row of U directly */
    lcolon[k] = coefc[k];
    lriga[k] = coefr[k] / coefr[0];
    if (k == 0) lriga[0] = coefr[0];
    ++k;
  } while (k < nterm && ia[k] == ia[k - 1]);
  while (k < nterm) {
    /* transposes of the previous row's upper entries into later rows */
    for (int i = k - 1; ja[i] > ia[i]; --i) {
      int j = topol[ja[i]];
      while (ja[j] < ia[i]) ++j;
      lriga[j] = lcolon[i];
      lcolon[j] = lriga[i];
    }
    while (ia[k] > ja[k]) ++k; /* to the diagonal of the new row */
    int kk = k;
    int r = ia[k];
    for (;;) {
      int j = topol[ja[k]];
      int i = topol[r];
      double a = 0.0, b = 0.0;
      while (!(ja[i] >= r || ja[j] >= r)) {
        if (ja[i] < ja[j]) {
          ++i;
        } else if (ja[i] == ja[j]) {
          a = a + lcolon[i] * lriga[j];
          if (i != j) b = b + lriga[i] * lcolon[j];
          ++i;
          ++j;
          if (i > nterm - 1) break;
        } else {
          ++j;
        }
      }
      lcolon[k] = coefc[k] - a;
      if (k == kk && lcolon[k] == 0.0) {
        int kkk = kk;
        do --kkk; while (ia[kkk] != ja[kkk]);
        lcolon[k] = lcolon[kkk];
        ++nfall;
      }
      if (ia[k] < ja[k]) lriga[k] = (coefr[k] - b) / lriga[kk];
      if (ia[k] == ja[k]) lriga[k] = lcolon[k];
      ++k;
      if (k >= nterm) return nfall;
      if (ia[k] != ia[k - 1]) break;
    }
  }
  return nfall;
}

void orc_prodl(double *z, const double *y, const int *ia, const int *ja,
               const double *coef2, int nterm) {
  /* SRC/solscal-extended.f:4205-4227 PRODL: L z = y. */
  z[0] = y[0] / coef2[0];
  double a = 0.0;
  for (int k = 1; k < nterm; ++k) {
    if (ja[k] < ia[k]) {
      a = a + coef2[k] * z[ja[k]];
    } else if (ja[k] == ia[k]) {
      z[ia[k]] = (y[ia[k]] - a) / coef2[k];
      a = 0.0;
    }
  }
}

void orc_produ(double *x, const double *h, int n, const int *ia,
               const int *ja, const double *coef2, const int *topol) {
  /* SRC/solscal-extended.f:4228-4250 PRODU: U x = h (unit diagonal). */
  x[n - 1] = h[n - 1];
  double a = 0.0;
  for (int i = topol[n - 1] - 1; i >= 0; --i) {
    if (ia[i] < ja[i]) {
      a = a + coef2[i] * x[ja[i]];
    } else if (ia[i] == ja[i]) {
      x[ia[i]] = h[ia[i]] - a;
      a = 0.0;
    }
  }
}

void orc_pru(double *y, const double *x, const int *ia, const int *ja,
             const double *coef2, int nterm) {
  /* SRC/solscal-extended.f:4251-4264 PRU: y = U x. */
  for (int k = 0; k < nterm; ++k) {
    if (ia[k] > ja[k]) continue;
    if (ia[k] == ja[k]) {
      y[ia[k]] = x[ia[k]];
      continue;
    }
    y[ia[k]] = y[ia[k]] + coef2[k] * x[ja[k]];
  }
}

void orc_prod1(double *x, const double *b, int n, const int *ia,
               const int *ja, const double *coef2, int nterm,
               const int *topol) {
  /* SRC/solscal-extended.f:4107-4141 PROD1: x = (LU)^-1 b. */
  orc_prodl(x, b, ia, ja, coef2, nterm);
  double a = 0.0;
  for (int i = topol[n - 1] - 1; i >= 0; --i) {
    if (ia[i] < ja[i]) {
      a = a + coef2[i] * x[ja[i]];
    } else if (ia[i] == ja[i]) {
      x[ia[i]] = x[ia[i]] - a;
      a = 0.0;
    }
  }
}

static void aperbn(int n, const int *topol, const int *ja, const double *coef,
                   const double *v, double *r) {
  /* SRC/solscal-extended.f:97-107 APERBN: r = A v. */
  for (int k = 0; k < n; ++k) {
    r[k] = 0.0;
    for (int i = topol[k]; i < topol[k + 1]; ++i) r[k] = r[k] + coef[i] * v[ja[i]];
  }
}

static void prodbh(double *z, const double *h, int n, const int *ia,
                   const int *ja, const double *coef1, const double *coef2,
                   int nterm, const int *topol, double *x, double *y) {
  /* SRC/solscal-extended.f:4142-4168 PRODBH: z = L^-1 A U^-1 h. */
  orc_produ(x, h, n, ia, ja, coef2, topol);
  for (int k = 0; k < n; ++k) {
    z[k] = 0.0;
    y[k] = 0.0;
  }
  for (int k = 0; k < nterm; ++k) y[ia[k]] = y[ia[k]] + coef1[k] * x[ja[k]];
  orc_prodl(z, y, ia, ja, coef2, nterm);
}

double orc_resnsy(const int *topol, const int *ja, const double *coef1,
                  const double *tnoti, const double *pnew, double *resv,
                  int n, int numdir, const int *noddir, double rmin) {
  /* SRC/solscal-extended.f:4332-4388 RESNSY. */
  for (int k = 0; k < n; ++k) resv[k] = tnoti[k];
  for (int k = 0; k < numdir; ++k) resv[noddir[k]] = 0.0;
  double xlung = 0.0;
  for (int k = 0; k < n; ++k) xlung = xlung + resv[k] * resv[k];
  for (int k = 0; k < n; ++k)
    for (int i = topol[k]; i < topol[k + 1]; ++i)
      resv[k] = resv[k] - coef1[i] * pnew[ja[i]];
  for (int k = 0; k < numdir; ++k) resv[noddir[k]] = 0.0;
  double res = 0.0;
  for (int k = 0; k < n; ++k) res = res + resv[k] * resv[k];
  return xlung <= rmin ? sqrt(res / n) : sqrt(res / xlung);
}

int orc_gcstas(int n, double *poten, const double *tnoti, double *scr,
               const int *topol, const int *ia, const int *ja,
               const double *coef1, const double *coef2, int nterm, int imax,
               double tol, int np, const int *contp, double *err_out) {
  /* SRC/solscal-extended.f:1010-1128 GCSTAS: BiCGSTAB on L^-1 A U^-1 with
   * y = U x, from x0 = poten; the relative error is that of the
   * preconditioned residual with the Dirichlet entries zeroed, over
   * ||b_free||.  scr: 10 n-vectors. */
  double *pjm = scr, *vjm = scr + n, *p = scr + 2 * (size_t)n,
         *v = scr + 3 * (size_t)n, *s = scr + 4 * (size_t)n,
         *t = scr + 5 * (size_t)n, *r = scr + 6 * (size_t)n,
         *rt = scr + 7 * (size_t)n, *scr1 = scr + 8 * (size_t)n,
         *scr2 = scr + 9 * (size_t)n;
  int niter = 1;
  double xlung = 0.0, err = 0.0;
  for (int i = 0; i < n; ++i) pjm[i] = tnoti[i];
  for (int i = 0; i < np; ++i) pjm[contp[i]] = 0.0;
  for (int i = 0; i < n; ++i) xlung = xlung + pjm[i] * pjm[i];
  aperbn(n, topol, ja, coef1, poten, scr1);
  for (int k = 0; k < n; ++k) scr1[k] = tnoti[k] - scr1[k];
  orc_prodl(r, scr1, ia, ja, coef2, nterm);
  for (int k = 0; k < n; ++k) rt[k] = r[k];
  orc_pru(scr1, poten, ia, ja, coef2, nterm);
  for (int k = 0; k < n; ++k) poten[k] = scr1[k];
  for (int i = 0; i < n; ++i) {
    pjm[i] = 0.0;
    vjm[i] = 0.0;
  }
  double rojm1 = 1.0, alfjm1 = 1.0, etaj = 1.0;
  for (;;) {
    double roj = 0.0;
    for (int i = 0; i < n; ++i) roj = roj + rt[i] * r[i];
    double beta = (roj / rojm1) * (alfjm1 / etaj);
    for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (pjm[i] - etaj * vjm[i]);
    prodbh(v, p, n, ia, ja, coef1, coef2, nterm, topol, scr1, scr2);
    double sigma = 0.0;
    for (int i = 0; i < n; ++i) sigma = sigma + rt[i] * v[i];
    double alfj = roj / sigma;
    for (int i = 0; i < n; ++i) s[i] = r[i] - alfj * v[i];
    prodbh(t, s, n, ia, ja, coef1, coef2, nterm, topol, scr1, scr2);
    double aa = 0.0, bb = 0.0;
    for (int i = 0; i < n; ++i) {
      aa = aa + s[i] * t[i];
      bb = bb + t[i] * t[i];
    }
    etaj = aa / bb;
    for (int i = 0; i < n; ++i) {
      poten[i] = poten[i] + alfj * p[i] + etaj * s[i];
      r[i] = s[i] - etaj * t[i];
    }
    for (int i = 0; i < np; ++i) r[contp[i]] = 0.0;
    aa = 0.0;
    for (int i = 0; i < n; ++i) aa = aa + r[i] * r[i];
    err = xlung > 0.0 ? sqrt(aa / xlung) : sqrt(aa / n);
    if (niter <= imax && err > tol) {
      niter = niter + 1;
      rojm1 = roj;
      alfjm1 = alfj;
      for (int i = 0; i < n; ++i) {
        pjm[i] = p[i];
        vjm[i] = v[i];
      }
      continue;
    }
    break;
  }
  for (int i = 0; i < n; ++i) {
    r[i] = poten[i];
    poten[i] = 0.0;
  }
  orc_produ(poten, r, n, ia, ja, coef2, topol);
  *err_out = err;
  return niter;
}

int orc_nsyslv(int n, int nterm, int numdir, const int *noddir, int itmxcg,
               double tolcg, double rmin, const int *ia, const int *ja,
               const int *topol, const double *coef1, double *coef2,
               double *coef3, double *scr, double *pnew, const double *tnoti,
               double *err, double *res, double *resini) {
  /* SRC/solscal-extended.f:3063-3174 NSYSLV, ISOLV=0: INCLU, x0 = (LU)^-1 b
   * (PROD1), GCSTAS, RESNSY.  scr: max(nterm, 11 n) doubles (LCOLON, then
   * the BiCGSTAB vectors).  coef3 is overwritten (A^T).  Returns NITER. */
  double rhsn = 0.0;
  for (int i = 0; i < n; ++i) rhsn = rhsn + tnoti[i] * tnoti[i];
  if (rhsn == 0.0) {
    for (int i = 0; i < n; ++i) pnew[i] = 0.0;
    *err = 0.0;
    *res = 0.0;
    *resini = 0.0;
    return 0;
  }
  orc_inclu(ia, ja, topol, coef1, coef3, coef2, scr, n, nterm);
  orc_prod1(pnew, tnoti, n, ia, ja, coef2, nterm, topol);
  *resini = orc_resnsy(topol, ja, coef1, tnoti, pnew, scr, n, numdir, noddir, rmin);
  int niter = orc_gcstas(n, pnew, tnoti, scr, topol, ia, ja, coef1, coef2, nterm,
                         itmxcg, tolcg, numdir, noddir, err);
  *res = orc_resnsy(topol, ja, coef1, tnoti, pnew, scr, n, numdir, noddir, rmin);
  return niter;
}

/* ===================================================================== */

struct orc_run {
  orc_params p;
  int nnod, nstr, n, ntri, nt, nterm, nzone;
  int *tetra, *topol, *ja, *tetja, *ivol, *tp;
  double *x, *y, *z, *ai, *bi, *ci, *di, *volu, *volur, *volnod, *arenod;
  double lmass[16];
  double *permx, *permy, *permz;
  double *pnodi, *snodi, *vg, *pel, *kznod;
  /* nodal state */
  double *pnew, *pold, *ptnew, *ptold, *ptimep, *pdiff;
  double *sw, *ckrw, *ckrwp, *etai, *et1, *et2, *swnew, *swtimep;
  double *ckrwe, *etae, *et1e;
  double *coef1, *coef2, *coef4, *coef5, *tnoti, *xt5, *scr, *qtranie;
  /* Newton (IOPT=2): full pattern row ids, derivative terms */
  int *ia;
  double *coef3, *dckrw, *detai;
  /* boundary conditions */
  int nq;
  int *contq;
  double *q;
  int *ifatm, *ifatmp, *noddir;
  double *atmpot, *atmact, *atmold, *lhsatm;
  int numdir;
  /* atmbc series (file replaced by arrays) */
  int n_atm, hspatm, atm_next, htiatm;
  double *atm_times, *atm_vals;
  double atmtim[3];
  double *atminp; /* [3][nnod] */
  /* time */
  double time, timep, deltat;
  int dtgmin, nstep, iter, nitert, itlin, itrtot, kbackt, klsfai;
  /* mass balance */
  double store1, store2, dstore, cdstor, cvin, cvout, vtot, cerras, caeras;
  double adinp, adoutp, ndinp, ndoutp, aninp, anoutp, nninp, nnoutp;
  double vin, vout, erras, errel;
  double adin, adout, ndin, ndout, anin, anout, nnin, nnout;
  double pinf, pl2, finf, fl2, pikmax;
  int lsfail, noback, bckstp, finished;
};

static void *xdup(const void *src, size_t bytes) {
  void *d = malloc(bytes ? bytes : 1);
  if (d && src) memcpy(d, src, bytes);
  return d;
}
static double *dalloc(size_t k) { return (double *)calloc(k ? k : 1, sizeof(double)); }
static int *ialloc(size_t k) { return (int *)calloc(k ? k : 1, sizeof(int)); }
#define DALLOC(k) dalloc((size_t)(k))
#define IALLOC(k) ialloc((size_t)(k))

/* ATMONE/ATMNXT file reader restated over arrays: the next record. */
static int atm_read(orc_run *r, double *timein, double *vals) {
  if (r->atm_next >= r->n_atm) return 0; /* END= */
  int k = r->atm_next++;
  *timein = r->atm_times[k];
  if (r->hspatm == 0)
    for (int i = 0; i < r->nnod; ++i)
      vals[i] = r->atm_vals[(size_t)k * r->nnod + i];
  else
    for (int i = 0; i < r->nnod; ++i) vals[i] = r->atm_vals[k];
  return 1;
}

static void atm_interp(orc_run *r, int set_act) {
  /* SRC/atmnxt.f:41-73 (label 300). */
  double *a2 = r->atminp + r->nnod, *a3 = r->atminp + 2 * (size_t)r->nnod;
  for (int i = 0; i < r->nnod; ++i) {
    if (r->atmtim[2] > r->atmtim[1]) {
      double slope = (a3[i] - a2[i]) / (r->atmtim[2] - r->atmtim[1]);
      if (r->p.ieto != 0) slope = 0.0;
      r->atmpot[i] = (a2[i] + slope * (r->time - r->atmtim[1])) * r->arenod[i];
    } else {
      r->atmpot[i] = a3[i] * r->arenod[i];
    }
    if (set_act && r->ifatm[i] == 0)
      r->atmact[i] = (r->atmpot[i] >= 0.0) ? r->atmpot[i]
                                           : (1.0 - r->p.scf) * r->atmpot[i];
  }
}

static void atm_shift_read(orc_run *r) {
  /* the label-200 loop of SRC/atmnxt.f:30-40 / SRC/atmone.f:77-93 */
  double *a1 = r->atminp, *a2 = a1 + r->nnod, *a3 = a2 + r->nnod;
  while (!(r->time <= r->atmtim[2])) {
    r->atmtim[0] = r->atmtim[1];
    r->atmtim[1] = r->atmtim[2];
    for (int i = 0; i < r->nnod; ++i) {
      a1[i] = a2[i];
      a2[i] = a3[i];
    }
    double tin;
    if (!atm_read(r, &tin, a3)) {
      r->htiatm = 1;
      break;
    }
    r->atmtim[2] = tin;
  }
}

static void atmnxt(orc_run *r) {
  /* SRC/atmnxt.f:9-104 (no Dirichlet/Neumann surface nodes, NSF=0). */
  if (r->htiatm != 0) return;
  atm_shift_read(r);
  atm_interp(r, 1);
}

static void atmbak(orc_run *r) {
  /* SRC/atmbak.f. */
  if (r->atmtim[0] >= r->atmtim[1]) return;
  double *a1 = r->atminp, *a2 = a1 + r->nnod;
  for (int i = 0; i < r->nnod; ++i) {
    double slope = (a2[i] - a1[i]) / (r->atmtim[1] - r->atmtim[0]);
    if (r->p.ieto != 0) slope = 0.0;
    r->atmpot[i] = (a1[i] + slope * (r->time - r->atmtim[0])) * r->arenod[i];
    if (r->ifatm[i] == 0)
      r->atmact[i] = (r->atmpot[i] >= 0.0) ? r->atmpot[i]
                                           : (1.0 - r->p.scf) * r->atmpot[i];
  }
}

static void atmone(orc_run *r) {
  /* SRC/atmone.f:39-163 (HSPATM read, NP=NQ=0 on the surface). */
  int nnod = r->nnod;
  double *a3 = r->atminp + 2 * (size_t)nnod;
  r->htiatm = 0;
  r->atmtim[0] = r->atmtim[1] = 0.0;
  for (int i = 0; i < nnod; ++i) {
    r->ifatm[i] = 0;
    r->ifatmp[i] = 0;
    r->atminp[i] = 0.0;
    r->atminp[nnod + i] = 0.0;
    r->atmpot[i] = r->atmact[i] = r->atmold[i] = 0.0;
  }
  double tin;
  if (!atm_read(r, &tin, a3)) {
    r->htiatm = 1;
  } else {
    r->atmtim[2] = tin;
    if (r->atmtim[2] <= 0.0)
      for (int i = 0; i < nnod; ++i) r->atmold[i] = a3[i] * r->arenod[i];
    atm_shift_read(r);
    atm_interp(r, 0);
  }
  for (int i = 0; i < nnod; ++i) {
    if (r->ifatm[i] == -1) continue;
    if (r->pnew[i] >= r->p.pondh_min) {
      r->ifatm[i] = 2;
      r->ifatmp[i] = 2;
      continue;
    }
    if (r->pnew[i] >= 0.0 && r->atmpot[i] > 0.0) r->ifatm[i] = 1;
    if (r->ptimep[i] >= 0.0 && r->atmold[i] > 0.0) r->ifatmp[i] = 1;
    if (r->pnew[i] <= r->p.pmin && r->atmpot[i] < 0.0) {
      r->pnew[i] = r->p.pmin;
      r->ifatm[i] = 1;
    }
    if (r->ptimep[i] <= r->p.pmin && r->atmold[i] < 0.0) {
      r->ptimep[i] = r->p.pmin;
      r->ifatmp[i] = 1;
    }
  }
  for (int i = 0; i < nnod; ++i) {
    if (r->ifatm[i] == 0)
      r->atmact[i] = (r->atmpot[i] >= 0.0) ? r->atmpot[i]
                                           : (1.0 - r->p.scf) * r->atmpot[i];
    else
      r->atmact[i] = 0.0;
    if (r->ifatmp[i] == 1 || r->ifatmp[i] == 2) r->atmold[i] = 0.0;
  }
}

static void weight(int n, double teta, const double *vcurr,
                   const double *vprev, double *vwght) {
  /* SRC/weight.f. */
  double teta1 = 1.0 - teta;
  for (int k = 0; k < n; ++k) vwght[k] = teta * vcurr[k] + teta1 * vprev[k];
}

static void neumann_update(orc_run *r, const double *ckrw) {
  if (r->nq > 0)
    orc_neumann_free(r->nnod, r->contq, r->arenod, ckrw, r->kznod, r->q);
}

orc_run *orc_run_create(int nnod, int nstr, int ntri, const int *tetra,
                        const double *x, const double *y, const double *z,
                        const double *arenod, int nzone,
                        const double *permx, const double *permy,
                        const double *permz, const double *elstor,
                        const double *poros, const double *vgn,
                        const double *vgrmc, const double *vgpsat,
                        const double *ptimep0, int n_atm, int hspatm,
                        const double *atm_times, const double *atm_vals,
                        const orc_params *prm) {
  orc_run *r = (orc_run *)calloc(1, sizeof(orc_run));
  if (!r) return NULL;
  r->p = *prm;
  r->nnod = nnod;
  r->nstr = nstr;
  r->ntri = ntri;
  r->n = nnod * (nstr + 1);
  r->nt = 3 * ntri * nstr;
  r->nzone = nzone;
  int n = r->n, nt = r->nt;
  size_t ns = (size_t)nstr * nzone;
  r->tetra = (int *)xdup(tetra, sizeof(int) * 5 * (size_t)nt);
  r->x = (double *)xdup(x, sizeof(double) * n);
  r->y = (double *)xdup(y, sizeof(double) * n);
  r->z = (double *)xdup(z, sizeof(double) * n);
  r->arenod = (double *)xdup(arenod, sizeof(double) * nnod);
  r->permx = (double *)xdup(permx, sizeof(double) * ns);
  r->permy = (double *)xdup(permy, sizeof(double) * ns);
  r->permz = (double *)xdup(permz, sizeof(double) * ns);
  /* geometry: SRC/grdsys.f -> VOLBAS, LOCMAS */
  r->ai = DALLOC(4 * (size_t)nt); r->bi = DALLOC(4 * (size_t)nt);
  r->ci = DALLOC(4 * (size_t)nt); r->di = DALLOC(4 * (size_t)nt);
  r->volu = DALLOC(nt); r->volur = DALLOC(nt); r->ivol = IALLOC(nt);
  r->volnod = DALLOC(n);
  orc_volbas(n, nt, r->tetra, r->x, r->y, r->z, r->ai, r->bi, r->ci, r->di,
             r->volnod, r->volu, r->volur, r->ivol);
  orc_locmas(r->lmass, prm->lump);
  /* pattern: SRC/cathy_main.f:2599-2601 STRPIC + TETPIC for IOPT=1,
   * STRNEW + TETNEW + TOPIA for IOPT=2 */
  int cap = 32 * n;
  r->topol = IALLOC(n + 1);
  r->ja = IALLOC(cap);
  r->tetja = IALLOC(16 * (size_t)nt);
  if (prm->iopt == 2) {
    r->nterm = orc_strnew(n, nt, r->tetra, r->topol, r->ja, cap);
    orc_tetnew(nt, r->tetra, r->topol, r->ja, r->tetja);
    r->ia = IALLOC(r->nterm);
    orc_topia(n, r->topol, r->ia);
  } else {
    r->nterm = orc_strpic(n, nt, r->tetra, r->topol, r->ja, cap);
    orc_tetpic(nt, r->tetra, r->topol, r->ja, r->tetja);
  }
  /* soil: SRC/inital.f -> TPNODI, CHPARM */
  r->pnodi = DALLOC(n); r->snodi = DALLOC(n);
  r->vg = DALLOC((size_t)VG_NFIELD * n); r->pel = DALLOC(nt);
  r->tp = IALLOC(n);
  orc_tpnodi(n, nt, ntri, r->tetra, nzone, poros, elstor, vgn, vgrmc, vgpsat,
             r->pnodi, r->snodi, r->vg, r->pel, r->tp);
  orc_chparm(n, r->pnodi, r->vg);
  /* state */
  r->pnew = DALLOC(n); r->pold = DALLOC(n); r->ptnew = DALLOC(n);
  r->ptold = DALLOC(n); r->ptimep = DALLOC(n); r->pdiff = DALLOC(n);
  r->sw = DALLOC(n); r->ckrw = DALLOC(n); r->ckrwp = DALLOC(n);
  r->etai = DALLOC(n); r->et1 = DALLOC(n); r->et2 = DALLOC(n);
  r->swnew = DALLOC(n); r->swtimep = DALLOC(n);
  r->ckrwe = DALLOC(nt); r->etae = DALLOC(nt); r->et1e = DALLOC(nt);
  r->coef1 = DALLOC(r->nterm); r->coef2 = DALLOC(r->nterm);
  r->coef4 = DALLOC(r->nterm); r->coef5 = DALLOC(r->nterm);
  r->tnoti = DALLOC(n); r->xt5 = DALLOC(n);
  r->scr = DALLOC(prm->iopt == 2 ? (r->nterm > 10 * n ? (size_t)r->nterm : 10 * (size_t)n)
                                 : 5 * (size_t)n);
  if (prm->iopt == 2) {
    r->coef3 = DALLOC(r->nterm);
    r->dckrw = DALLOC(n);
    r->detai = DALLOC(n);
  }
  r->qtranie = DALLOC(n);
  r->ifatm = IALLOC(nnod); r->ifatmp = IALLOC(nnod);
  r->noddir = IALLOC(nnod + 1);
  r->atmpot = DALLOC(nnod); r->atmact = DALLOC(nnod); r->atmold = DALLOC(nnod);
  r->lhsatm = DALLOC(nnod); r->atminp = DALLOC(3 * (size_t)nnod);
  r->n_atm = n_atm;
  r->hspatm = hspatm;
  r->atm_times = (double *)xdup(atm_times, sizeof(double) * (size_t)n_atm);
  r->atm_vals = (double *)xdup(
      atm_vals, sizeof(double) * (size_t)n_atm * (hspatm ? 1 : nnod));
  /* initial condition and INIT1 (SRC/init1.f:37-60) */
  memcpy(r->ptimep, ptimep0, sizeof(double) * n);
  if (r->p.deltat > r->p.dtmax) r->p.deltat = r->p.dtmax;
  if (r->p.deltat <= r->p.dtmin) {
    r->p.deltat = r->p.dtmin;
    r->dtgmin = 0;
  } else {
    r->dtgmin = 1;
  }
  r->deltat = r->p.deltat;
  r->timep = 0.0;
  r->time = r->deltat;
  r->iter = 1;
  r->nstep = 1;
  memcpy(r->pnew, r->ptimep, sizeof(double) * n);
  /* free drainage: all bottom nodes (BCONE with NNEU<0) */
  r->nq = prm->free_drainage ? nnod : 0;
  r->contq = IALLOC(nnod);
  r->q = DALLOC(nnod);
  for (int i = 0; i < nnod; ++i) r->contq[i] = nstr * nnod + i;
  /* SRC/inital.f:216-223 ATMONE, POLD/PTNEW/PTOLD */
  atmone(r);
  memcpy(r->pold, r->pnew, sizeof(double) * n);
  weight(n, r->p.tetaf, r->pnew, r->ptimep, r->ptnew);
  memcpy(r->ptold, r->ptnew, sizeof(double) * n);
  /* MBINIT (SRC/mbinit.f:52-74) */
  r->adinp = r->adoutp = r->ndinp = r->ndoutp = 0.0;
  r->aninp = r->anoutp = 0.0;
  for (int k = 0; k < nnod; ++k)
    if (r->ifatmp[k] == 0) {
      if (r->atmold[k] > 0.0) r->aninp += r->atmold[k];
      else r->anoutp += r->atmold[k];
    }
  /* SRC/cathy_main.f:2663-2704: CHVELO, STORCAL, KZNOD, NEUMANN */
  orc_chvelo(n, r->ptimep, r->pnodi, r->vg, r->sw, r->ckrw);
  r->store1 = 0.0;
  for (int i = 0; i < n; ++i)
    r->store1 = r->store1 + r->sw[i] * r->volnod[i] * r->pnodi[i];
  r->store2 = r->store1;
  double *kz = DALLOC(nt);
  for (int e = 0; e < nt; ++e) {
    int istr = e / (3 * ntri), mt = r->tetra[(size_t)e * 5 + 4] - 1;
    kz[e] = r->permz[(size_t)istr * nzone + mt];
  }
  r->kznod = DALLOC(n);
  orc_eltnod(n, nt, r->tp, r->tetra, kz, r->kznod);
  free(kz);
  r->nninp = r->nnoutp = 0.0;
  if (r->nq > 0) {
    memcpy(r->ckrwp, r->ckrw, sizeof(double) * n);
    neumann_update(r, r->ckrw);
    for (int k = 0; k < r->nq; ++k) {
      if (r->q[k] > 0.0) r->nninp += r->q[k];
      else r->nnoutp += r->q[k];
    }
  }
  return r;
}

void orc_run_free(orc_run *r) {
  if (!r) return;
  void *ptrs[] = {r->tetra, r->topol, r->ja, r->tetja, r->ivol, r->tp,
                  r->x, r->y, r->z, r->ai, r->bi, r->ci, r->di, r->volu,
                  r->volur, r->volnod, r->arenod, r->permx, r->permy,
                  r->permz, r->pnodi, r->snodi, r->vg, r->pel, r->kznod,
                  r->pnew, r->pold, r->ptnew, r->ptold, r->ptimep, r->pdiff,
                  r->sw, r->ckrw, r->ckrwp, r->etai, r->et1, r->et2,
                  r->swnew, r->swtimep, r->ckrwe, r->etae, r->et1e,
                  r->coef1, r->coef2, r->coef4, r->coef5, r->tnoti, r->xt5,
                  r->scr, r->qtranie, r->contq, r->q, r->ifatm, r->ifatmp,
                  r->noddir, r->atmpot, r->atmact, r->atmold, r->lhsatm,
                  r->atm_times, r->atm_vals, r->atminp, r->ia, r->coef3,
                  r->dckrw, r->detai};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  free(r);
}

/* One Picard iteration: SRC/picard.f:64-203 (IPEAT=0, KSLOPE=0, IVGHU=0,
 * no nudging, NSF=0). */
static void picard(orc_run *r, double *err_out) {
  int n = r->n, nt = r->nt;
  orc_chpic0(n, r->ptnew, r->pnew, r->ptimep, r->swnew, r->swtimep,
             r->snodi, r->pnodi, r->vg, r->sw, r->ckrw, r->etai, r->et1,
             r->et2);
  orc_nodelt(nt, r->tetra, r->ckrw, r->ckrwe);
  orc_nodelt(nt, r->tetra, r->etai, r->etae);
  orc_nodelt(nt, r->tetra, r->et1, r->et1e);
  /* PEL here is NODELT(PNODI) (SRC/picuns.f:48) */
  orc_nodelt(nt, r->tetra, r->pnodi, r->pel);
  memset(r->coef1, 0, sizeof(double) * r->nterm);
  memset(r->coef2, 0, sizeof(double) * r->nterm);
  memset(r->coef4, 0, sizeof(double) * r->nterm);
  memset(r->coef5, 0, sizeof(double) * r->nterm);
  orc_asspic(nt, r->ntri, r->nzone, r->tetra, r->tetja, r->lmass, r->coef1,
             r->coef2, r->coef4, r->et1e, r->pel, r->permx, r->permy,
             r->permz, r->ckrwe, r->volu, r->volur, r->bi, r->ci, r->di);
  orc_rhspic(n, r->ja, r->topol, r->deltat, r->ptnew, r->pnew, r->ptimep,
             r->swnew, r->swtimep, r->coef1, r->coef2, r->coef4, r->tnoti);
  orc_cfmatp(n, r->nterm, r->p.tetaf, r->deltat, r->coef1, r->coef2,
             r->coef4, r->coef5, r->topol, r->et2);
  orc_rhsgrv(nt, r->ntri, r->nzone, r->tetra, r->tnoti, r->di, r->permz,
             r->ivol, r->ckrwe);
  memcpy(r->xt5, r->tnoti, sizeof(double) * n);
  r->numdir = orc_bcpic(0, r->nq, NULL, r->contq, NULL, r->q, r->nnod, n,
                        r->lhsatm, r->ifatm, r->atmact, r->atmold, r->topol,
                        r->coef1, r->tnoti, r->p.tetaf, 1.7e100, r->noddir,
                        r->qtranie);
  double err = 0.0, res = 0.0;
  int niter = orc_symslv(n, r->nterm, r->numdir, r->p.itmxcg, r->p.tolcg,
                         1.7e-100, r->noddir, r->ja, r->topol, r->pdiff,
                         r->tnoti, r->coef1, r->coef2, r->scr, &err, &res);
  r->itlin += niter;
  r->nitert += niter;
  if (niter >= r->p.itmxcg) {
    r->lsfail = 1;
    r->klsfai += 1;
  } else {
    r->lsfail = 0;
  }
  for (int k = 0; k < n; ++k) r->pnew[k] = r->pnew[k] + r->pdiff[k];
  orc_shlpic(0, r->topol, NULL, NULL, NULL, r->coef1, r->pnew, r->pold,
             r->nnod, r->ifatm, r->lhsatm);
  if (err_out) *err_out = err;
}

/* One Newton iteration: SRC/newton.f (IVGHU=0, KSLOPE=0, ISOLV=0, NP=0,
 * NSF=0). */
static void newton(orc_run *r, double *err_out) {
  int n = r->n, nt = r->nt;
  orc_chnew0(n, r->ptnew, r->snodi, r->pnodi, r->vg, r->sw, r->ckrw, r->etai,
             r->dckrw, r->detai);
  orc_nodelt(nt, r->tetra, r->ckrw, r->ckrwe);
  orc_nodelt(nt, r->tetra, r->etai, r->etae);
  memset(r->coef1, 0, sizeof(double) * r->nterm);
  memset(r->coef2, 0, sizeof(double) * r->nterm);
  memset(r->coef3, 0, sizeof(double) * r->nterm);
  orc_assnew(nt, r->ntri, r->nzone, r->tetra, r->tetja, r->lmass, r->coef1,
             r->coef2, r->coef3, r->permx, r->permy, r->permz, r->ckrwe,
             r->etae, r->ptnew, r->ptimep, r->pnew, r->dckrw, r->detai,
             r->p.tetaf, r->deltat, r->volu, r->volur, r->ivol, r->bi, r->ci,
             r->di);
  orc_rhsnew(n, r->ja, r->topol, r->deltat, r->ptnew, r->pnew, r->ptimep,
             r->coef1, r->coef2, r->tnoti);
  orc_cfmatn(r->nterm, r->p.tetaf, r->deltat, r->coef1, r->coef2, r->coef3);
  orc_rhsgrv(nt, r->ntri, r->nzone, r->tetra, r->tnoti, r->di, r->permz,
             r->ivol, r->ckrwe);
  memcpy(r->xt5, r->tnoti, sizeof(double) * n);
  r->numdir = orc_bcnew(r->nq, r->contq, r->q, r->nnod, r->lhsatm, r->ifatm,
                        r->atmact, r->atmold, r->ja, r->topol, r->coef1,
                        r->tnoti, r->p.tetaf, 1.7e100, r->noddir);
  double err = 0.0, res = 0.0, resini = 0.0;
  int niter = orc_nsyslv(n, r->nterm, r->numdir, r->noddir, r->p.itmxcg,
                         r->p.tolcg, 1.7e-100, r->ia, r->ja, r->topol,
                         r->coef1, r->coef2, r->coef3, r->scr, r->pdiff,
                         r->tnoti, &err, &res, &resini);
  r->itlin += niter;
  r->nitert += niter;
  if (niter >= r->p.itmxcg) {
    r->lsfail = 1;
    r->klsfai += 1;
  } else {
    r->lsfail = 0;
  }
  for (int k = 0; k < n; ++k) r->pnew[k] = r->pnew[k] + r->pdiff[k];
  orc_shlnew(r->topol, r->ja, r->coef1, r->pnew, r->pold, r->nnod, r->ifatm,
             r->lhsatm);
  if (err_out) *err_out = err;
}

/* MASBAL restated for IOPT=1, NP=0, NSF=0, no nudging
 * (SRC/masbal.f:37-107 with SRC/fluxmb.f). */
static void masbal(orc_run *r) {
  int n = r->n, nnod = r->nnod;
  if (r->p.iopt == 2) {
    orc_bknew(nnod, r->p.tetaf, r->topol, r->ja, r->ifatm, r->pdiff,
              r->atmact, r->atmold, r->coef1, r->xt5);
    /* The reference's Newton branch never sets FLOW3D's SWNEW / SWTIMEP
     * (SRC/flow3d.f:83 is a local of FLOW3D that only CHPIC0 writes), so
     * its STORMB reads stale values; here they are Sw at PNEW and at
     * PTIMEP, as the Picard branch has them. */
    for (int i = 0; i < n; ++i) {
      double pnot = VGF(r->vg, n, VG_PNOT, i), rmc = VGF(r->vg, n, VG_RMC, i);
      r->swnew[i] = pnot * orc_fvgse(r->pnew[i], r->vg, n, i) + rmc / r->pnodi[i];
      r->swtimep[i] = pnot * orc_fvgse(r->ptimep[i], r->vg, n, i) + rmc / r->pnodi[i];
    }
  } else {
    orc_bkpic(n, 0, nnod, r->p.tetaf, r->topol, r->ja, NULL, r->ifatm,
              r->pdiff, NULL, NULL, r->atmact, r->atmold, r->coef1, r->xt5,
              r->scr);
  }
  /* CHMASS(PTIMEP) into SCR and CHMASS(PNEW) into ETAI: only ETAI's new
   * value is used later (SRC/masbal.f:70-73). */
  orc_chmass(n, r->ptimep, r->snodi, r->pnodi, r->vg, r->scr);
  orc_chmass(n, r->pnew, r->snodi, r->pnodi, r->vg, r->etai);
  double deltam = 0.5 * r->deltat;
  r->dstore = orc_stormb(r->pnew, r->ptimep, r->swnew, r->swtimep, n,
                         r->volnod, r->snodi, r->pnodi);
  double ndin = 0.0, ndout = 0.0, nnin = 0.0, nnout = 0.0;
  for (int k = 0; k < r->nq; ++k) {
    if (r->q[k] > 0.0) nnin += r->q[k];
    else nnout += r->q[k];
  }
  double adin = 0.0, adout = 0.0, anin = 0.0, anout = 0.0;
  for (int k = 0; k < nnod; ++k) {
    if (r->ifatm[k] == -1) continue;
    if (r->ifatm[k] == 1 || r->ifatm[k] == 2) {
      if (r->atmact[k] > 0.0) adin += r->atmact[k];
      else adout += r->atmact[k];
    } else {
      if (r->atmact[k] > 0.0) anin += r->atmact[k];
      else anout += r->atmact[k];
    }
  }
  double vadin = (adin + r->adinp) * deltam, vadout = (adout + r->adoutp) * deltam;
  double vndin = (ndin + r->ndinp) * deltam, vndout = (ndout + r->ndoutp) * deltam;
  double vanin = (anin + r->aninp) * deltam, vanout = (anout + r->anoutp) * deltam;
  double vnnin = (nnin + r->nninp) * deltam, vnnout = (nnout + r->nnoutp) * deltam;
  r->vin = vadin + vndin + vanin + vnnin + 0.0;
  r->vout = vadout + vndout + vanout + vnnout + 0.0 + 0.0;
  r->erras = r->vin + r->vout - r->dstore;
  r->errel = ((r->vin + r->vout) != 0.0) ? 100.0 * r->erras / (r->vin + r->vout)
                                         : 0.0;
  /* current-step fluxes, promoted to "previous" once the step is accepted
   * (SRC/cathy_main.f:3283-3290) */
  r->adin = adin; r->adout = adout; r->ndin = ndin; r->ndout = ndout;
  r->anin = anin; r->anout = anout; r->nnin = nnin; r->nnout = nnout;
}

/* FLOW3D: SRC/flow3d.f:93-294 (NLRELX=0, NSF=0), Picard (IOPT=1) or
 * Newton (IOPT=2) iterations. */
static void flow3d(orc_run *r) {
  int n = r->n;
  for (;;) {
    if (r->p.iopt == 2) newton(r, NULL);
    else picard(r, NULL);
    masbal(r);
    /* CONVER: SRC/conver.f:24-48 */
    orc_norms(n, &r->pikmax, &r->pinf, &r->pl2, &r->finf, &r->fl2, r->pnew,
              r->pold, r->tnoti);
    if ((r->p.l2norm == 0 && r->pinf <= r->p.tolswi) ||
        (r->p.l2norm != 0 && r->pl2 <= r->p.tolswi))
      orc_switch_old(r->nnod, r->ifatm, r->atmact, r->atmpot, r->pnew,
                     r->p.pmin);
    r->bckstp = 0;
    int itagen = r->iter < r->p.ituns;
    int errgmx = (r->pl2 >= r->p.ernlmx || r->pinf >= r->p.ernlmx ||
                  r->fl2 >= r->p.ernlmx || r->finf >= r->p.ernlmx);
    int normcv = (r->p.l2norm == 0) ? (r->pinf <= r->p.toluns)
                                    : (r->pl2 <= r->p.toluns);
    if (!r->lsfail && !errgmx && !normcv && itagen) {
      memcpy(r->pold, r->pnew, sizeof(double) * n);
      memcpy(r->ptold, r->ptnew, sizeof(double) * n);
      weight(n, r->p.tetaf, r->pnew, r->ptimep, r->ptnew);
      r->iter += 1;
      continue;
    }
    r->itrtot += r->iter;
    if (!r->lsfail && !errgmx && normcv) {
      r->noback = 0;
    } else if (r->dtgmin) {
      r->bckstp = 1;
      r->noback = 0;
    } else {
      r->noback = 1;
    }
    return;
  }
}

/* BKSTEP: SRC/bkstep.f:58-147 (constant non-atmospheric BCs, NSF=0). */
static void bkstep(orc_run *r) {
  int n = r->n, nnod = r->nnod;
  memcpy(r->pnew, r->ptimep, sizeof(double) * n);
  memcpy(r->ifatm, r->ifatmp, sizeof(int) * nnod);
  memcpy(r->atmact, r->atmold, sizeof(double) * nnod);
  r->time = r->timep;
  r->deltat = r->deltat * r->p.dtredm - r->p.dtreds;
  if (r->deltat <= r->p.dtmin) {
    r->deltat = r->p.dtmin;
    r->dtgmin = 0;
  } else {
    r->dtgmin = 1;
  }
  r->time = r->time + r->deltat;
  r->kbackt += 1;
  r->iter = 1;
  r->nitert = 0;
  if (r->time > r->atmtim[1]) atmnxt(r);
  else atmbak(r);
  orc_switch_old(nnod, r->ifatm, r->atmact, r->atmpot, r->pnew, r->p.pmin);
  memcpy(r->pold, r->pnew, sizeof(double) * n);
  weight(n, r->p.tetaf, r->pnew, r->ptimep, r->ptnew);
  memcpy(r->ptold, r->ptnew, sizeof(double) * n);
}

int orc_run_step(orc_run *r, double *st) {
  if (r->finished) return r->finished;
  int n = r->n, nnod = r->nnod;
  /* top of the time loop, SRC/cathy_main.f:2892-2980 */
  neumann_update(r, r->ckrw);
  atmnxt(r);
  for (int i = 0; i < n; ++i) r->qtranie[i] = 0.0; /* ETRAN, no ET */
  orc_switch_old(nnod, r->ifatm, r->atmact, r->atmpot, r->pnew, r->p.pmin);
  memcpy(r->pold, r->pnew, sizeof(double) * n);
  weight(n, r->p.tetaf, r->pnew, r->ptimep, r->ptnew);
  memcpy(r->ptold, r->ptnew, sizeof(double) * n);
  int backs = 0;
  for (;;) { /* label 100 */
    flow3d(r);
    if (!r->bckstp) break;
    bkstep(r);
    neumann_update(r, r->ckrwp);
    ++backs;
  }
  /* SRC/cathy_main.f:3233-3262: CHVELO, STORCAL, cumulative balance */
  orc_chvelo(n, r->pnew, r->pnodi, r->vg, r->sw, r->ckrw);
  r->store1 = 0.0;
  for (int i = 0; i < n; ++i)
    r->store1 = r->store1 + r->sw[i] * r->volnod[i] * r->pnodi[i];
  r->store2 = r->store2 + r->dstore;
  r->vtot = r->vtot + (r->vin + r->vout);
  r->cvin = r->cvin + r->vin;
  r->cvout = r->cvout + r->vout;
  r->cdstor = r->cdstor + r->dstore;
  r->cerras = r->cerras + r->erras;
  r->caeras = r->caeras + fabs(r->erras);
  if (st) {
    st[0] = r->nstep; st[1] = r->deltat; st[2] = r->time; st[3] = r->iter;
    st[4] = r->nitert; st[5] = backs; st[6] = r->store1; st[7] = r->store2;
    st[8] = r->dstore; st[9] = r->cdstor; st[10] = r->vin; st[11] = r->cvin;
    st[12] = r->vout; st[13] = r->cvout; st[14] = r->vin + r->vout;
    st[15] = r->vtot; st[16] = r->erras; st[17] = r->errel;
    st[18] = r->cerras; st[19] = r->caeras; st[20] = r->pinf;
    st[21] = r->itrtot; st[22] = r->itlin; st[23] = r->klsfai;
  }
  r->adinp = r->adin; r->adoutp = r->adout; r->ndinp = r->ndin;
  r->ndoutp = r->ndout; r->aninp = r->anin; r->anoutp = r->anout;
  r->nninp = r->nnin; r->nnoutp = r->nnout;
  /* SRC/cathy_main.f:3765-3770 */
  memcpy(r->ifatmp, r->ifatm, sizeof(int) * nnod);
  memcpy(r->atmold, r->atmact, sizeof(double) * nnod);
  memcpy(r->ptimep, r->pnew, sizeof(double) * n);
  memcpy(r->ckrwp, r->ckrw, sizeof(double) * n);
  if (r->noback) return r->finished = 2;
  if (fabs(r->time - r->p.tmax) <= 0.001 * r->deltat) return r->finished = 1;
  r->dtgmin = orc_timupd(&r->timep, &r->time, &r->deltat, r->p.dtmin,
                         r->p.dtmax, r->p.tmax, r->p.dtmaga, r->p.dtmagm,
                         r->p.dtreds, r->p.dtredm, r->iter, r->p.ituns1,
                         r->p.ituns2);
  r->nstep += 1; /* TIMNXT */
  r->iter = 1;
  r->nitert = 0;
  r->kbackt = 0;
  return 0;
}

void orc_run_get_psi(const orc_run *r, double *psi) {
  memcpy(psi, r->pnew, sizeof(double) * r->n);
}
void orc_run_get_sw(const orc_run *r, double *sw) {
  memcpy(sw, r->sw, sizeof(double) * r->n);
}
void orc_run_get_ifatm(const orc_run *r, int *ifatm) {
  memcpy(ifatm, r->ifatm, sizeof(int) * r->nnod);
}
int orc_run_n(const orc_run *r) { return r->n; }
int orc_run_nterm(const orc_run *r) { return r->nterm; }
const int *orc_run_topol(const orc_run *r) { return r->topol; }
const int *orc_run_ja(const orc_run *r) { return r->ja; }

/* ===================================================================== */
/*                            Darcy velocities                           */
/* ===================================================================== */

void orc_vel3d(int nt, int ntri, int nzone, const int *tetra, const double *pnew,
               const double *permx, const double *permy, const double *permz,
               double *uu, double *vv, double *ww, const double *bi,
               const double *ci, const double *di, const double *ckrwe,
               const double *volur, const int *ivol) {
  /* SRC/vel3d.f:28-52: element velocities -kr K grad(psi + z) from the P1
   * gradient; perm* are [nstr][nzone] here, PERMX(ISTR,MTYPE) there. */
  for (int e = 0; e < nt; ++e) {
    int istr = e / (3 * ntri);
    int mt = tetra[(size_t)e * 5 + 4] - 1;
    size_t ix = (size_t)istr * nzone + mt;
    double bb = 0.0, cc = 0.0, dd = 0.0;
    double xyz = -ckrwe[e] * volur[e] * (double)ivol[e];
    for (int j = 0; j < 4; ++j) {
      int i = tetra[(size_t)e * 5 + j];
      bb = bb + pnew[i] * bi[(size_t)e * 4 + j];
      cc = cc + pnew[i] * ci[(size_t)e * 4 + j];
      dd = dd + pnew[i] * di[(size_t)e * 4 + j];
    }
    uu[e] = bb * xyz * permx[ix];
    vv[e] = cc * xyz * permy[ix];
    ww[e] = (dd * xyz - ckrwe[e]) * permz[ix];
  }
}

void orc_vnod3d(int n, int nt, const int *tp, const int *tetra, const double *uu,
                const double *vv, const double *ww, double *unod, double *vnod,
                double *wnod) {
  /* SRC/vnod3d.f:17-32: nodal means of the element velocities, summed in
   * element order. */
  for (int i = 0; i < n; ++i) unod[i] = vnod[i] = wnod[i] = 0.0;
  for (int k = 0; k < nt; ++k)
    for (int ii = 0; ii < 4; ++ii) {
      int nd = tetra[(size_t)k * 5 + ii];
      unod[nd] = unod[nd] + uu[k];
      vnod[nd] = vnod[nd] + vv[k];
      wnod[nd] = wnod[nd] + ww[k];
    }
  for (int i = 0; i < n; ++i) {
    unod[i] = unod[i] / tp[i];
    vnod[i] = vnod[i] / tp[i];
    wnod[i] = wnod[i] / tp[i];
  }
}

/* ---------------------------------------------------------------- routing */

/* MC (SRC/mc.f:8-63): Muskingum-Cunge coefficients of one branch and its
 * outflow at the new time level. */
void orc_mc(double local_slope, double epl, double ks, double w, double b1, double y1, double deltat,
            double q_in_kk, double q_in_kkp1, double q_out_kk, double q_overland, double *cu, double *ak,
            double *q_out_kkp1) {
  double qc = 1.0 / 3.0 * (q_in_kk + q_in_kkp1 + q_out_kk);
  const double qcmin = 1.0e-05;
  if (qc <= qcmin) qc = qcmin;
  const double beta = atan(local_slope);
  const double g = (1.0 - y1 + 2.0 / 3.0 * b1);
  const double ck = 5.0 / (3.0 * g) * pow(ks, 3.0 / 5.0) * pow(w, -2.0 / 5.0) * pow(sin(beta), 3.0 / 1.0e1) *
                    pow(qc, 1.0 - 3.0 * g / 5.0);
  *ak = ck / epl;
  *cu = ck * deltat / epl;
  double dh = pow(qc, 1.0 - b1) / (2 * g * w * tan(beta));
  if (dh < (1.0 - *cu)) dh = 1.0 - *cu;
  const double x = 0.50 - dh / (ck * epl);
  const double den = 2.0 * (1.0 - x) + *cu;
  const double c1 = (*cu - 2.0 * x) / den;
  const double c2 = (*cu + 2.0 * x) / den;
  const double c3 = (2.0 * (1.0 - x) - *cu) / den;
  const double c4 = (2.0 * ck * deltat) / den;
  *q_out_kkp1 = c1 * q_in_kkp1 + c2 * q_in_kk + c3 * q_out_kk + c4 * q_overland;
}

/* CALCOLI (SRC/calcoli.f:4-76): the level's rate of change dd at time t of
 * the substep.  Returns -1 where the reference stops (level at or below
 * the table's first height). */
static int calcoli(double hh, double *dd, double delta_t, int n_ha, const double *ah_r, const double *ha_r,
                   double q_in_kk, double t, double h_sfioro, double q_in_kkp1, double surface_water,
                   double h_fondo_rr) {
  double q_inflow_p = q_in_kk + (q_in_kkp1 - q_in_kk) / delta_t * t;
  q_inflow_p = q_inflow_p + surface_water;
  if (hh == h_fondo_rr && q_inflow_p < 0.0) {
    *dd = 0.0;
    return 0;
  }
  double ah_r_p = 0.0;
  int found = 0;
  for (int j = 0; j < n_ha - 1; ++j) {
    if (ha_r[j] < hh && ha_r[j + 1] >= hh) {
      ah_r_p = ah_r[j] + (ah_r[j + 1] - ah_r[j]) / (ha_r[j + 1] - ha_r[j]) * (hh - ha_r[j]);
      found = 1;
      break;
    }
  }
  if (!found) {
    if (hh > ha_r[n_ha - 1]) ah_r_p = ah_r[n_ha - 1];
    else return -1;
  }
  double qh_r_p;
  if (hh <= h_sfioro) qh_r_p = 0.0;
  else qh_r_p = q_inflow_p >= 0.0 ? q_inflow_p : 0.0;
  *dd = 1 / ah_r_p * (q_inflow_p - qh_r_p);
  return 0;
}

int orc_serbatoio(const orc_network *net, int r, double q_in_kk, double q_in_kkp1, double h_pool_kk,
                  double delta_t, double surface_water, double *q_out_kkp1, double *h_pool_kkp1) {
  const int n_ha = net->n_ha[r];
  double ha_r[30], ah_r[30];
  for (int i = 0; i < n_ha && i < 30; ++i) {
    ha_r[i] = net->hres[r + (size_t)i * net->maxres];
    ah_r[i] = net->ares[r + (size_t)i * net->maxres];
  }
  const double h_sfioro = net->h_soglia[r], h_fondo = net->h_fondo[r];
  const double t_iniz = 0.0, t_fin = delta_t;
  double dd, hh = h_pool_kk;
  if (calcoli(hh, &dd, delta_t, n_ha, ah_r, ha_r, q_in_kk, t_iniz, h_sfioro, q_in_kkp1, surface_water, h_fondo))
    return -1;
  const double dh1 = delta_t * dd;
  hh = h_pool_kk + 0.50 * dh1;
  if (hh < h_fondo) hh = h_fondo;
  if (calcoli(hh, &dd, delta_t, n_ha, ah_r, ha_r, q_in_kk, t_iniz + 0.50 * delta_t, h_sfioro, q_in_kkp1,
              surface_water, h_fondo))
    return -1;
  const double dh2 = delta_t * dd;
  hh = h_pool_kk + 0.50 * dh2;
  if (hh < h_fondo) hh = h_fondo;
  if (calcoli(hh, &dd, delta_t, n_ha, ah_r, ha_r, q_in_kk, t_iniz + 0.50 * delta_t, h_sfioro, q_in_kkp1,
              surface_water, h_fondo))
    return -1;
  const double dh3 = delta_t * dd;
  hh = h_pool_kk + 0.50 * dh3; /* the reference's fourth stage: half of dh3 */
  if (hh < h_fondo) hh = h_fondo;
  if (calcoli(hh, &dd, delta_t, n_ha, ah_r, ha_r, q_in_kk, t_fin, h_sfioro, q_in_kkp1, surface_water, h_fondo))
    return -1;
  const double dh4 = delta_t * dd;
  double h = h_pool_kk + dh1 / 6 + dh2 / 3 + dh3 / 3 + dh4 / 6;
  if (h < h_fondo) h = h_fondo;
  *h_pool_kkp1 = h;
  if (h <= h_sfioro) *q_out_kkp1 = 0.0;
  else *q_out_kkp1 = q_in_kkp1 + surface_water >= 0.0 ? q_in_kkp1 + surface_water : 0.0;
  return 0;
}

/* ROUTE (SRC/route.f:4-234): cells in QOI_SN order, each routing the
 * cardinal (W_1) and diagonal (W_2) shares of its inflow and of the surface
 * water over its cell -- by Muskingum-Cunge (TIPO_R = 0) or through its
 * reservoir (0 < TIPO_R < 1000) -- then adding each outflow to the
 * downstream cell's Q_IN_KKP1 (the last cell's cardinal outflow leaves). */
int orc_route(const orc_network *net, double deltat, const double *surface_water, const double *q_in_kk,
              const double *q_out_kk1, const double *q_out_kk2, double *q_in_kkp1, double *q_out_kkp1_1,
              double *q_out_kkp1_2, double *cu_ak, const double *h_pool_kk, double *h_pool_kkp1) {
  const int nrow = net->nrow;
  double cu_max = 0.0;
  for (int q = 0; q < net->ncell; ++q) {
    const int ib = net->qoi[q];           /* 1-based I_BASIN */
    const int jr = ib % nrow;
    const int j = jr != 0 ? jr : nrow;
    const int i = jr != 0 ? (ib - j) / nrow + 1 : ib / nrow;
    const size_t g = (size_t)(i - 1) + (size_t)(j - 1) * net->ldc;
    const double nrc = net->nrc[g];
    const double sw = surface_water[ib - 1] / nrc;
    const int tipo = net->tipo_r ? net->tipo_r[ib - 1] : 0;
    for (int br = 0; br < 2; ++br) {
      const double w = br ? net->w2[g] : net->w1[g];
      if (w == 0) continue;
      const double epl = br ? net->epl2[g] : net->epl1[g];
      const double cr = 1.0 / epl;
      const double q_overland = sw * w * cr;
      const int pout = br ? net->pout2[g] : net->pout1[g];
      /* NINT((P_OUTFLOW-5)/3.): a default-real quotient, rounded half away from zero */
      const int iii = (int)lroundf((float)(pout - 5) / 3.0f);
      const int jjj = pout - 5 - 3 * iii;
      const int ib_cv = (i + iii - 1) * nrow + (j + jjj);
      const double qin_kk = q_in_kk[ib - 1] * w / nrc;
      const double qout_kk = (br ? q_out_kk2[ib - 1] : q_out_kk1[ib - 1]) / nrc;
      const double qin_kkp1 = q_in_kkp1[ib - 1] * w / nrc;
      double *out = br ? q_out_kkp1_2 : q_out_kkp1_1;
      if (tipo >= 1000) return -1; /* 'INCOMPRENSIBILE CASE!' and STOP */
      if (tipo != 0) {
        const int r = net->reservr[ib - 1] - 1;
        double qout, h;
        if (orc_serbatoio(net, r, qin_kk, qin_kkp1, h_pool_kk[r], deltat, sw, &qout, &h)) return -1;
        if (qout < 0.0) qout = 0.0;
        h_pool_kkp1[r] = h;
        out[ib - 1] = qout;
      } else {
        double cu, ak, qout;
        orc_mc(br ? net->slope2[g] : net->slope1[g], epl, br ? net->kss2[g] : net->kss1[g],
               br ? net->ws2[g] : net->ws1[g], net->b1[g], net->y1[g], deltat, qin_kk, qin_kkp1, qout_kk,
               q_overland, &cu, &ak, &qout);
        if (qout < 0.0) qout = 0.0;
        out[ib - 1] = qout * nrc;
        if (cu >= cu_max) {
          cu_max = cu;
          cu_ak[1] = ak;
        }
      }
      if (br == 1 || q != net->ncell - 1) q_in_kkp1[ib_cv - 1] = q_in_kkp1[ib_cv - 1] + out[ib - 1];
    }
  }
  cu_ak[0] = cu_max;
  return 0;
}

/* ALTEZZE (SRC/altezze.f:4-107): cell volume and water depth at the new
 * time level; reservoir cells (0 < TIPO_R < 1000) and lake cells
 * (TIPO_R >= 1000) take the depth of their reservoir's pool. */
void orc_altezze(const orc_network *net, const int *celtype, int ncell, double dx, double dy, double deltat,
                 const double *q_in_kk, const double *q_in_kkp1, const double *q_out_kk1, const double *q_out_kk2,
                 const double *q_out_kkp1_1, const double *q_out_kkp1_2, const double *volume_kk,
                 const double *surface_water, double *volume_kkp1, double *h_water_kkp1,
                 const double *h_pool_kkp1) {
  for (int c = 0; c < ncell; ++c) {
    const int tipo = celtype ? celtype[c] : 0;
    if (tipo == 0) {
      const double dv = (q_in_kk[c] + q_in_kkp1[c]) / 2 * deltat + surface_water[c] * deltat -
                        (q_out_kk1[c] + q_out_kk2[c]) / 2 * deltat -
                        (q_out_kkp1_1[c] + q_out_kkp1_2[c]) / 2 * deltat;
      const double v = volume_kk[c] + dv;
      volume_kkp1[c] = v;
      if (v >= 0.0) {
        h_water_kkp1[c] = v / (dx * dy);
      } else {
        volume_kkp1[c] = 0.0;
        h_water_kkp1[c] = 0.0;
      }
    } else if (tipo < 1000) {
      h_water_kkp1[c] = h_pool_kkp1[net->reservr[c] - 1] - net->elevation[c];
      if (h_water_kkp1[c] < 0) h_water_kkp1[c] = 0.0;
    } else {
      const int aa = net->reservr[tipo - 1000 - 1];
      const double h = h_pool_kkp1[aa - 1] - net->elevation[c];
      h_water_kkp1[c] = h > 0.0 ? h : 0.0;
    }
  }
}
