/*
 * tetwt.c -- FP64 CPU restatement of the tetrahedron histogram weights of
 * the chi0 path (gettetwt + tetwt5x_dtet4).  TEST INFRASTRUCTURE ONLY: the
 * checker of the HIP implementation, loaded only by tests/, smoke() and the
 * cpu_baseline leg of bench.py.
 *
 * Scope: npm = 1 (time reversal) or 2, nctot core states, nmtet = 1,
 * tetrahedron_matrix_linear = F or T (S/switch.f90:253-263), no tetrakbt,
 * no --interbandonly/--intrabandonly filters, no Fermi-level shift.
 *
 * References (SRC/subroutines = S/):
 *   gettetwt          S/m_tetwt.f90:36-256
 *   tetwt5x_dtet4     S/tetwt5.f90:9-486
 *   hisrange          S/tetwt5.f90:487-513
 *   lindtet6          S/tetwt5.f90:514-600
 *   inttetra6         S/tetwt5.f90:601-672
 *   inttetrac6        S/tetwt5.f90:673-713
 *   intttvc6          S/tetwt5.f90:714-787
 *   mkwcg, wab        S/tetwt5.f90:788-840
 *   minv33tp          S/util.f90:202-212
 *   midk3, midk       S/tetwt5.f90:848-857, 913-919
 *   chkdgn            S/tetwt5.f90:868-912
 *   rsvwwk00_4        S/tetwt5.f90:920-942
 *   integtetn         S/m_tetrakbt.f90:441-466
 *   sortea            S/util.f90:574-594
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PI 3.14159265358979323846

/* sortea: stable insertion sort, ascending; ord is 0-based */
static void sortea(const double* ea, int* ord, int n) {
  for (int i = 0; i < n; ++i) ord[i] = i;
  for (int ix = 1; ix < n; ++ix)
    for (int i = ix; i >= 1; --i) {
      if (ea[ord[i - 1]] > ea[ord[i]]) {
        int t = ord[i - 1];
        ord[i - 1] = ord[i];
        ord[i] = t;
        continue;
      }
      break;
    }
}

static double det33(const double am[3][3]) { /* am[col][row] = am(row+1, col+1) */
  return am[0][0] * am[1][1] * am[2][2] - am[0][0] * am[1][2] * am[2][1] - am[0][1] * am[1][0] * am[2][2] +
         am[0][1] * am[1][2] * am[2][0] + am[0][2] * am[1][0] * am[2][1] - am[0][2] * am[1][1] * am[2][0];
}

/* primitive of the cross-section area of a linear function over a tetrahedron, e ascending */
static int integtetn(const double e[4], double ee, double* out) {
  const double e1 = e[0] - 3e-8, e2 = e[1] - 2e-8, e3 = e[2] - 1e-8, e4 = e[3];
  const double V1 = ee - e1, V2 = ee - e2, V4 = ee - e4;
  if (ee < e1) {
    *out = 0.0;
  } else if (e1 <= ee && ee < e2) {
    *out = V1 * V1 * V1 / ((e4 - e1) * (e3 - e1) * (e2 - e1));
  } else if (e2 <= ee && ee < e3) {
    const double a = V1 / pow((e4 - e1) * (e3 - e1) * (e2 - e1), 1.0 / 3.0);
    const double b = V2 / pow(-(e4 - e2) * (e3 - e2) * (e1 - e2), 1.0 / 3.0);
    *out = (a - b) * (a * a + a * b + b * b);
  } else if (e3 <= ee && ee < e4) {
    *out = 1.0 - V4 * V4 * V4 / ((e1 - e4) * (e2 - e4) * (e3 - e4));
  } else if (ee == e4) {
    *out = 1.0;
  } else {
    return -1;
  }
  return 0;
}

/* What the per-tetrahedron routines share: the main tetrahedron's corners kkv (in idtetf order) and
 * whether the weights are split over its corners (tetrahedron_matrix_linear).  The histogram wtthis
 * is wtthis(nwhis, 4) column-major; without matrix_linear only column 1 is used. */
typedef struct {
  int ml;
  double kkv[4][3];
} tetctx;

/* mkwcg: weight of each corner (ascending energies e) in the point that represents the
 * cross-section of the tetrahedron at energy ee; the weights sum to 1 (S/tetwt5.f90:788-833) */
static void mkwcg(const double e[4], double ee, double wcg[4]) {
  const double e1 = e[0] - 3e-8, e2 = e[1] - 2e-8, e3 = e[2] - 1e-8, e4 = e[3];
#define WAB(ea, eb, wa, wb) \
  do {                      \
    const double t_ = (eb) - (ea); \
    wa = ((eb) - ee) / t_;  \
    wb = (ee - (ea)) / t_;  \
  } while (0)
  double w12_1, w12_2, w13_1, w13_3, w14_1, w14_4, w23_2, w23_3, w24_2, w24_4, w34_3, w34_4;
  if (ee <= e[0]) {
    wcg[0] = 1.0;
    wcg[1] = wcg[2] = wcg[3] = 0.0;
  } else if (e1 <= ee && ee < e2) {
    WAB(e1, e2, w12_1, w12_2);
    WAB(e1, e3, w13_1, w13_3);
    WAB(e1, e4, w14_1, w14_4);
    wcg[0] = (w12_1 + w13_1 + w14_1) / 3.0;
    wcg[1] = w12_2 / 3.0;
    wcg[2] = w13_3 / 3.0;
    wcg[3] = w14_4 / 3.0;
  } else if (e2 <= ee && ee < e3) {
    WAB(e1, e3, w13_1, w13_3);
    WAB(e1, e4, w14_1, w14_4);
    WAB(e2, e3, w23_2, w23_3);
    WAB(e2, e4, w24_2, w24_4);
    wcg[0] = (w13_1 + w14_1) / 4.0;
    wcg[1] = (w23_2 + w24_2) / 4.0;
    wcg[2] = (w13_3 + w23_3) / 4.0;
    wcg[3] = (w14_4 + w24_4) / 4.0;
  } else if (e3 <= ee && ee < e4) {
    WAB(e1, e4, w14_1, w14_4);
    WAB(e2, e4, w24_2, w24_4);
    WAB(e3, e4, w34_3, w34_4);
    wcg[0] = w14_1 / 3.0;
    wcg[1] = w24_2 / 3.0;
    wcg[2] = w34_3 / 3.0;
    wcg[3] = (w24_4 + w34_4 + w14_4) / 3.0;
  } else {
    /* ee > e(4); the reference leaves wcg as it was for ee == e(4) exactly, taken here as the top corner */
    wcg[3] = 1.0;
    wcg[0] = wcg[1] = wcg[2] = 0.0;
  }
#undef WAB
}

/* intttvc6: wtthis(ihis) += int_bin d omega int d^3k delta(omega + v(k)); with matrix_linear the
 * bin's weight (times 4) is split over the main corners by kkvkin(ix, i), the coordinates of this
 * tetrahedron's corner i in the main corners, applied to mkwcg at the bin centre */
static int intttvc6(const tetctx* cx, const double kkvkin[4][4], const double v[4], double voltet,
                    const double* frhis, int nwhis, double* wtthis) {
  double mv[4], WW[4];
  int ord[4];
  for (int i = 0; i < 4; ++i) mv[i] = -v[i];
  sortea(mv, ord, 4);
  for (int i = 0; i < 4; ++i) WW[i] = -v[ord[i]];
  if (!(WW[0] <= WW[1]) || !(WW[1] <= WW[2]) || !(WW[2] <= WW[3])) return -2;
  int inihis = -999, iedhis = -999, ix = 1;
  for (int ihis = 1; ihis <= nwhis; ++ihis) {
    if (ix == 1 && WW[0] < frhis[ihis]) { /* frhis(ihis+1) */
      inihis = ihis;
      ix = 4;
    }
    if (ix == 4 && WW[3] < frhis[ihis]) {
      iedhis = ihis;
      break;
    }
  }
  if (iedhis == -999 || inihis == -999) return -3;
  double norm;
  if (integtetn(WW, WW[3], &norm)) return -4;
  const double pvn = voltet / norm;
  double intega = 0.0, integb;
  for (int ihis = inihis; ihis <= iedhis; ++ihis) {
    if (frhis[ihis] > WW[3]) {
      integb = norm;
    } else if (integtetn(WW, frhis[ihis], &integb)) {
      return -4;
    }
    if (cx->ml) {
      double wcg[4];
      mkwcg(WW, 0.5 * (frhis[ihis - 1] + frhis[ihis]), wcg);
      const double wttt = pvn * (integb - intega) * 4.0;
      for (int c = 0; c < 4; ++c) {
        double www = 0.0;
        for (int j = 0; j < 4; ++j) www += kkvkin[c][ord[j]] * wcg[j];
        wtthis[ihis - 1 + (size_t)nwhis * c] += wttt * www;
      }
    } else {
      wtthis[ihis - 1] += pvn * (integb - intega);
    }
    intega = integb;
  }
  return 0;
}

/* minv33tp: the reciprocal vectors of the columns of a (columns of b; b(:,j).a(:,i) = delta_ij) */
static void minv33tp(const double a[3][3], double b[3][3]) {
  for (int j = 0; j < 3; ++j) {
    const double* u = a[(j + 1) % 3];
    const double* w = a[(j + 2) % 3];
    b[j][0] = u[1] * w[2] - u[2] * w[1];
    b[j][1] = u[2] * w[0] - u[0] * w[2];
    b[j][2] = u[0] * w[1] - u[1] * w[0];
  }
  const double det = a[0][0] * b[0][0] + a[0][1] * b[0][1] + a[0][2] * b[0][2];
  for (int j = 0; j < 3; ++j)
    for (int d = 0; d < 3; ++d) b[j][d] /= det;
}

/* corners: kk[8][3], xx[8]; itetx 1-based as in the reference */
static int inttetrac6(const tetctx* cx, double kk[8][3], const double xx[8], const int itetx[4],
                      const double* frhis, int nwhis, double* wtthis) {
  double xin[4], kin[4][3], am[3][3], kkvkin[4][4] = {{0}};
  for (int i = 0; i < 4; ++i) {
    xin[i] = xx[itetx[i] - 1];
    for (int d = 0; d < 3; ++d) kin[i][d] = kk[itetx[i] - 1][d];
  }
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) am[i][d] = kin[i][d] - kin[3][d];
  const double voltet = fabs(det33(am) / 6.0);
  if (cx->ml) { /* kin(:, i) = sum_c kkv(:, c) kkvkin(c, i) */
    double bm[3][3], kkv4bm[3];
    for (int i = 0; i < 3; ++i)
      for (int d = 0; d < 3; ++d) am[i][d] = cx->kkv[i][d] - cx->kkv[3][d];
    minv33tp(am, bm);
    for (int j = 0; j < 3; ++j) kkv4bm[j] = cx->kkv[3][0] * bm[j][0] + cx->kkv[3][1] * bm[j][1] + cx->kkv[3][2] * bm[j][2];
    for (int i = 0; i < 4; ++i) {
      for (int j = 0; j < 3; ++j)
        kkvkin[j][i] = kin[i][0] * bm[j][0] + kin[i][1] * bm[j][1] + kin[i][2] * bm[j][2] - kkv4bm[j];
      kkvkin[3][i] = 1.0 - (kkvkin[0][i] + kkvkin[1][i] + kkvkin[2][i]);
    }
  }
  return intttvc6(cx, (const double(*)[4])kkvkin, xin, voltet, frhis, nwhis, wtthis);
}

/* point on the edge k(i)--k(j) where ee crosses zero (i, j 1-based) */
static void midk(double kk[8][3], const double* ee, double* xx, int i, int j, int out) {
  const double ratio = ee[i - 1] / (ee[i - 1] - ee[j - 1]);
  xx[out - 1] = xx[i - 1] + ratio * (xx[j - 1] - xx[i - 1]);
  for (int d = 0; d < 3; ++d) kk[out - 1][d] = kk[i - 1][d] + ratio * (kk[j - 1][d] - kk[i - 1][d]);
}

static void midk3(double kk[8][3], const double* ee, double* xx, double* yy, int i, int j, int out) {
  const double ratio = ee[i - 1] / (ee[i - 1] - ee[j - 1]);
  xx[out - 1] = xx[i - 1] + ratio * (xx[j - 1] - xx[i - 1]);
  yy[out - 1] = yy[i - 1] + ratio * (yy[j - 1] - yy[i - 1]);
  for (int d = 0; d < 3; ++d) kk[out - 1][d] = kk[i - 1][d] + ratio * (kk[j - 1][d] - kk[i - 1][d]);
}

/* the unoccupied cut (eb > E_F) of one occupied sub-tetrahedron */
static int inttetra6(const tetctx* cx, double kk_[8][3], const double xx_[8], const double ebf[8],
                     const int itetx[4], const double* frhis, int nwhis, double* wtthis) {
  double xin[4], kin[4][3], ebfin[4], kk[8][3], xx[8], ee[4];
  int ord[4];
  for (int i = 0; i < 4; ++i) {
    xin[i] = xx_[itetx[i] - 1];
    ebfin[i] = ebf[itetx[i] - 1];
    for (int d = 0; d < 3; ++d) kin[i][d] = kk_[itetx[i] - 1][d];
  }
  sortea(ebfin, ord, 4); /* ee(1) <= ... <= ee(4) */
  for (int i = 0; i < 4; ++i) {
    ee[i] = ebfin[ord[i]];
    xx[i] = xin[ord[i]];
    for (int d = 0; d < 3; ++d) kk[i][d] = kin[ord[i]][d];
  }
  int rc = 0;
  if (0.0 >= ee[3]) {
    /* nothing unoccupied */
  } else if (ee[3] > 0.0 && 0.0 >= ee[2]) { /* Fig 1 */
    midk(kk, ee, xx, 4, 2, 5);
    midk(kk, ee, xx, 4, 1, 6);
    midk(kk, ee, xx, 4, 3, 7);
    rc |= inttetrac6(cx, kk, xx, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
  } else if (ee[2] > 0.0 && 0.0 >= ee[1]) { /* Fig 2 */
    midk(kk, ee, xx, 4, 2, 5);
    midk(kk, ee, xx, 4, 1, 6);
    midk(kk, ee, xx, 3, 1, 7);
    midk(kk, ee, xx, 3, 2, 8);
    rc |= inttetrac6(cx, kk, xx, (const int[]){4, 3, 5, 6}, frhis, nwhis, wtthis);
    rc |= inttetrac6(cx, kk, xx, (const int[]){3, 6, 7, 5}, frhis, nwhis, wtthis);
    rc |= inttetrac6(cx, kk, xx, (const int[]){3, 5, 7, 8}, frhis, nwhis, wtthis);
  } else if (ee[1] > 0.0 && 0.0 >= ee[0]) { /* Fig 3 */
    midk(kk, ee, xx, 1, 4, 5);
    midk(kk, ee, xx, 1, 2, 6);
    midk(kk, ee, xx, 1, 3, 7);
    rc |= inttetrac6(cx, kk, xx, (const int[]){3, 4, 7, 2}, frhis, nwhis, wtthis);
    rc |= inttetrac6(cx, kk, xx, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
    rc |= inttetrac6(cx, kk, xx, (const int[]){4, 2, 6, 7}, frhis, nwhis, wtthis);
  } else { /* Fig 4 */
    rc |= inttetrac6(cx, kk, xx, (const int[]){1, 2, 3, 4}, frhis, nwhis, wtthis);
  }
  return rc;
}

/* occupied cut (ea < E_F) of one tetrahedron, then the unoccupied cut of each piece:
 * wtthis(ihis) = int_bin d omega int d^3k f(ea)(1-f(eb)) delta(omega + x(k)) */
static int lindtet6(const tetctx* cx, const double kvec[4][3], const double ea[4], const double eb[4],
                    const double x[4], double efermia, double efermib, const double* frhis, int nwhis,
                    double* wtthis) {
  int ord[4];
  sortea(ea, ord, 4);
  int r[4] = {ord[3], ord[2], ord[1], ord[0]}; /* descending ea */
  double kk[8][3], xx[8], ee[4], ebf[8];
  for (int i = 0; i < 4; ++i) {
    for (int d = 0; d < 3; ++d) kk[i][d] = kvec[r[i]][d];
    ee[i] = ea[r[i]] - efermia;
    xx[i] = x[r[i]];
    ebf[i] = eb[r[i]] - efermib;
  }
  int rc = 0;
  if (0.0 <= ee[3]) {
    /* nothing occupied */
  } else if (ee[3] < 0.0 && 0.0 <= ee[2]) { /* Fig 1 */
    midk3(kk, ee, xx, ebf, 4, 2, 5);
    midk3(kk, ee, xx, ebf, 4, 1, 6);
    midk3(kk, ee, xx, ebf, 4, 3, 7);
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
  } else if (ee[2] < 0.0 && 0.0 <= ee[1]) { /* Fig 2 */
    midk3(kk, ee, xx, ebf, 4, 2, 5);
    midk3(kk, ee, xx, ebf, 4, 1, 6);
    midk3(kk, ee, xx, ebf, 3, 1, 7);
    midk3(kk, ee, xx, ebf, 3, 2, 8);
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){4, 3, 5, 6}, frhis, nwhis, wtthis);
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){3, 6, 7, 5}, frhis, nwhis, wtthis);
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){3, 5, 7, 8}, frhis, nwhis, wtthis);
  } else if (ee[1] < 0.0 && 0.0 <= ee[0]) { /* Fig 3 */
    midk3(kk, ee, xx, ebf, 1, 4, 5);
    midk3(kk, ee, xx, ebf, 1, 2, 6);
    midk3(kk, ee, xx, ebf, 1, 3, 7);
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){3, 4, 7, 2}, frhis, nwhis, wtthis);
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){4, 2, 6, 7}, frhis, nwhis, wtthis);
  } else { /* Fig 4 */
    rc |= inttetra6(cx, kk, xx, ebf, (const int[]){1, 2, 3, 4}, frhis, nwhis, wtthis);
  }
  return rc;
}

int orc_lindtet6(const double kvec[4][3], const double ea[4], const double eb[4], const double x[4],
                 double efermia, double efermib, const double* frhis, int nwhis, double* wtthis) {
  tetctx cx;
  cx.ml = 0;
  memcpy(cx.kkv, kvec, sizeof(cx.kkv));
  return lindtet6(&cx, kvec, ea, eb, x, efermia, efermib, frhis, nwhis, wtthis);
}

/* matrix_linear: wtthis(nwhis, 4), column c for corner c of kvec (nmtet = 1: kkv = kvec) */
int orc_lindtet6_ml(const double kvec[4][3], const double ea[4], const double eb[4], const double x[4],
                    double efermia, double efermib, const double* frhis, int nwhis, double* wtthis) {
  tetctx cx;
  cx.ml = 1;
  memcpy(cx.kkv, kvec, sizeof(cx.kkv));
  return lindtet6(&cx, kvec, ea, eb, x, efermia, efermib, frhis, nwhis, wtthis);
}

int orc_hisrange(const double* frhis, int nwhis, double demin, double demax, int* ihw, int* nhw) {
  *ihw = -9999;
  *nhw = -9999;
  for (int ihis = 1; ihis <= nwhis; ++ihis)
    if (demin < frhis[ihis]) {
      *ihw = ihis;
      break;
    }
  if (*ihw == -9999) return -1;
  for (int ihis = *ihw; ihis <= nwhis; ++ihis)
    if (demax < frhis[ihis]) {
      *nhw = ihis - *ihw + 1;
      break;
    }
  return *nhw == -9999 ? -1 : 0;
}

/* degenerate groups of consecutive energies (|de| <= 1e-4); 1-based ranges */
int orc_chkdgn(const double* ene, int ndat, int* ixini, int* ixend) {
  if (ndat < 1) return 0;
  ixini[0] = 1;
  if (ndat == 1) {
    ixend[0] = 1;
    return 1;
  }
  int i = 1;
  for (int ix = 2; ix <= ndat; ++ix) {
    if (fabs(ene[ix - 1] - ene[ix - 2]) > 1e-4) {
      ixend[i - 1] = ix - 1;
      i = i + 1;
      ixini[i - 1] = ixend[i - 2] + 1;
      if (ix == ndat) ixend[i - 1] = ix;
    } else if (ix == ndat) {
      ixend[i - 1] = ndat;
    }
  }
  return i;
}

/* Band index of pair coordinate (ibx, jbx) for jpm (0: occupied at k, unoccupied at k+q; 1: the reverse),
 * S/tetwt5.f90:252-264: beyond the valence bands below E_F come the core states (index nband + c). */
static void pair_bands(int jpm, int ibx, int jbx, int nocc_k, int nocc_kq, int nband, int* ib, int* jb) {
  if (jpm == 0) {
    *ib = ibx < nocc_k ? ibx : ibx - nocc_k + nband;
    *jb = jbx;
  } else {
    *ib = ibx;
    *jb = jbx < nocc_kq ? jbx : jbx - nocc_kq + nband;
  }
}

/* IDX4: ib over nband + nctot, jb over nband + ncc, then k, then jpm */
#define IDX4(ib, jb, k, jpm) \
  ((size_t)(ib) + (size_t)n1 * ((size_t)(jb) + (size_t)n2 * ((size_t)(k) + (size_t)nqbz * (size_t)(jpm))))

int orc_tetwt_full(int nband, int nqbz, int ntetf, const int* idtetf, const double* qbzw, const int* ib1bz,
                   const double* ek1, const double* ek2, double ef, double ebmx, const double* qlat,
                   const double* frhis, int nwhis, int npm, int nctot, const double* ecore, int matrix_linear,
                   int nbnbx_cap, int nhw_cap, int* nbnb, int* n1b, int* n2b, int* ihw, int* nhw, int* jhw, double* whw,
                   int* nbnbx_out, int* nhwtot_out) {
  if (npm < 1 || npm > 2 || nctot < 0 || (nctot > 0 && !ecore)) return -5;
  double qb[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) qb[c][r] = qlat[r + 3 * c];
  const double voltot = fabs(det33(qb));
  const double piofvoltot = PI / 4.0 / voltot;
  const int ncc = npm == 2 ? nctot : 0, n1 = nband + nctot, n2 = nband + ncc;
  /* energies with the core states appended at every k (S/tetwt5.f90:179-185) */
  double* e1 = (double*)malloc(sizeof(double) * (size_t)n1 * nqbz);
  double* e2 = (double*)malloc(sizeof(double) * (size_t)n1 * nqbz);
  for (int k = 0; k < nqbz; ++k) {
    for (int b = 0; b < nband; ++b) {
      e1[b + (size_t)n1 * k] = ek1[b + (size_t)nband * k];
      e2[b + (size_t)n1 * k] = ek2[b + (size_t)nband * k];
    }
    for (int c = 0; c < nctot; ++c) e1[nband + c + (size_t)n1 * k] = e2[nband + c + (size_t)n1 * k] = ecore[c];
  }
  const size_t nbb = (size_t)n1 * n2 * nqbz * npm;
  unsigned char* iwgt = (unsigned char*)calloc(nbb, 1);
  double* demin = (double*)malloc(sizeof(double) * nbb);
  double* demax = (double*)malloc(sizeof(double) * nbb);
  int* ibjb = (int*)calloc(nbb, sizeof(int));
  int rc = 0;
  for (size_t e = 0; e < nbb; ++e) {
    demin[e] = 1e10;
    demax[e] = -1e10;
  }
  for (int job = 0; job <= 1 && rc == 0; ++job) {
    for (int itet = 0; itet < ntetf && rc == 0; ++itet) {
      int kk[4];
      double kvec[4][3];
      for (int c = 0; c < 4; ++c) {
        const int w = idtetf[c + 4 * itet] - 1;
        kk[c] = ib1bz[w] - 1;
        for (int d = 0; d < 3; ++d) kvec[c][d] = qbzw[d + 3 * w];
      }
      /* valence bands below E_F (max over corners) and below the cutoff (min over corners) */
      int nocc_k = 0, nocc_kq = 0, nbmx_k = nband, nbmx_kq = nband;
      for (int c = 0; c < 4; ++c) {
        int o1 = 0, o2 = 0, m1 = 0, m2 = 0;
        for (int b = 0; b < nband; ++b) {
          o1 += ek1[b + (size_t)nband * kk[c]] < ef;
          o2 += ek2[b + (size_t)nband * kk[c]] < ef;
          m1 += ek1[b + (size_t)nband * kk[c]] < fmin(1e10, ebmx);
          m2 += ek2[b + (size_t)nband * kk[c]] < fmin(1e10, ebmx);
        }
        if (o1 > nocc_k) nocc_k = o1;
        if (o2 > nocc_kq) nocc_kq = o2;
        if (m1 < nbmx_k) nbmx_k = m1;
        if (m2 < nbmx_kq) nbmx_kq = m2;
      }
      for (int jpm = 0; jpm < npm && rc == 0; ++jpm) {
        const int ni = jpm == 0 ? nocc_k + nctot : nbmx_k, nj = jpm == 0 ? nbmx_kq : nocc_kq + nctot;
        for (int ibx = 0; ibx < ni && rc == 0; ++ibx)
          for (int jbx = 0; jbx < nj; ++jbx) {
            int ib, jb;
            pair_bands(jpm, ibx, jbx, nocc_k, nocc_kq, nband, &ib, &jb);
            double eocc[4], eun[4], x[4];
            double mino = 1e300, maxu = -1e300, maxd = -1e300;
            for (int c = 0; c < 4; ++c) {
              const double a = e1[ib + (size_t)n1 * kk[c]], b = e2[jb + (size_t)n1 * kk[c]];
              eocc[c] = jpm == 0 ? a : b;
              eun[c] = jpm == 0 ? b : a;
              mino = fmin(mino, eocc[c]);
              maxu = fmax(maxu, eun[c]);
              maxd = fmax(maxd, eun[c] - eocc[c]);
            }
            if (!(mino <= ef && maxu >= ef)) continue;
            if (maxd < 0.0) continue;
            for (int c = 0; c < 4; ++c) x[c] = 0.5 * (eocc[c] - eun[c]);
            if (job == 0) {
              double dmx = -x[0], dmn = -x[0];
              for (int c = 1; c < 4; ++c) {
                dmx = fmax(dmx, -x[c]);
                dmn = fmin(dmn, -x[c]);
              }
              for (int c = 0; c < 4; ++c) {
                const size_t e = IDX4(ib, jb, kk[c], jpm);
                iwgt[e] = 1;
                if (dmx > demax[e]) demax[e] = dmx;
                if (dmn < demin[e]) demin[e] = dmn;
              }
              continue;
            }
            double* wt = (double*)calloc(4 * (size_t)nwhis, sizeof(double));
            tetctx cx;
            cx.ml = matrix_linear != 0;
            memcpy(cx.kkv, kvec, sizeof(cx.kkv));
            if (lindtet6(&cx, (const double(*)[3])kvec, eocc, eun, x, ef, ef, frhis, nwhis, wt)) {
              rc = -3;
              free(wt);
              break;
            }
            /* S/tetwt5.f90:329-340: corner c takes column c with matrix_linear, else column 1 times
             * 4 wtet (wtet = 1/4, :180) */
            for (int c = 0; c < 4; ++c) {
              const int ibib = ibjb[IDX4(ib, jb, kk[c], jpm)] - 1;
              const size_t o = (size_t)ibib + (size_t)nbnbx_cap * (kk[c] + (size_t)nqbz * jpm);
              const int jini = jhw[o] - 1, iini = ihw[o] - 1, nnn = nhw[o];
              if (cx.ml)
                for (int i = 0; i < nnn; ++i) whw[jini + i] += wt[iini + i + (size_t)nwhis * c] * piofvoltot;
              else
                for (int i = 0; i < nnn; ++i) whw[jini + i] += wt[iini + i] * piofvoltot * 4.0 * 0.25;
            }
            free(wt);
          }
      }
    }
    if (rc) break;
    /* symmetrisation over degenerate bands (S/tetwt5.f90:349-427): valence groups of each side,
     * plus the core groups on the occupied side (jpm = 0: k side; jpm = 1: k+q side) */
    int* i1 = (int*)malloc(sizeof(int) * 4 * (size_t)n1);
    int* e1g = i1 + n1;
    int* i2 = e1g + n1;
    int* e2g = i2 + n1;
    int ncr = 0;
    int* ic = (int*)malloc(sizeof(int) * 2 * (size_t)(nctot > 0 ? nctot : 1));
    int* ec = ic + (nctot > 0 ? nctot : 1);
    if (nctot) ncr = orc_chkdgn(ecore, nctot, ic, ec);
    for (int kx = 0; kx < nqbz && rc == 0; ++kx) {
      const int nr1 = orc_chkdgn(ek1 + (size_t)nband * kx, nband, i1, e1g);
      const int nr2 = orc_chkdgn(ek2 + (size_t)nband * kx, nband, i2, e2g);
      for (int jpm = 0; jpm < npm && rc == 0; ++jpm) {
        const int g1 = nr1 + (jpm == 0 ? ncr : 0), g2 = nr2 + (jpm == 1 ? ncr : 0);
        for (int r1 = 0; r1 < g1 && rc == 0; ++r1)
          for (int r2 = 0; r2 < g2; ++r2) {
            const int a0 = r1 < nr1 ? i1[r1] - 1 : nband + ic[r1 - nr1] - 1;
            const int a1 = r1 < nr1 ? e1g[r1] - 1 : nband + ec[r1 - nr1] - 1;
            const int b0 = r2 < nr2 ? i2[r2] - 1 : nband + ic[r2 - nr2] - 1;
            const int b1 = r2 < nr2 ? e2g[r2] - 1 : nband + ec[r2 - nr2] - 1;
            if (job == 0) {
              unsigned char w = 0;
              double mx = -1e300, mn = 1e300;
              for (int a = a0; a <= a1; ++a)
                for (int b = b0; b <= b1; ++b) {
                  w |= iwgt[IDX4(a, b, kx, jpm)];
                  mx = fmax(mx, demax[IDX4(a, b, kx, jpm)]);
                  mn = fmin(mn, demin[IDX4(a, b, kx, jpm)]);
                }
              for (int a = a0; a <= a1; ++a)
                for (int b = b0; b <= b1; ++b) {
                  iwgt[IDX4(a, b, kx, jpm)] = w;
                  demax[IDX4(a, b, kx, jpm)] = mx;
                  demin[IDX4(a, b, kx, jpm)] = mn;
                }
            } else {
              int isum = 0;
              for (int a = a0; a <= a1; ++a)
                for (int b = b0; b <= b1; ++b) isum += ibjb[IDX4(a, b, kx, jpm)] != 0;
              const int bs = (a1 - a0 + 1) * (b1 - b0 + 1);
              if (isum != 0 && isum != bs) {
                rc = -4;
                break;
              }
              if (isum == 0 || bs == 1) continue;
              double* wt = (double*)calloc(nwhis, sizeof(double));
              for (int a = a0; a <= a1; ++a)
                for (int b = b0; b <= b1; ++b) {
                  const size_t o = (size_t)(ibjb[IDX4(a, b, kx, jpm)] - 1) + (size_t)nbnbx_cap * (kx + (size_t)nqbz * jpm);
                  for (int i = 0; i < nhw[o]; ++i) wt[ihw[o] - 1 + i] += whw[jhw[o] - 1 + i];
                }
              for (int i = 0; i < nwhis; ++i) wt[i] /= bs;
              for (int a = a0; a <= a1; ++a)
                for (int b = b0; b <= b1; ++b) {
                  const size_t o = (size_t)(ibjb[IDX4(a, b, kx, jpm)] - 1) + (size_t)nbnbx_cap * (kx + (size_t)nqbz * jpm);
                  for (int i = 0; i < nhw[o]; ++i) whw[jhw[o] - 1 + i] = wt[ihw[o] - 1 + i];
                }
              free(wt);
            }
          }
      }
    }
    free(i1);
    free(ic);
    if (rc || job == 1) break;
    /* index sets per (k, jpm): nbnb, n1b/n2b (rsvwwk00_4: ib-major), ihw/nhw (hisrange), jhw, ibjb */
    int nbnbx = 1;
    for (int jpm = 0; jpm < npm; ++jpm)
      for (int kx = 0; kx < nqbz; ++kx) {
        int c = 0;
        for (int ib = 0; ib < n1; ++ib)
          for (int jb = 0; jb < n2; ++jb) c += iwgt[IDX4(ib, jb, kx, jpm)];
        nbnb[kx + (size_t)nqbz * jpm] = c;
        if (c > nbnbx) nbnbx = c;
      }
    *nbnbx_out = nbnbx;
    if (nbnbx > nbnbx_cap) {
      rc = -1;
      break;
    }
    int jtot = 1; /* whw pointers run over jpm, then k, then the pairs (S/m_tetwt.f90:182-196) */
    for (int jpm = 0; jpm < npm && rc == 0; ++jpm)
      for (int kx = 0; kx < nqbz && rc == 0; ++kx) {
        int ix = 0;
        for (int ib = 0; ib < n1 && rc == 0; ++ib)
          for (int jb = 0; jb < n2; ++jb)
            if (iwgt[IDX4(ib, jb, kx, jpm)]) {
              const size_t o = (size_t)ix + (size_t)nbnbx_cap * (kx + (size_t)nqbz * jpm);
              n1b[o] = ib + 1;
              n2b[o] = jb + 1;
              if (orc_hisrange(frhis, nwhis, demin[IDX4(ib, jb, kx, jpm)], demax[IDX4(ib, jb, kx, jpm)], &ihw[o],
                               &nhw[o])) {
                rc = -2;
                break;
              }
              jhw[o] = jtot;
              jtot += nhw[o];
              ibjb[IDX4(ib, jb, kx, jpm)] = ix + 1;
              ++ix;
            }
      }
    *nhwtot_out = jtot - 1;
    if (rc == 0 && jtot - 1 > nhw_cap) rc = -1;
    if (rc == 0) memset(whw, 0, sizeof(double) * (size_t)(jtot - 1));
  }
  free(e1);
  free(e2);
  free(iwgt);
  free(demin);
  free(demax);
  free(ibjb);
  return rc;
}

int orc_tetwt(int nband, int nqbz, int ntetf, const int* idtetf, const double* qbzw, const int* ib1bz,
              const double* ek1, const double* ek2, double ef, double ebmx, const double* qlat,
              const double* frhis, int nwhis, int nbnbx_cap, int nhw_cap, int* nbnb, int* n1b, int* n2b,
              int* ihw, int* nhw, int* jhw, double* whw, int* nbnbx_out, int* nhwtot_out) {
  return orc_tetwt_full(nband, nqbz, ntetf, idtetf, qbzw, ib1bz, ek1, ek2, ef, ebmx, qlat, frhis, nwhis, 1, 0, NULL, 0,
                        nbnbx_cap, nhw_cap, nbnb, n1b, n2b, ihw, nhw, jhw, whw, nbnbx_out, nhwtot_out);
}
