/*
 * tetwt.c -- FP64 CPU restatement of the tetrahedron histogram weights of
 * the chi0 path (gettetwt + tetwt5x_dtet4).  TEST INFRASTRUCTURE ONLY: the
 * checker of the HIP implementation, loaded only by tests/, smoke() and the
 * cpu_baseline leg of bench.py.
 *
 * Scope (the reference's default path): npm = 1 (time reversal), no core
 * states (nctot = 0), nmtet = 1, tetrahedron_matrix_linear = F, no tetrakbt,
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

/* intttvc6, matrix_linear = F: wtthis(ihis) += int_bin d omega int d^3k delta(omega + v(k)) */
static int intttvc6(const double v[4], double voltet, const double* frhis, int nwhis, double* wtthis) {
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
    wtthis[ihis - 1] += pvn * (integb - intega);
    intega = integb;
  }
  return 0;
}

/* corners: kk[8][3], xx[8]; itetx 1-based as in the reference */
static int inttetrac6(double kk[8][3], const double xx[8], const int itetx[4], const double* frhis, int nwhis,
                      double* wtthis) {
  double xin[4], kin[4][3], am[3][3];
  for (int i = 0; i < 4; ++i) {
    xin[i] = xx[itetx[i] - 1];
    for (int d = 0; d < 3; ++d) kin[i][d] = kk[itetx[i] - 1][d];
  }
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) am[i][d] = kin[i][d] - kin[3][d];
  const double voltet = fabs(det33(am) / 6.0);
  return intttvc6(xin, voltet, frhis, nwhis, wtthis);
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
static int inttetra6(double kk_[8][3], const double xx_[8], const double ebf[8], const int itetx[4],
                     const double* frhis, int nwhis, double* wtthis) {
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
    rc |= inttetrac6(kk, xx, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
  } else if (ee[2] > 0.0 && 0.0 >= ee[1]) { /* Fig 2 */
    midk(kk, ee, xx, 4, 2, 5);
    midk(kk, ee, xx, 4, 1, 6);
    midk(kk, ee, xx, 3, 1, 7);
    midk(kk, ee, xx, 3, 2, 8);
    rc |= inttetrac6(kk, xx, (const int[]){4, 3, 5, 6}, frhis, nwhis, wtthis);
    rc |= inttetrac6(kk, xx, (const int[]){3, 6, 7, 5}, frhis, nwhis, wtthis);
    rc |= inttetrac6(kk, xx, (const int[]){3, 5, 7, 8}, frhis, nwhis, wtthis);
  } else if (ee[1] > 0.0 && 0.0 >= ee[0]) { /* Fig 3 */
    midk(kk, ee, xx, 1, 4, 5);
    midk(kk, ee, xx, 1, 2, 6);
    midk(kk, ee, xx, 1, 3, 7);
    rc |= inttetrac6(kk, xx, (const int[]){3, 4, 7, 2}, frhis, nwhis, wtthis);
    rc |= inttetrac6(kk, xx, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
    rc |= inttetrac6(kk, xx, (const int[]){4, 2, 6, 7}, frhis, nwhis, wtthis);
  } else { /* Fig 4 */
    rc |= inttetrac6(kk, xx, (const int[]){1, 2, 3, 4}, frhis, nwhis, wtthis);
  }
  return rc;
}

/* occupied cut (ea < E_F) of one tetrahedron, then the unoccupied cut of each piece:
 * wtthis(ihis) = int_bin d omega int d^3k f(ea)(1-f(eb)) delta(omega + x(k)) */
int orc_lindtet6(const double kvec[4][3], const double ea[4], const double eb[4], const double x[4],
                 double efermia, double efermib, const double* frhis, int nwhis, double* wtthis) {
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
    rc |= inttetra6(kk, xx, ebf, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
  } else if (ee[2] < 0.0 && 0.0 <= ee[1]) { /* Fig 2 */
    midk3(kk, ee, xx, ebf, 4, 2, 5);
    midk3(kk, ee, xx, ebf, 4, 1, 6);
    midk3(kk, ee, xx, ebf, 3, 1, 7);
    midk3(kk, ee, xx, ebf, 3, 2, 8);
    rc |= inttetra6(kk, xx, ebf, (const int[]){4, 3, 5, 6}, frhis, nwhis, wtthis);
    rc |= inttetra6(kk, xx, ebf, (const int[]){3, 6, 7, 5}, frhis, nwhis, wtthis);
    rc |= inttetra6(kk, xx, ebf, (const int[]){3, 5, 7, 8}, frhis, nwhis, wtthis);
  } else if (ee[1] < 0.0 && 0.0 <= ee[0]) { /* Fig 3 */
    midk3(kk, ee, xx, ebf, 1, 4, 5);
    midk3(kk, ee, xx, ebf, 1, 2, 6);
    midk3(kk, ee, xx, ebf, 1, 3, 7);
    rc |= inttetra6(kk, xx, ebf, (const int[]){3, 4, 7, 2}, frhis, nwhis, wtthis);
    rc |= inttetra6(kk, xx, ebf, (const int[]){4, 5, 6, 7}, frhis, nwhis, wtthis);
    rc |= inttetra6(kk, xx, ebf, (const int[]){4, 2, 6, 7}, frhis, nwhis, wtthis);
  } else { /* Fig 4 */
    rc |= inttetra6(kk, xx, ebf, (const int[]){1, 2, 3, 4}, frhis, nwhis, wtthis);
  }
  return rc;
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

#define IDX3(ib, jb, k) ((size_t)(ib) + (size_t)nband * ((size_t)(jb) + (size_t)nband * (size_t)(k)))

int orc_tetwt(int nband, int nqbz, int ntetf, const int* idtetf, const double* qbzw, const int* ib1bz,
              const double* ek1, const double* ek2, double ef, double ebmx, const double* qlat,
              const double* frhis, int nwhis, int nbnbx_cap, int nhw_cap, int* nbnb, int* n1b, int* n2b,
              int* ihw, int* nhw, int* jhw, double* whw, int* nbnbx_out, int* nhwtot_out) {
  double qb[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) qb[c][r] = qlat[r + 3 * c];
  const double voltot = fabs(det33(qb));
  const double piofvoltot = PI / 4.0 / voltot;
  const size_t nbb = (size_t)nband * nband * nqbz;
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
      int noccx_k = 0, nbmx_kq = nband;
      for (int c = 0; c < 4; ++c) {
        int n1 = 0, n2 = 0;
        for (int b = 0; b < nband; ++b) {
          n1 += ek1[b + (size_t)nband * kk[c]] < ef;
          n2 += ek2[b + (size_t)nband * kk[c]] < fmin(1e10, ebmx);
        }
        if (n1 > noccx_k) noccx_k = n1;
        if (n2 < nbmx_kq) nbmx_kq = n2;
      }
      for (int ib = 0; ib < noccx_k; ++ib)
        for (int jb = 0; jb < nbmx_kq; ++jb) {
          double eocc[4], eun[4], x[4];
          double mino = 1e300, maxu = -1e300, maxd = -1e300;
          for (int c = 0; c < 4; ++c) {
            eocc[c] = ek1[ib + (size_t)nband * kk[c]];
            eun[c] = ek2[jb + (size_t)nband * kk[c]];
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
              const size_t e = IDX3(ib, jb, kk[c]);
              iwgt[e] = 1;
              if (dmx > demax[e]) demax[e] = dmx;
              if (dmn < demin[e]) demin[e] = dmn;
            }
            continue;
          }
          double* wt = (double*)calloc(nwhis, sizeof(double));
          if (orc_lindtet6((const double(*)[3])kvec, eocc, eun, x, ef, ef, frhis, nwhis, wt)) {
            rc = -3;
            free(wt);
            break;
          }
          for (int c = 0; c < 4; ++c) {
            const int ibib = ibjb[IDX3(ib, jb, kk[c])] - 1;
            const size_t o = (size_t)ibib + (size_t)nbnbx_cap * kk[c];
            const int jini = jhw[o] - 1, iini = ihw[o] - 1, nnn = nhw[o];
            for (int i = 0; i < nnn; ++i) whw[jini + i] += wt[iini + i] * piofvoltot * 4.0 * 0.25;
          }
          free(wt);
        }
    }
    if (rc) break;
    /* symmetrisation over degenerate bands */
    int* i1 = (int*)malloc(sizeof(int) * 4 * nband);
    int* e1 = i1 + nband;
    int* i2 = e1 + nband;
    int* e2 = i2 + nband;
    for (int kx = 0; kx < nqbz && rc == 0; ++kx) {
      const int nr1 = orc_chkdgn(ek1 + (size_t)nband * kx, nband, i1, e1);
      const int nr2 = orc_chkdgn(ek2 + (size_t)nband * kx, nband, i2, e2);
      for (int r1 = 0; r1 < nr1 && rc == 0; ++r1)
        for (int r2 = 0; r2 < nr2; ++r2) {
          const int a0 = i1[r1] - 1, a1 = e1[r1] - 1, b0 = i2[r2] - 1, b1 = e2[r2] - 1;
          if (job == 0) {
            unsigned char w = 0;
            double mx = -1e300, mn = 1e300;
            for (int a = a0; a <= a1; ++a)
              for (int b = b0; b <= b1; ++b) {
                w |= iwgt[IDX3(a, b, kx)];
                mx = fmax(mx, demax[IDX3(a, b, kx)]);
                mn = fmin(mn, demin[IDX3(a, b, kx)]);
              }
            for (int a = a0; a <= a1; ++a)
              for (int b = b0; b <= b1; ++b) {
                iwgt[IDX3(a, b, kx)] = w;
                demax[IDX3(a, b, kx)] = mx;
                demin[IDX3(a, b, kx)] = mn;
              }
          } else {
            int isum = 0;
            for (int a = a0; a <= a1; ++a)
              for (int b = b0; b <= b1; ++b) isum += ibjb[IDX3(a, b, kx)] != 0;
            const int bs = (a1 - a0 + 1) * (b1 - b0 + 1);
            if (isum != 0 && isum != bs) {
              rc = -4;
              break;
            }
            if (isum == 0 || bs == 1) continue;
            double* wt = (double*)calloc(nwhis, sizeof(double));
            for (int a = a0; a <= a1; ++a)
              for (int b = b0; b <= b1; ++b) {
                const size_t o = (size_t)(ibjb[IDX3(a, b, kx)] - 1) + (size_t)nbnbx_cap * kx;
                for (int i = 0; i < nhw[o]; ++i) wt[ihw[o] - 1 + i] += whw[jhw[o] - 1 + i];
              }
            for (int i = 0; i < nwhis; ++i) wt[i] /= bs;
            for (int a = a0; a <= a1; ++a)
              for (int b = b0; b <= b1; ++b) {
                const size_t o = (size_t)(ibjb[IDX3(a, b, kx)] - 1) + (size_t)nbnbx_cap * kx;
                for (int i = 0; i < nhw[o]; ++i) whw[jhw[o] - 1 + i] = wt[ihw[o] - 1 + i];
              }
            free(wt);
          }
        }
    }
    free(i1);
    if (rc || job == 1) break;
    /* index sets: nbnb, n1b/n2b (rsvwwk00_4), ihw/nhw (hisrange), jhw, ibjb */
    int nbnbx = 1;
    for (int kx = 0; kx < nqbz; ++kx) {
      int c = 0;
      for (int ib = 0; ib < nband; ++ib)
        for (int jb = 0; jb < nband; ++jb) c += iwgt[IDX3(ib, jb, kx)];
      nbnb[kx] = c;
      if (c > nbnbx) nbnbx = c;
    }
    *nbnbx_out = nbnbx;
    if (nbnbx > nbnbx_cap) {
      rc = -1;
      break;
    }
    int jtot = 1;
    for (int kx = 0; kx < nqbz && rc == 0; ++kx) {
      int ix = 0;
      for (int ib = 0; ib < nband; ++ib)
        for (int jb = 0; jb < nband; ++jb)
          if (iwgt[IDX3(ib, jb, kx)]) {
            const size_t o = (size_t)ix + (size_t)nbnbx_cap * kx;
            n1b[o] = ib + 1;
            n2b[o] = jb + 1;
            if (orc_hisrange(frhis, nwhis, demin[IDX3(ib, jb, kx)], demax[IDX3(ib, jb, kx)], &ihw[o], &nhw[o])) {
              rc = -2;
              break;
            }
            jhw[o] = jtot;
            jtot += nhw[o];
            ibjb[IDX3(ib, jb, kx)] = ix + 1;
            ++ix;
          }
    }
    *nhwtot_out = jtot - 1;
    if (rc == 0 && jtot - 1 > nhw_cap) rc = -1;
    if (rc == 0) memset(whw, 0, sizeof(double) * (size_t)(jtot - 1));
  }
  free(iwgt);
  free(demin);
  free(demax);
  free(ibjb);
  return rc;
}
