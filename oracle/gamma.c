/*
 * gamma.c -- FP64 CPU restatement of the Gamma-point pieces of the W build:
 * the q = 0 body of WVIllwI / WVRllwR (the mu = 1 row and column left out of
 * the inversion), the head llw = 1/epsinv(1,1) at the offset points q0i with
 * the wing of the static inverse, getw0 (the Gamma-cell average of W - v from
 * llw), and modifyWV0 (W(1,1) of the q = 0 records replaced, then E -> M).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  S/ = SRC/subroutines/ of ecalj.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PI 3.14159265358979323846

/* ---------------------------------------------------------------------- */
/* sylm (S/sylm.f90:1-81): unnormalized real harmonic polynomials r^l Y_L / cy_L
 * at r, L = l*l + l + m + 1 (m = -l..l): cos(m phi) for m > 0, sin(|m| phi)
 * for m < 0, no Condon-Shortley sign.  Restated with the same recursions
 * (z-recursion for P(l,l-1..1) when x^2+y^2 <= z^2, the other otherwise).   */
void orc_sylm(const double* r, double* yl, int lmx) {
  const int n = (lmx + 1) * (lmx + 1);
  double c[32], s[32], p[32][32];
  const double x = r[0], y = r[1], z = r[2];
  const double st = x * x + y * y, z2 = z * z, r2 = st + z2;
  yl[0] = 1.0;
  if (n < 2) return;
  if (r2 <= 1e-28) {
    for (int i = 1; i < n; ++i) yl[i] = 0.0;
    return;
  }
  c[1] = 1.0; s[1] = 0.0; p[1][1] = 1.0; p[2][2] = 1.0;  /* 1-based as in the reference */
  c[2] = x; s[2] = y; p[2][1] = z;
  yl[1] = y; yl[2] = z; yl[3] = x;
  int nt = 1;
  for (int l = 2; l <= lmx; ++l) {
    const int lp1 = l + 1, lm1 = l - 1, lav = l * lp1 + 1;
    p[lp1][1] = ((l + lm1) * z * p[l][1] - lm1 * r2 * p[lm1][1]) / l;
    yl[lav - 1] = p[lp1][1];
    nt += 2;
    p[lp1][lp1] = p[l][l] * nt;
    c[lp1] = x * c[l] - y * s[l];
    s[lp1] = x * s[l] + y * c[l];
    yl[lav + l - 1] = p[lp1][lp1] * c[lp1];
    yl[lav - l - 1] = p[lp1][lp1] * s[lp1];
    if (st <= z2) {
      for (int lmm = 1; lmm <= lm1; ++lmm) {
        const int m = l - lmm, mp1 = m + 1;
        p[lp1][mp1] = (r2 * (l + m) * p[l][mp1] - st * p[lp1][mp1 + 1]) / (z * (l - m));
        yl[lav + m - 1] = p[lp1][mp1] * c[mp1];
        yl[lav - m - 1] = p[lp1][mp1] * s[mp1];
      }
    } else {
      for (int m = 1; m <= lm1; ++m) {
        const int mp1 = m + 1;
        p[lp1][mp1] = ((lm1 + m) * r2 * p[l][m] - (lp1 - m) * z * p[lp1][m]) / st;
        yl[lav + m - 1] = p[lp1][mp1] * c[mp1];
        yl[lav - m - 1] = p[lp1][mp1] * s[mp1];
      }
    }
  }
}

/* sylmnc (S/sylm.f90:82-121): the normalization cy_L of the polynomials above. */
void orc_sylmnc(double* cy, int lmx) {
  const double tpi = 2.0 * PI, fpi = 4.0 * PI;
  cy[0] = 1.0 / sqrt(fpi);
  for (int l = 1; l <= lmx; ++l) {
    const int lp1 = l + 1, lav = l * lp1 + 1;
    const double tlp1 = l + lp1;
    cy[lav - 1] = sqrt(tlp1 / fpi);
    for (int m = 1; m <= l; ++m) {
      double fn2 = lp1 - m;
      for (int i = lp1 - m + 1; i <= l + m; ++i) fn2 *= i;
      cy[lav + m - 1] = sqrt(tlp1 / (fn2 * tpi));
      cy[lav - m - 1] = cy[lav + m - 1];
    }
  }
}

/* Real Gaunt coefficients g(L1, L2, L3) = int Y_L1 Y_L2 Y_L3 dOmega for
 * L1, L2 < (l12+1)^2 and L3 < (l3+1)^2, stored g[L1 + n12*(L2 + n12*L3)]:
 * what rotcg with the identity returns (S/rotcg.f90:54-69, from scg,
 * S/scg.f90:11-113).  Computed by exact quadrature of the harmonics of
 * orc_sylm: Gauss-Legendre in cos(theta) (40 points: exact to degree 79) and
 * the trapezoid rule in phi (96 points: exact to trigonometric degree 95).  */
int orc_gaunt(int l12, int l3, double* g) {
  if (l12 < 0 || l3 < 0 || 2 * l12 + l3 > 40) return -1;
  const int n12 = (l12 + 1) * (l12 + 1), n3 = (l3 + 1) * (l3 + 1);
  const int lm = l12 > l3 ? l12 : l3, nl = (lm + 1) * (lm + 1);
  const int nth = 40, nph = 96;
  double xt[40], wt[40];
  orc_gauss(nth, -1.0, 1.0, xt, wt);
  double* cy = (double*)malloc(sizeof(double) * nl);
  double* yl = (double*)malloc(sizeof(double) * nl);
  orc_sylmnc(cy, lm);
  memset(g, 0, sizeof(double) * (size_t)n12 * n12 * n3);
  for (int it = 0; it < nth; ++it) {
    const double ct = xt[it], stt = sqrt(fmax(0.0, 1.0 - ct * ct));
    for (int ip = 0; ip < nph; ++ip) {
      const double ph = 2.0 * PI * ip / nph;
      const double r[3] = {stt * cos(ph), stt * sin(ph), ct};
      orc_sylm(r, yl, lm);
      for (int i = 0; i < nl; ++i) yl[i] *= cy[i];
      const double w = wt[it] * 2.0 * PI / nph;
      for (int c = 0; c < n3; ++c)
        for (int b = 0; b < n12; ++b) {
          const double wb = w * yl[b] * yl[c];
          for (int a = 0; a < n12; ++a) g[a + n12 * (b + (size_t)n12 * c)] += wb * yl[a];
        }
    }
  }
  free(cy);
  free(yl);
  return 0;
}

static int lval(int lm) { /* ll (S/llnew.f90): l of the 1-based index lm */
  int l = 0;
  while ((l + 1) * (l + 1) < lm) ++l;
  return l;
}

/* getw0 (S/m_w0w0i.f90:146-311).  llw(ii:ie, nq0i) complex, column-major
 * with leading dimension ie-ii+1; dmlx(nq0i, 9); epinvq0i(nq0i, nq0i);
 * wklm((lmxax+1)^2); epinv(3, 3, nq0i).  Out: w0(ii:ie) and, when ii <= 0,
 * llmat(3, 3) at iw = 0.  Returns 0, or 2 when a matrix is singular. */
int orc_getw0(int ii, int ie, int nq0i, int lmxax, const double complex* llw, const double* dmlx,
              const double* epinvq0i, const double* wklm, double wbz, const double* epinv,
              double complex* w0, double complex* llmat) {
  const int nw = ie - ii + 1;
  const int nlklm = (lmxax + 1) * (lmxax + 1);
  const double fpi = 4.0 * PI;
  /* epinvq0i^-1 (:168-169) */
  double complex* em1 = (double complex*)malloc(sizeof(double complex) * nq0i * nq0i);
  for (int e = 0; e < nq0i * nq0i; ++e) em1[e] = epinvq0i[e];
  if (orc_zinv(nq0i, em1, nq0i) != 0) {
    free(em1);
    return 2;
  }
  double complex* invr = (double complex*)calloc((size_t)nw * nq0i, sizeof(double complex));
  double complex* llwyl = (double complex*)calloc((size_t)nw * 9, sizeof(double complex));
  for (int iw = 0; iw < nw; ++iw) { /* :171-180 */
    for (int i = 0; i < nq0i; ++i) {
      double complex s = 0.0;
      for (int j = 0; j < nq0i; ++j) s += creal(em1[i + nq0i * j]) * llw[iw + (long)nw * j];
      invr[iw + (long)nw * i] = s;
    }
    for (int lm = 1; lm <= 9; ++lm) {
      if (lm >= 2 && lm <= 4) continue;
      double complex s = 0.0;
      for (int i = 0; i < nq0i; ++i) s += dmlx[i + nq0i * (lm - 1)] * invr[iw + (long)nw * i];
      llwyl[(lm - 1) + 9 * iw] = s;
    }
  }
  if (ii <= 0 && llmat) { /* :182-188 */
    const int iw = -ii;
    for (int e = 0; e < 9; ++e) llmat[e] = 0.0;
    for (int i = 0; i < nq0i; ++i)
      for (int e = 0; e < 9; ++e) llmat[e] += invr[iw + (long)nw * i] * epinv[e + 9 * i];
  }
  const int lxx = lmxax > 2 ? lmxax : 2; /* :189-191 */
  const int n12 = (lxx + 1) * (lxx + 1);
  double* cg = (double*)malloc(sizeof(double) * (size_t)n12 * n12 * 9);
  orc_gaunt(lxx, 2, cg);
  int neven = 0;
  for (int lm = 1; lm <= nlklm; ++lm) neven += (lval(lm) % 2 == 0);
  double complex* m = (double complex*)malloc(sizeof(double complex) * neven * neven);
  int rc = 0;
  for (int iw = 0; iw < nw; ++iw) { /* :246-277 */
    int lm2x = 0;
    for (int lm2 = 1; lm2 <= nlklm; ++lm2) {
      if (lval(lm2) % 2 == 1) continue;
      int lm1x = 0;
      for (int lm1 = 1; lm1 <= nlklm; ++lm1) {
        if (lval(lm1) % 2 == 1) continue;
        /* cg(lm2, k, lm1) of rotcg; the coefficient is symmetric in its three
         * indices, so it is read as g[lm2, lm1, k] from the table with k <= 9 */
        double complex s = cg[(lm2 - 1) + n12 * ((lm1 - 1) + (size_t)n12 * 0)] * llwyl[0 + 9 * iw];
        double complex s2 = 0.0;
        for (int k = 5; k <= 9; ++k)
          s2 += cg[(lm2 - 1) + n12 * ((lm1 - 1) + (size_t)n12 * (k - 1))] * llwyl[(k - 1) + 9 * iw];
        m[lm1x + neven * lm2x] = s + s2;
        ++lm1x;
      }
      ++lm2x;
    }
    if (orc_zinv(neven, m, neven) != 0) {
      rc = 2;
      break;
    }
    double complex w = 0.0;
    int lm1x = 0;
    for (int lm1 = 1; lm1 <= nlklm; ++lm1) {
      if (lval(lm1) % 2 == 1) continue;
      w += wklm[lm1 - 1] * fpi * sqrt(fpi) * m[lm1x + neven * 0] / wbz;
      if (lm1 == 1) w -= wklm[0] * pow(fpi, 1.5) / wbz;
      ++lm1x;
    }
    w0[iw] = w;
  }
  free(em1);
  free(invr);
  free(llwyl);
  free(cg);
  free(m);
  return rc;
}

/* epstilde of the W build from chi (M or E basis) at one frequency:
 * x = fac*chi, M -> E (S/m_llw.f90:336-337), eps = delta - s_i x_ij s_j;
 * head > 0: eps(1,1) = 1 - head*x(1,1) (:219-222, :413-416). */
static void build_eps(int ngb, int nspin, int is_x0_m_basis, const double complex* c, const double* vcousq,
                      const double complex* m2e, double head, double complex* x, double complex* t,
                      double complex* eps) {
  const long n2 = (long)ngb * ngb;
  const double fac = (nspin == 1) ? 2.0 : 1.0;
  for (long e = 0; e < n2; ++e) x[e] = fac * c[e];
  if (is_x0_m_basis) {
    orc_zgemm('N', 'N', ngb, ngb, ngb, 1.0, x, ngb, m2e, ngb, 0.0, t, ngb);
    orc_zgemm('C', 'N', ngb, ngb, ngb, 1.0, m2e, ngb, t, ngb, 0.0, x, ngb);
  }
  for (int j = 0; j < ngb; ++j)
    for (int i = 0; i < ngb; ++i) {
      double complex e = -vcousq[i] * x[i + (long)ngb * j] * vcousq[j];
      if (i == j) e = 1.0 + e;
      eps[i + (long)ngb * j] = e;
    }
  if (head > 0.0) eps[0] = 1.0 - head * x[0];
}

/* q = 0 body (iq = 1) of WVRllwR / WVIllwI (S/m_llw.f90:98-178, 299-373 with
 * ix = 1): epsilon restricted to mu, nu >= 2 is inverted, W(1,:) = W(:,1) = 0,
 * and W stays in the E basis (the E -> M step skips iq = 1).  The real axis
 * (zminv with pivoting) and the imaginary axis give the same inverse up to
 * rounding, so one routine serves both.  wc: ldw x ngb x nw. */
int orc_wc_build_gamma(int ngb, int nw, int nspin, int is_x0_m_basis, const double complex* chi,
                       const double* vcousq, const double complex* m2e, double complex* wc, int ldw) {
  const long n2 = (long)ngb * ngb;
  const int m = ngb - 1;
  double complex* x = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* t = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* eps = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* sub = (double complex*)malloc(sizeof(double complex) * (m > 0 ? (long)m * m : 1));
  int rc = 0;
  for (int iw = 0; iw < nw; ++iw) {
    build_eps(ngb, nspin, is_x0_m_basis, chi + n2 * iw, vcousq, m2e, 0.0, x, t, eps);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) sub[i + (long)m * j] = eps[(i + 1) + (long)ngb * (j + 1)];
    if (m > 0 && orc_zinv(m, sub, m) != 0) rc = 2;
    double complex* w = wc + (long)ldw * ngb * iw;
    for (int j = 0; j < ngb; ++j)
      for (int i = 0; i < ngb; ++i) {
        double complex e = 0.0;
        if (i > 0 && j > 0) {
          e = vcousq[i] * sub[(i - 1) + (long)m * (j - 1)] * vcousq[j];
          if (i == j) e -= vcousq[i] * vcousq[j];
        }
        w[i + (long)ldw * j] = e;
      }
  }
  free(x);
  free(t);
  free(eps);
  free(sub);
  return rc;
}

/* Offset points q0i (iq > nqibz) of WVRllwR / WVIllwI (S/m_llw.f90:183-256,
 * 378-440): epsilon with the head 1 - vcou1*chi(1,1), full inverse,
 * llw(iw) = 1/epsinv(1,1); when 0 <= wing_iw < nw,
 * wing(mu-1) = epsinv(1,mu)/epsinv(1,1), mu = 2..ngb, at that frequency. */
int orc_eps_head(int ngb, int nw, int nspin, int is_x0_m_basis, const double complex* chi, const double* vcousq,
                 const double complex* m2e, double vcou1, double complex* llw, int wing_iw, double complex* wing) {
  const long n2 = (long)ngb * ngb;
  double complex* x = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* t = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* eps = (double complex*)malloc(sizeof(double complex) * n2);
  int rc = 0;
  for (int iw = 0; iw < nw; ++iw) {
    build_eps(ngb, nspin, is_x0_m_basis, chi + n2 * iw, vcousq, m2e, vcou1, x, t, eps);
    if (orc_zinv(ngb, eps, ngb) != 0) {
      rc = 2;
      continue;
    }
    llw[iw] = 1.0 / eps[0];
    if (iw == wing_iw && wing)
      for (int j = 1; j < ngb; ++j) wing[j - 1] = eps[(long)ngb * j] / eps[0];
  }
  free(x);
  free(t);
  free(eps);
  return rc;
}

/* modifyWV0 (S/m_w0w0i.f90:89-140) on records held in memory: W(1,1,iw) =
 * w0(iw), then W <- m2e W m2e^H when is_wc_m_basis.  w: ldw x ngb x nw. */
void orc_wv0_patch(int ngb, int nw, double complex* w, int ldw, const double complex* w0, int is_wc_m_basis,
                   const double complex* m2e) {
  const long n2 = (long)ngb * ngb;
  double complex* a = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* b = (double complex*)malloc(sizeof(double complex) * n2);
  for (int iw = 0; iw < nw; ++iw) {
    double complex* z = w + (long)ldw * ngb * iw;
    z[0] = w0[iw];
    if (!is_wc_m_basis) continue;
    orc_zgemm('N', 'C', ngb, ngb, ngb, 1.0, z, ldw, m2e, ngb, 0.0, a, ngb);
    orc_zgemm('N', 'N', ngb, ngb, ngb, 1.0, m2e, ngb, a, ngb, 0.0, b, ngb);
    for (int j = 0; j < ngb; ++j)
      for (int i = 0; i < ngb; ++i) z[i + (long)ldw * j] = b[i + (long)ngb * j];
  }
  free(a);
  free(b);
}
