/*
 * oracle.c -- FP64 CPU restatement of ecalj's QSGW chi0 -> W -> Sigma_c path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of the
 * product path.  Each function cites the reference routine it follows
 * (S/ = SRC/subroutines/ of ecalj).  OpenMP is used only to parallelise over
 * independent output columns; summation order inside one output element is
 * the reference's order.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846

/* ---------------------------------------------------------------------- */
/* zmm_h (S/m_blas.f90:205-232): C = alpha op(A) op(B) + beta C, column-major.
 * The reference calls zgemm3m; this restates plain zgemm semantics.       */
static inline double complex opget(char op, const double complex* a, long lda, long i, long l) {
  /* element (i,l) of op(A) */
  if (op == 'N') return a[i + lda * l];
  if (op == 'T') return a[l + lda * i];
  return conj(a[l + lda * i]);
}

/* Optional LP64 Fortran-ABI BLAS/LAPACK (orc_use_blas).  The reference calls
 * zgemm(3m) and zgetrf+zgetri from MKL (S/m_blas.f90:225, S/m_lapack.f90:21-47);
 * with the hooks set the oracle calls the same routines of another BLAS, and
 * without them it runs the loops below (same results up to rounding).     */
typedef void (*zgemm_fn)(char*, char*, int*, int*, int*, double complex*, const double complex*, int*,
                         const double complex*, int*, double complex*, double complex*, int*);
typedef void (*zgetrf_fn)(int*, int*, double complex*, int*, int*, int*);
typedef void (*zgetri_fn)(int*, double complex*, int*, int*, double complex*, int*, int*);
static zgemm_fn blas_zgemm = NULL;
static zgetrf_fn lapack_zgetrf = NULL;
static zgetri_fn lapack_zgetri = NULL;

void orc_use_blas(void* zgemm, void* zgetrf, void* zgetri) {
  blas_zgemm = (zgemm_fn)zgemm;
  lapack_zgetrf = (zgetrf_fn)zgetrf;
  lapack_zgetri = (zgetri_fn)zgetri;
}

int orc_blas_active(void) { return blas_zgemm != NULL; }

static int fits_int(long v) { return v >= 0 && v <= 2147483647L; }

void orc_zgemm(char opa, char opb, int m, int n, int k, double complex alpha,
               const double complex* a, long lda, const double complex* b, long ldb,
               double complex beta, double complex* c, long ldc) {
  if (m <= 0 || n <= 0) return;
  if (blas_zgemm && fits_int(lda) && fits_int(ldb) && fits_int(ldc) && k > 0) {
    int im = m, in = n, ik = k, ilda = lda > 0 ? (int)lda : 1, ildb = ldb > 0 ? (int)ldb : 1, ildc = (int)ldc;
    char ta = opa, tb = opb;
    blas_zgemm(&ta, &tb, &im, &in, &ik, &alpha, a, &ilda, b, &ildb, &beta, c, &ildc);
    return;
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < n; ++j) {
    double complex* cj = c + ldc * (long)j;
    if (beta == 0.0) {
      for (int i = 0; i < m; ++i) cj[i] = 0.0;
    } else if (beta != 1.0) {
      for (int i = 0; i < m; ++i) cj[i] *= beta;
    }
    if (opa == 'N') {
      /* axpy form: contiguous in i */
      for (int l = 0; l < k; ++l) {
        double complex blj = alpha * opget(opb, b, ldb, l, j);
        if (blj == 0.0) continue;
        const double complex* al = a + lda * (long)l;
        for (int i = 0; i < m; ++i) cj[i] += al[i] * blj;
      }
    } else {
      /* dot form: op(A)(i,l) = A(l,i) or conj(A(l,i)), contiguous in l */
      double complex* bcol = NULL;
      double complex bbuf[512];
      if (k <= 512) bcol = bbuf; else bcol = (double complex*)malloc(sizeof(double complex) * k);
      for (int l = 0; l < k; ++l) bcol[l] = opget(opb, b, ldb, l, j);
      for (int i = 0; i < m; ++i) {
        const double complex* ai = a + lda * (long)i;
        double sr = 0.0, si = 0.0;
        if (opa == 'C') {
          for (int l = 0; l < k; ++l) {
            double ar = creal(ai[l]), aim = -cimag(ai[l]);
            double br = creal(bcol[l]), bi = cimag(bcol[l]);
            sr += ar * br - aim * bi;
            si += ar * bi + aim * br;
          }
        } else {
          for (int l = 0; l < k; ++l) {
            double ar = creal(ai[l]), aim = cimag(ai[l]);
            double br = creal(bcol[l]), bi = cimag(bcol[l]);
            sr += ar * br - aim * bi;
            si += ar * bi + aim * br;
          }
        }
        cj[i] += alpha * (sr + I * si);
      }
      if (bcol != bbuf) free(bcol);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* gauss (S/mate.f90:1-206).  The reference tabulates 15-digit nodes for
 * n in {6,10,12,16,20,24,32,40,48} and stops otherwise (:113-124).  Here the
 * nodes are the roots of P_n found by Newton iteration (textbook method),
 * weights 2/((1-x^2) P_n'(x)^2), mapped to [a,b] as x = e2 +- e1*t
 * (S/mate.f90:110-111).  Node order follows the reference: ascending.   */
int orc_gauss(int n, double a, double b, double* x, double* w) {
  static const int ok[] = {6, 10, 12, 16, 20, 24, 32, 40, 48};
  int supported = 0;
  for (int i = 0; i < 9; ++i) supported |= (ok[i] == n);
  if (!supported) return -1;
  double e1 = 0.5 * (b - a), e2 = e1 + a;
  int half = n / 2;
  for (int i = 1; i <= half; ++i) {
    /* i-th positive root counted from the largest */
    double t = cos(PI * (i - 0.25) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int j = 2; j <= n; ++j) {
        double p2 = ((2.0 * j - 1.0) * t * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      double dt = p1 / dp;
      t -= dt;
      if (fabs(dt) < 1e-16) break;
    }
    {
      double p0 = 1.0, p1 = t;
      for (int j = 2; j <= n; ++j) {
        double p2 = ((2.0 * j - 1.0) * t * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
    }
    double wt = 2.0 / ((1.0 - t * t) * dp * dp);
    /* roots t_i descending in i; reference stores x(half+i')=e2+e1*t ascending */
    int ip = half - i + 1; /* 1-based index of this root among positive roots, ascending */
    x[half + ip - 1] = e2 + e1 * t;
    x[half - ip] = e2 - e1 * t;
    w[half + ip - 1] = wt * e1;
    w[half - ip] = wt * e1;
  }
  return 0;
}

/* freq01 (S/m_freq.f90:164-186); expa as in readfreq_r (S/m_readfreq_r.f90:44-49). */
int orc_freq01(int nx, double ua, double* freqx, double* freqw, double* wx, double* expa) {
  if (orc_gauss(nx, 0.0, 1.0, freqx, wx) != 0) return -1;
  for (int i = 0; i < nx; ++i) {
    freqw[i] = (1.0 - freqx[i]) / freqx[i];
    if (expa) expa[i] = exp(-ua * ua * freqw[i] * freqw[i]);
  }
  return 0;
}

/* Histogram bins (S/m_freq.f90:95-108). */
int orc_frhis(double oratio, double dw, double omg2max, double* frhis) {
  double aa = oratio - 1.0, bb = dw / aa;
  int iw = 0;
  for (;;) {
    iw += 1;
    if (bb * (exp(aa * (iw - 1)) - 1.0) > omg2max + 1e-6) break;
  }
  int nwhis = iw + 2;
  if (frhis)
    for (int i = 1; i <= nwhis + 1; ++i) frhis[i - 1] = bb * (exp(aa * (i - 1)) - 1.0);
  return nwhis;
}

/* ---------------------------------------------------------------------- */
/* hilbertmat (S/hilbertmat.f90:14-79); inner hilbertmat2 (:56-70).        */
static void hilbertmat2(double complex zz, double a, double b, double complex* facm,
                        double complex* facn) {
  const double eps = 1e-8;
  double complex db = zz - b, da = zz - a;
  *facm = 0.0;
  *facn = 0.0;
  if (cabs(da) > eps && cabs(db) > eps) {
    *facn = clog(db / da);
    *facm = zz * (*facn) + (b - a);
  }
}

void orc_hilbertmat(double complex zz, int nwhis, const double* his_L, const double* his_C,
                    const double* his_R, double complex* rmat) {
  const double epsz = 1e-13;
  const double complex zze = zz + I * epsz;
  /* all arrays indexed -nwhis..nwhis -> offset nwhis */
  const double* L = his_L + nwhis;
  const double* C = his_C + nwhis;
  const double* R = his_R + nwhis;
  double complex* rm = rmat + nwhis;
  for (int i = -nwhis; i <= nwhis; ++i) rm[i] = 0.0;
  for (int iw = -nwhis; iw <= nwhis; ++iw) {
    if (iw == 0) continue;
    double ci = C[iw], cim, cip;
    double complex facm, facn;
    if (iw == -nwhis) cim = L[iw];
    else if (iw == 1) cim = 0.0;
    else cim = C[iw - 1];
    double cicim = ci - cim;
    hilbertmat2(zze, L[iw], C[iw], &facm, &facn);
    rm[iw] += (facm - facn * cim) / cicim;
    if (iw != -nwhis && iw != 1) rm[iw - 1] += (-facm + facn * ci) / cicim;
    if (iw == nwhis) cip = R[iw];
    else if (iw == -1) cip = 0.0;
    else cip = C[iw + 1];
    double cipci = cip - ci;
    hilbertmat2(zze, C[iw], R[iw], &facm, &facn);
    rm[iw] += (-facm + facn * cip) / cipci;
    if (iw != nwhis && iw != -1) rm[iw + 1] += (facm - facn * ci) / cipci;
  }
}

/* his_L/his_C/his_R from frhis (S/dpsion5.f90:84-86). */
static void his_edges(int nwhis, const double* frhis, double* L, double* C, double* R) {
  /* index j = iw + nwhis, iw in -nwhis..nwhis; frhis is 1-based frhis(1:nwhis+1) */
  for (int iw = -nwhis; iw <= nwhis; ++iw) {
    int j = iw + nwhis;
    if (iw < 0) {
      L[j] = -frhis[-iw];     /* -frhis(-iw+1) */
      R[j] = -frhis[-iw - 1]; /* -frhis(-iw)   */
    } else if (iw == 0) {
      L[j] = 0.0; /* middle element of both source lists is 0 */
      R[j] = 0.0;
    } else {
      L[j] = frhis[iw - 1]; /* frhis(iw)   */
      R[j] = frhis[iw];     /* frhis(iw+1) */
    }
    C[j] = 0.5 * (L[j] + R[j]);
  }
}

/* dpsion_init imaginary branch, npm=1 (S/dpsion5.f90:121-129). */
void orc_imatt(int niw, const double* freqi, int nwhis, const double* frhis, double* imatt) {
  int n2 = 2 * nwhis + 1;
  double* L = (double*)malloc(sizeof(double) * n2 * 3);
  double* C = L + n2;
  double* R = C + n2;
  his_edges(nwhis, frhis, L, C, R);
#pragma omp parallel for
  for (int it = 0; it < niw; ++it) {
    double complex* rrr = (double complex*)malloc(sizeof(double complex) * n2);
    orc_hilbertmat(I * freqi[it], nwhis, L, C, R, rrr);
    for (int iw = 1; iw <= nwhis; ++iw)
      imatt[it + (long)niw * (iw - 1)] = creal(rrr[nwhis + iw] - rrr[nwhis - iw]) / PI;
    free(rrr);
  }
  free(L);
}

/* dpsion_init real branch, npm=1 without chipm (S/dpsion5.f90:91-103):
 * rmatt(it,iw) = Re(rmat(iw) - rmat(-iw))/pi of hilbertmat at zz = freqr(it)
 * (real; hilbertmat adds i*1e-13).  rmatt is (nw+1) x nwhis column-major. */
void orc_rmatt(int nw, const double* freqr, int nwhis, const double* frhis, double* rmatt) {
  int n2 = 2 * nwhis + 1;
  double* L = (double*)malloc(sizeof(double) * n2 * 3);
  double* C = L + n2;
  double* R = C + n2;
  his_edges(nwhis, frhis, L, C, R);
#pragma omp parallel for
  for (int it = 0; it <= nw; ++it) {
    double complex* rrr = (double complex*)malloc(sizeof(double complex) * n2);
    orc_hilbertmat(freqr[it], nwhis, L, C, R, rrr);
    for (int iw = 1; iw <= nwhis; ++iw)
      rmatt[it + (long)(nw + 1) * (iw - 1)] = creal(rrr[nwhis + iw] - rrr[nwhis - iw]) / PI;
    free(rrr);
  }
  free(L);
}

/* dpsion_chiq real axis, npm=1 without chipm (S/dpsion5.f90:193-200,
 * 244-256).  Unless `scaled`, first scales rcxq(:,:,1:nwhis) in place by
 * -wfac/(R-L) as dpsion_chiq does once for both branches; then
 * zxq(:,:,it) = i*rcxq(:,:,it) + sum_iw rmatt(it,iw) rcxq(:,:,iw), it = 0..nw
 * (slice it of rcxq is bin it: freq_r(it) is that bin's centre, and slice 0
 * is the unaccumulated zero-frequency slot).                             */
void orc_dpsion_real(int npr, int npr_col, int nwhis, int nw, const double* frhis,
                     const double* rmatt, double ecut, int scaled, double complex* rcxq,
                     double complex* zxq) {
  long slice = (long)npr * npr_col;
  if (!scaled)
    for (int iw = 1; iw <= nwhis; ++iw) {
      double hl = frhis[iw - 1], hr = frhis[iw], hc = 0.5 * (hl + hr);
      double wfac = (ecut < 1e9) ? exp(-(hc / ecut) * (hc / ecut)) : 1.0;
      double s = -wfac / (hr - hl);
      double complex* r = rcxq + slice * iw;
      for (long e = 0; e < slice; ++e) r[e] *= s;
    }
#pragma omp parallel for
  for (long e = 0; e < slice; ++e)
    for (int it = 0; it <= nw; ++it) {
      double complex s = 0.0;
      for (int iw = 1; iw <= nwhis; ++iw) s += rcxq[e + slice * iw] * rmatt[it + (long)(nw + 1) * (iw - 1)];
      zxq[e + slice * it] = I * rcxq[e + slice * it] + s;
    }
}

/* zmel IPW block (S/m_zmel.f90:418-538).  gggmat(X, G) = ggg(igggi(s G - X -
 * nadd)) (0 outside the cube or for index 0); branch 1 (G1 first, taken when
 * nm > ntp0): ggitp = gggmat(X, G1) geigq, zmelp0(:, n, itp) =
 * ggitp(igcgp2i(Gc + G2), itp) dgeigqk(:, n); branch 2: ggit = gggmat(X, G2)
 * dgeigqk with s = -1, zmelp0(:, it, :) = ggit(igcgp2i(Gc - G1), it) geigq;
 * zmel(igc, n, m) = phase(igc) zmelp0.  branch 0 = the reference's choice.
 * Returns -1 when an igcgp2i lookup leaves its cube or finds no entry. */
static long cube_at(const int* v, const int* lo, const int* n) {
  long i0 = v[0] - lo[0], i1 = v[1] - lo[1], i2 = v[2] - lo[2];
  if (i0 < 0 || i0 >= n[0] || i1 < 0 || i1 >= n[1] || i2 < 0 || i2 >= n[2]) return -1;
  return i0 + (long)n[0] * (i1 + (long)n[1] * i2);
}

int orc_zmel_ipw(int ngc, const int* ngveccR, const int* nadd, const double complex* phase, int ngpmx, int ngp1,
                 const int* ngvecpB1, const double complex* geigq, int ntp0, int ngp2, const int* ngvecpB2,
                 const double complex* dgeigqk, int nm, const double complex* ggg, const int* glo, const int* gn,
                 const int* igggi, int ngcgp, const int* nvgcgp2, const int* clo, const int* cn,
                 const int* igcgp2i, int branch, double complex* zmel, long ldz) {
  if (branch == 0) branch = (nm > ntp0) ? 1 : 2;
  const int s = (branch == 1) ? 1 : -1;
  const int ngp = (branch == 1) ? ngp1 : ngp2;
  const int* gp = (branch == 1) ? ngvecpB1 : ngvecpB2;
  const double complex* B = (branch == 1) ? geigq : dgeigqk; /* first contraction */
  const int nB = (branch == 1) ? ntp0 : nm;
  double complex* mat = (double complex*)calloc((size_t)ngcgp * ngp, sizeof(double complex));
  for (int c = 0; c < ngp; ++c)
    for (int r = 0; r < ngcgp; ++r) {
      int v[3];
      for (int d = 0; d < 3; ++d) v[d] = s * gp[d + 3 * c] - nvgcgp2[d + 3 * r] - nadd[d];
      long at = cube_at(v, glo, gn);
      if (at >= 0 && igggi[at] > 0) mat[r + (long)ngcgp * c] = ggg[igggi[at] - 1];
    }
  double complex* gg = (double complex*)malloc(sizeof(double complex) * (size_t)ngcgp * nB);
  orc_zgemm('N', 'N', ngcgp, nB, ngp, 1.0, mat, ngcgp, B, ngpmx, 0.0, gg, ngcgp);
  free(mat);
  /* second contraction over the other plane-wave basis */
  const int ngq = (branch == 1) ? ngp2 : ngp1;
  const int* gq = (branch == 1) ? ngvecpB2 : ngvecpB1;
  const int t = (branch == 1) ? 1 : -1; /* Gc + G2 or Gc - G1 */
  const double complex* C2 = (branch == 1) ? dgeigqk : geigq;
  const int nC = (branch == 1) ? nm : ntp0;
  double complex* work = (double complex*)malloc(sizeof(double complex) * (size_t)ngc * ngq);
  double complex* out = (double complex*)malloc(sizeof(double complex) * (size_t)ngc * nC);
  int rc = 0;
  for (int j = 0; j < nB && rc == 0; ++j) {
    for (int q = 0; q < ngq && rc == 0; ++q)
      for (int c = 0; c < ngc; ++c) {
        int v[3];
        for (int d = 0; d < 3; ++d) v[d] = ngveccR[d + 3 * c] + t * gq[d + 3 * q];
        long at = cube_at(v, clo, cn);
        if (at < 0 || igcgp2i[at] <= 0) {
          rc = -1;
          break;
        }
        work[c + (long)ngc * q] = gg[(igcgp2i[at] - 1) + (long)ngcgp * j];
      }
    if (rc) break;
    orc_zgemm('N', 'N', ngc, nC, ngq, 1.0, work, ngc, C2, ngpmx, 0.0, out, ngc);
    for (int i = 0; i < nC; ++i)
      for (int c = 0; c < ngc; ++c) {
        /* branch 1: j = itp, i = n; branch 2: j = it, i = itp */
        const int n = (branch == 1) ? i : j, m = (branch == 1) ? j : i;
        zmel[c + ldz * (n + (long)nm * m)] = phase[c] * out[c + (long)ngc * i];
      }
  }
  free(gg);
  free(work);
  free(out);
  return rc;
}

/* ---------------------------------------------------------------------- */
/* zmel (S/m_zmel.f90:146-597).  MT valence-valence block per atom
 * (:323-371):  ppbvphiq(:,:,i) = ppb(:,:,i)^T cphiq  (:361-362),
 *              zmelt_d(:,:,i)  = phase * cphim^H ppbvphiq(:,:,i) (:365-366),
 *              zmelt(ims+i-1,:,:) = zmelt_d(:,:,i) (:367-371);
 * then M->E (:559-560) and conj (:590-594).                              */
int orc_zmel(int natom, const int* nv, const int* mdim, const double complex* phase,
             const double* ppb, int ndima, int nm, const double complex* cphim, int nq,
             const double complex* cphiq, const double complex* m2e, int nbb, int conjg,
             double complex* zmel) {
  int ngb = 0, nvsum = 0;
  for (int a = 0; a < natom; ++a) {
    ngb += mdim[a];
    nvsum += nv[a];
  }
  if (nvsum > ndima) return 1;
  double complex* zmelt = (double complex*)calloc((size_t)ngb * nm * nq, sizeof(double complex));
  long ias = 0, ims = 0, poff = 0;
  for (int a = 0; a < natom; ++a) {
    int v = nv[a], md = mdim[a];
    const double* pa = ppb + poff;
    /* the reference's two batched GEMMs over i (:353-360):
     *   ppbvphiq(b, m, i) = sum_c ppb(c, b, i) cphiq(ias+c, m)              (op T)
     *   zmelt(ims+i, n, m) = phase * sum_b conj(cphim(ias+b, n)) ppbvphiq(b, m, i)  (op C) */
    double complex* pc = (double complex*)malloc(sizeof(double complex) * (size_t)v * v);
    double complex* pv = (double complex*)malloc(sizeof(double complex) * (size_t)v * (nq > 0 ? nq : 1));
    double complex* zt = (double complex*)malloc(sizeof(double complex) * (size_t)(nm > 0 ? nm : 1) * (nq > 0 ? nq : 1));
    for (int i = 0; i < md; ++i) {
      for (long e = 0; e < (long)v * v; ++e) pc[e] = pa[e + (long)v * v * i];
      orc_zgemm('T', 'N', v, nq, v, 1.0, pc, v, cphiq + ias, ndima, 0.0, pv, v);
      orc_zgemm('C', 'N', nm, nq, v, phase[a], cphim + ias, ndima, pv, v, 0.0, zt, nm);
      for (int m = 0; m < nq; ++m)
        for (int n = 0; n < nm; ++n) zmelt[ims + i + (long)ngb * (n + (long)nm * m)] = zt[n + (long)nm * m];
    }
    free(pc);
    free(pv);
    free(zt);
    ias += v;
    ims += md;
    poff += (long)v * v * md;
  }
  long nrow = m2e ? nbb : ngb;
  if (m2e) {
    /* zmel = m2e^H zmelt  (nbb x ngb) (ngb x nm*nq) */
    orc_zgemm('C', 'N', nbb, nm * nq, ngb, 1.0, m2e, ngb, zmelt, ngb, 0.0, zmel, nbb);
  } else {
    memcpy(zmel, zmelt, sizeof(double complex) * (size_t)ngb * nm * nq);
  }
  if (conjg)
    for (long i = 0; i < nrow * nm * nq; ++i) zmel[i] = conj(zmel[i]);
  free(zmelt);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* x0gemm (S/x0gemm.f90:2-97): bin lists (:36-64) then per (jpm, iw)
 * rcxq(:,:,iw*(3-2jpm)) += zw^H wzw with zw(t,I)=zmel(I,it,itp),
 * wzw(t,J)=zmel(J+ipr_col-1,it,itp)*whw(t) (:72-90).                    */
void orc_x0gemm(double complex* rcxq, int npr, int ipr_col, int npr_col, int nwhis, int npm,
                int ns1, int ns2, const double complex* zmel, int nbb, int nqtot,
                int icoun_lo, int icoun_hi, const int* iwini, const int* iwend, const int* itc,
                const int* itpc, const int* jpmc, const int* icouini, const double* whwc) {
  int nm = ns2 - ns1 + 1;
  long slice = (long)npr * npr_col;
  int s0 = (1 - npm) * nwhis; /* Fortran index of slice 0 */
  (void)nqtot;
  /* count pairs per (iw, jpm) and build lists in icoun order */
  int nb = nwhis * npm;
  int* cnt = (int*)calloc(nb, sizeof(int));
  for (int ic = icoun_lo; ic <= icoun_hi; ++ic) {
    int it = itc[ic - 1];
    if (it < ns1 || it > ns2) continue;
    for (int iw = iwini[ic - 1]; iw <= iwend[ic - 1]; ++iw) cnt[(iw - 1) + nwhis * (jpmc[ic - 1] - 1)]++;
  }
  long* start = (long*)malloc(sizeof(long) * (nb + 1));
  start[0] = 0;
  for (int b = 0; b < nb; ++b) start[b + 1] = start[b] + cnt[b];
  long tot = start[nb];
  long* col = (long*)malloc(sizeof(long) * (tot + 1));
  double* wt = (double*)malloc(sizeof(double) * (tot + 1));
  memset(cnt, 0, sizeof(int) * nb);
  for (int ic = icoun_lo; ic <= icoun_hi; ++ic) {
    int it = itc[ic - 1], itp = itpc[ic - 1];
    if (it < ns1 || it > ns2) continue;
    for (int iw = iwini[ic - 1]; iw <= iwend[ic - 1]; ++iw) {
      int b = (iw - 1) + nwhis * (jpmc[ic - 1] - 1);
      long p = start[b] + cnt[b]++;
      col[p] = (long)nbb * ((it - ns1) + (long)nm * (itp - 1));
      wt[p] = whwc[iw - iwini[ic - 1] + icouini[ic - 1] - 1];
    }
  }
  /* per bin: zw(t,I) and wzw(t,J) gathered, rcxq(:,:,iw) += zw^H wzw (:72-90) */
  long nmax = 0;
  for (int b = 0; b < nb; ++b)
    if (start[b + 1] - start[b] > nmax) nmax = start[b + 1] - start[b];
  double complex* zw = (double complex*)malloc(sizeof(double complex) * (size_t)(nmax > 0 ? nmax : 1) * npr);
  double complex* wzw = (double complex*)malloc(sizeof(double complex) * (size_t)(nmax > 0 ? nmax : 1) * npr_col);
  for (int jpm = 1; jpm <= npm; ++jpm)
    for (int iw = 1; iw <= nwhis; ++iw) {
      int b = (iw - 1) + nwhis * (jpm - 1);
      long nt = start[b + 1] - start[b];
      if (nt == 0) continue;
      int fidx = iw * (3 - 2 * jpm);
      double complex* r = rcxq + slice * (fidx - s0);
      for (long t = 0; t < nt; ++t) {
        const double complex* z = zmel + col[start[b] + t];
        double w = wt[start[b] + t];
        for (int i = 0; i < npr; ++i) zw[t + nt * i] = z[i];
        for (int j = 0; j < npr_col; ++j) wzw[t + nt * j] = z[j + ipr_col - 1] * w;
      }
      orc_zgemm('C', 'N', npr, npr_col, (int)nt, 1.0, zw, nt, wzw, nt, 1.0, r, npr);
    }
  free(zw);
  free(wzw);
  free(cnt);
  free(start);
  free(col);
  free(wt);
}

/* dpsion_chiq imaginary axis, npm=1 (S/dpsion5.f90:211-232). */
void orc_dpsion_imag(int npr, int npr_col, int nwhis, int niw, const double* frhis,
                     const double* imatt, double ecut, double complex* rcxq, double complex* zxqi) {
  long slice = (long)npr * npr_col;
  for (int iw = 1; iw <= nwhis; ++iw) {
    double hl = frhis[iw - 1], hr = frhis[iw], hc = 0.5 * (hl + hr);
    double wfac = (ecut < 1e9) ? exp(-(hc / ecut) * (hc / ecut)) : 1.0;
    double s = -wfac / (hr - hl);
    double complex* r = rcxq + slice * iw;
    for (long e = 0; e < slice; ++e) r[e] *= s;
  }
  /* zxqi = rcxq(:, 1:nwhis) imatt^T, the reference's GEMM (:229) */
  double complex* im = (double complex*)malloc(sizeof(double complex) * (size_t)niw * nwhis);
  for (long e = 0; e < (long)niw * nwhis; ++e) im[e] = imatt[e];
  orc_zgemm('N', 'T', (int)slice, niw, nwhis, 1.0, rcxq + slice, slice, im, niw, 0.0, zxqi, slice);
  free(im);
}

/* Composed direct-frequency chi0 (associativity of the x0gemm bin sum and
 * the dpsion_chiq GEMM, S/x0gemm.f90:88-89 with S/dpsion5.f90:229).      */
void orc_chi0_freq(int ni, int npairs, int nf, const double complex* z, const double* w,
                   double complex* chi) {
  long slice = (long)ni * ni;
#pragma omp parallel for
  for (int j = 0; j < ni; ++j)
    for (int v = 0; v < nf; ++v) {
      double complex* c = chi + slice * v + (long)ni * j;
      for (int i = 0; i < ni; ++i) c[i] = 0.0;
      for (int p = 0; p < npairs; ++p) {
        const double complex* zp = z + (long)ni * p;
        double complex wz = w[v + (long)nf * p] * zp[j];
        for (int i = 0; i < ni; ++i) c[i] += conj(zp[i]) * wz;
      }
    }
}

/* ---------------------------------------------------------------------- */
/* zminv_h (S/m_lapack.f90:21-47): zgetrf + zgetri.  Restated as LU with
 * partial pivoting followed by inversion; same inverse up to rounding.   */
int orc_zinv(int n, double complex* a, long lda) {
  if (n <= 0) return 0;
  if (lapack_zgetrf && lapack_zgetri && fits_int(lda)) {
    int in = n, ilda = (int)lda, info = 0, lwork = 64 * n;
    int* ipiv = (int*)malloc(sizeof(int) * n);
    double complex* work = (double complex*)malloc(sizeof(double complex) * (size_t)lwork);
    lapack_zgetrf(&in, &in, a, &ilda, ipiv, &info);
    if (info == 0) lapack_zgetri(&in, a, &ilda, ipiv, work, &lwork, &info);
    free(ipiv);
    free(work);
    return info == 0 ? 0 : 2;
  }
  int* piv = (int*)malloc(sizeof(int) * n);
  /* LU factorisation, right-looking, partial pivoting */
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = cabs(a[k + lda * k]);
    for (int i = k + 1; i < n; ++i) {
      double v = cabs(a[i + lda * k]);
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (best == 0.0) { free(piv); return 2; }
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double complex t = a[k + lda * j];
        a[k + lda * j] = a[p + lda * j];
        a[p + lda * j] = t;
      }
    double complex inv = 1.0 / a[k + lda * k];
    for (int i = k + 1; i < n; ++i) a[i + lda * k] *= inv;
#pragma omp parallel for
    for (int j = k + 1; j < n; ++j) {
      double complex akj = a[k + lda * j];
      if (akj == 0.0) continue;
      for (int i = k + 1; i < n; ++i) a[i + lda * j] -= a[i + lda * k] * akj;
    }
  }
  /* inv(U) in place (upper triangle) */
  for (int j = 0; j < n; ++j) {
    a[j + lda * j] = 1.0 / a[j + lda * j];
    double complex ajj = -a[j + lda * j];
    for (int i = 0; i < j; ++i) {
      double complex s = 0.0;
      for (int l = i; l < j; ++l) s += a[i + lda * l] * a[l + lda * j];
      a[i + lda * j] = s;
    }
    for (int i = 0; i < j; ++i) a[i + lda * j] *= ajj;
  }
  /* solve inv(A) * L = inv(U) for inv(A) (getri), column by column from the right */
  double complex* work = (double complex*)malloc(sizeof(double complex) * n);
  for (int j = n - 1; j >= 0; --j) {
    for (int i = j + 1; i < n; ++i) {
      work[i] = a[i + lda * j];
      a[i + lda * j] = 0.0;
    }
    for (int l = j + 1; l < n; ++l) {
      double complex wl = work[l];
      if (wl == 0.0) continue;
      for (int i = 0; i < n; ++i) a[i + lda * j] -= a[i + lda * l] * wl;
    }
  }
  free(work);
  /* undo the row interchanges as column interchanges */
  for (int k = n - 1; k >= 0; --k) {
    int p = piv[k];
    if (p != k)
      for (int i = 0; i < n; ++i) {
        double complex t = a[i + lda * k];
        a[i + lda * k] = a[i + lda * p];
        a[i + lda * p] = t;
      }
  }
  free(piv);
  return 0;
}

/* WVIllwI body, q /= 0 (S/m_llw.f90:299-373). */
int orc_wc_build(int ngb, int nw, int nspin, int is_x0_m_basis, int is_wc_m_basis,
                 const double complex* chi, const double* vcousq, const double complex* m2e,
                 double complex* wc, int ldw) {
  long n2 = (long)ngb * ngb;
  double complex* x = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* t = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* eps = (double complex*)malloc(sizeof(double complex) * n2);
  int rc = 0;
  for (int iw = 0; iw < nw; ++iw) {
    const double complex* c = chi + n2 * iw;
    double fac = (nspin == 1) ? 2.0 : 1.0; /* :299-303 */
    for (long e = 0; e < n2; ++e) x[e] = fac * c[e];
    if (is_x0_m_basis) { /* :336-337 */
      orc_zgemm('N', 'N', ngb, ngb, ngb, 1.0, x, ngb, m2e, ngb, 0.0, t, ngb);
      orc_zgemm('C', 'N', ngb, ngb, ngb, 1.0, m2e, ngb, t, ngb, 0.0, x, ngb);
    }
    for (int j = 0; j < ngb; ++j) /* :341-348 */
      for (int i = 0; i < ngb; ++i) {
        double complex e = -vcousq[i] * x[i + (long)ngb * j] * vcousq[j];
        if (i == j) e = 1.0 + e;
        eps[i + (long)ngb * j] = e;
      }
    if (orc_zinv(ngb, eps, ngb) != 0) rc = 2; /* :350 */
    double complex* w = wc + (long)ldw * ngb * iw;
    for (int j = 0; j < ngb; ++j) /* :353-360 */
      for (int i = 0; i < ngb; ++i) {
        double complex e = vcousq[i] * eps[i + (long)ngb * j] * vcousq[j];
        if (i == j) e -= vcousq[i] * vcousq[j];
        t[i + (long)ngb * j] = e;
      }
    if (is_wc_m_basis) { /* :366-367 */
      orc_zgemm('N', 'C', ngb, ngb, ngb, 1.0, t, ngb, m2e, ngb, 0.0, x, ngb);
      orc_zgemm('N', 'N', ngb, ngb, ngb, 1.0, m2e, ngb, x, ngb, 0.0, t, ngb);
    }
    for (int j = 0; j < ngb; ++j)
      for (int i = 0; i < ngb; ++i) w[i + (long)ldw * j] = t[i + (long)ngb * j];
  }
  free(x);
  free(t);
  free(eps);
  return rc;
}

/* ---------------------------------------------------------------------- */
/* Imaginary-axis weights (S/m_sxcf_gemm.f90:432-458). */
void orc_wgtim(int niw, int ns1, int ns2, int ntq, int nctot, const double* omega,
               const double* ekc, double esmr, double ua, double wkkr, const double* x,
               const double* wt, const double* expa, double* wgtim) {
  const double rmax = 2.0;
  int nm = ns2 - ns1 + 1;
  double sig = 0.5 * esmr;
  double sig2 = 2.0 * sig * sig;
  double* w_ = (double*)malloc(sizeof(double) * (niw + 1));
  for (int itp = 0; itp < ntq; ++itp)
    for (int it = ns1; it <= ns2; ++it) {
      double we = 0.5 * (omega[itp] - ekc[it - ns1]);
      double aw = fabs(ua * we);
      double aw2 = aw * aw;
      double sgn = signbit(we) ? -1.0 : 1.0;
      if (it <= nctot) {
        double s = 0.0;
        for (int i = 0; i < niw; ++i) {
          double cons = 1.0 / (we * we * x[i] * x[i] + (1.0 - x[i]) * (1.0 - x[i]));
          w_[1 + i] = we * cons * wt[i] * (-1.0 / PI);
          s += w_[1 + i] * expa[i];
        }
        w_[0] = (fabs(we) < rmax / ua) ? (-s - 0.5 * sgn * exp(we * we * ua * ua) * erfc(ua * fabs(we))) : 0.0;
      } else {
        double s = 0.0;
        for (int i = 0; i < niw; ++i) {
          double omd = 1.0 / x[i] - 1.0;
          double omd2w = omd * omd + we * we;
          double r = omd2w / sig2, cons;
          if (r > 5e-3)
            cons = (1.0 - exp(-r)) / omd2w;
          else
            cons = (1.0 / sig2 - omd2w / (sig2 * sig2) / 2.0 + omd2w * omd2w / (sig2 * sig2 * sig2) / 6.0 -
                    omd2w * omd2w * omd2w / pow(sig2, 4) / 24.0 + pow(omd2w, 4) / pow(sig2, 5) / 120.0 -
                    pow(omd2w, 5) / pow(sig2, 6) / 720.0);
          w_[1 + i] = -we * cons * wt[i] / (x[i] * x[i]) / PI;
          s += cons * expa[i] * wt[i] / (x[i] * x[i]);
        }
        w_[0] = we * s / PI + sgn * 0.5 * exp(aw2) * (erfc(sqrt(aw2 + we * we / sig2)) - erfc(aw));
      }
      double* out = wgtim + (long)(niw + 1) * ((it - ns1) + (long)nm * itp);
      for (int i = 0; i <= niw; ++i) out[i] = wkkr * w_[i];
    }
  free(w_);
}

/* Imaginary-axis Sigma_c + final GEMM + clip (S/m_sxcf_gemm.f90:463-496, 599-608). */
/* czmelwc of the imaginary axis (:463-496): cz = sum_iw Wt(iw)^H (wgtim(iw) o zmel) */
static void sigc_imag_cz(int ngb, int nm, int ntq, int niw, const double complex* zmel, const double complex* wc,
                         int ldw, const double* wgtim, double complex* cz) {
  long nc = (long)nm * ntq;
  long n2 = (long)ngb * ngb;
  double complex* w = (double complex*)malloc(sizeof(double complex) * n2);
  double complex* wz = (double complex*)malloc(sizeof(double complex) * ngb * nc);
  for (int iw = 0; iw <= niw; ++iw) {
    const double complex* src = wc + (long)ldw * ngb * iw;
    /* keepwv unpack: wc(i,j)=upper(i,j), wc(j,i)=conj(upper(i,j)) for i<=j (:468-475) */
    for (int j = 0; j < ngb; ++j)
      for (int i = 0; i <= j; ++i) {
        double complex u = src[i + (long)ldw * j];
        w[i + (long)ngb * j] = u;
        w[j + (long)ngb * i] = conj(u);
      }
    /* wzmel = wgtim(iw,it,itp) * zmel (:482-488) */
    for (long c = 0; c < nc; ++c) {
      double g = wgtim[iw + (long)(niw + 1) * c];
      for (int i = 0; i < ngb; ++i) wz[i + (long)ngb * c] = g * zmel[i + (long)ngb * c];
    }
    /* czmelwc (+)= wc^H wzmel (:492) */
    orc_zgemm('C', 'N', ngb, (int)nc, ngb, 1.0, w, ngb, wz, ngb, iw == 0 ? 0.0 : 1.0, cz, ngb);
  }
  free(w);
  free(wz);
}

/* zsec += czmelwc^H zmel, K = ngb*nm (:599); Im(zsec_nn) <= 0 (:601-608) */
static void sigc_final(int ngb, int nm, int ntq, const double complex* cz, const double complex* zmel,
                       double complex* zsec, int ldz) {
  orc_zgemm('C', 'N', ntq, ntq, ngb * nm, 1.0, cz, (long)ngb * nm, zmel, (long)ngb * nm, 1.0, zsec, ldz);
  for (int i = 0; i < ntq; ++i) {
    double complex z = zsec[i + (long)ldz * i];
    double zi = cimag(z);
    if (zi > 0.0) zi = 0.0;
    zsec[i + (long)ldz * i] = creal(z) + I * zi;
  }
}

void orc_sigc_imag(int ngb, int nm, int ntq, int niw, const double complex* zmel,
                   const double complex* wc, int ldw, const double* wgtim,
                   double complex* zsec, int ldz) {
  double complex* cz = (double complex*)malloc(sizeof(double complex) * ngb * (long)nm * ntq);
  sigc_imag_cz(ngb, nm, ntq, niw, zmel, wc, ldw, wgtim, cz);
  sigc_final(ngb, nm, ntq, cz, zmel, zsec, ldz);
  free(cz);
}

/* Real-axis pole term of Sigma_c for one internal k (S/m_sxcf_gemm.f90:501-594).
 * Entry lists as the reference builds them (:512-552): for each external
 * state itp and middle state it in the window [itini, itend] (:514-515), the
 * pole weight wfac = wfacx2(omg, ef, ekc(it), smearing) * wkkr * sign(omg-ef)
 * (skipped below 1e-8), the frequency we = |omg - weavx2(...)|/2 (Hartree),
 * the first mesh point ixs with freq_r(ixs) > we, and the 3-point weights
 * wfac * [1, we^2, we^4] * inverse([1, xi^2, xi^4]) of xi = freq_r(ixs-1..ixs+1)
 * (:543-548).  Then for every real frequency iw in increasing order, the
 * entries' columns get czmelwc(:,it,itp) += wgt * Wt(iw)^H zmel(:,it,itp)
 * (:556-590) with the keepwv Hermitian packing of W (:560-566), and
 * zsec += czmelwc^H zmel (:599), then Im(zsec_nn) <= 0.
 * ekc(1..nm): middle-state energies (Ry), the first nctot are core; omega:
 * ntq external energies (Ry); ef, esmr in Ry; freq_r(0..nw) in Hartree;
 * wr: ldw x ngb x (nw+1).  Returns the number of entries, or -1 if a pole
 * frequency falls outside freq_r(1..nw-1). */
static void inv33(const double a[3][3], double r[3][3]) {
  /* adjugate / determinant (a[i][j] = row i, column j) */
  double c[3][3];
  c[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  c[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
  c[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
  c[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  c[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
  c[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
  c[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  c[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
  c[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
  double det = a[0][0] * c[0][0] + a[0][1] * c[1][0] + a[0][2] * c[2][0];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = c[i][j] / det;
}

static int sigc_real_cz(int ngb, int nm, int ntq, const double complex* zmel, const double complex* wr, int ldw,
                        int nw, const double* freq_r, const double* omega, const double* ekc, double ef,
                        double esmr, int nctot, int ns2r, double wkkr, double complex* cz) {
  const double ddw = 10.0, wfaccut = 1e-8;
  int nt0p = 0, nt0m = 0;
  for (int i = 0; i < nm; ++i) {
    if (ekc[i] < ef + ddw * esmr) nt0p++;
    if (ekc[i] < ef - ddw * esmr) nt0m++;
  }
  const long cap = 3L * nm * ntq + 3;
  int* e_iw = (int*)malloc(sizeof(int) * cap);
  int* e_col = (int*)malloc(sizeof(int) * cap);
  double* e_w = (double*)malloc(sizeof(double) * cap);
  long ne = 0;
  for (int itp = 0; itp < ntq; ++itp) {
    const double omg = omega[itp];
    const int itini = (omg >= ef) ? (nt0m + 1 > 1 ? nt0m + 1 : 1) : 1;
    const int itend = (omg >= ef) ? ns2r : (nt0p < ns2r ? nt0p : ns2r);
    for (int it = itini; it <= itend; ++it) {
      double wfac = orc_wfacx2(omg, ef, ekc[it - 1], it <= nctot ? 0.0 : esmr);
      if (wfac < wfaccut) continue;
      wfac *= wkkr * ((omg - ef) >= 0.0 ? 1.0 : -1.0);
      const double we = 0.5 * fabs(omg - orc_weavx2(omg, ef, ekc[it - 1], esmr));
      int ixs = 0;
      for (int i = 1; i <= nw; ++i)
        if (freq_r[i] > we) {
          ixs = i;
          break;
        }
      if (ixs < 1 || ixs + 1 > nw) {
        free(e_iw);
        free(e_col);
        free(e_w);
        return -1;
      }
      double a[3][3], ai[3][3];
      for (int i = 0; i < 3; ++i) {
        const double xi = freq_r[ixs - 1 + i];
        a[i][0] = 1.0;
        a[i][1] = xi * xi;
        a[i][2] = xi * xi * xi * xi;
      }
      inv33(a, ai);
      const double xv[3] = {1.0, we * we, we * we * we * we};
      for (int j = 0; j < 3; ++j) {
        double g = 0.0;
        for (int i = 0; i < 3; ++i) g += xv[i] * ai[i][j];
        e_iw[ne] = ixs - 1 + j;
        e_col[ne] = (it - 1) + nm * itp;
        e_w[ne] = wfac * g;
        ne++;
      }
    }
  }
  const long n2 = (long)ngb * ngb;
  double complex* w = (double complex*)malloc(sizeof(double complex) * n2);
  for (int iw = 0; iw <= nw; ++iw) {
    int any = 0;
    for (long e = 0; e < ne; ++e) any |= (e_iw[e] == iw);
    if (!any) continue;
    const double complex* src = wr + (long)ldw * ngb * iw;
    for (int j = 0; j < ngb; ++j)
      for (int i = 0; i <= j; ++i) {
        double complex u = src[i + (long)ldw * j];
        w[i + (long)ngb * j] = u;
        w[j + (long)ngb * i] = conj(u);
      }
    for (long e = 0; e < ne; ++e) {
      if (e_iw[e] != iw) continue;
      const double complex* z = zmel + (long)ngb * e_col[e];
      double complex* c = cz + (long)ngb * e_col[e];
      for (int i = 0; i < ngb; ++i) {
        double complex s = 0.0;
        for (int k = 0; k < ngb; ++k) s += conj(w[k + (long)ngb * i]) * (e_w[e] * z[k]);
        c[i] += s;
      }
    }
  }
  free(w);
  free(e_iw);
  free(e_col);
  free(e_w);
  return (int)ne;
}

int orc_sigc_real(int ngb, int nm, int ntq, const double complex* zmel, const double complex* wr, int ldw, int nw,
                  const double* freq_r, const double* omega, const double* ekc, double ef, double esmr, int nctot,
                  int ns2r, double wkkr, double complex* zsec, int ldz) {
  double complex* cz = (double complex*)calloc((size_t)ngb * nm * ntq, sizeof(double complex));
  const int ne = sigc_real_cz(ngb, nm, ntq, zmel, wr, ldw, nw, freq_r, omega, ekc, ef, esmr, nctot, ns2r, wkkr, cz);
  if (ne >= 0) sigc_final(ngb, nm, ntq, cz, zmel, zsec, ldz);
  free(cz);
  return ne;
}

/* get_correlation_block (S/m_sxcf_gemm.f90:423-608): imaginary-axis and
 * real-axis parts in one czmelwc, one final product, then the clip. */
int orc_sigc_full(int ngb, int nm, int ntq, int niw, const double complex* zmel, const double complex* wc, int ldw,
                  const double* wgtim, const double complex* wr, int ldwr, int nw, const double* freq_r,
                  const double* omega, const double* ekc, double ef, double esmr, int nctot, int ns2r, double wkkr,
                  double complex* zsec, int ldz) {
  double complex* cz = (double complex*)malloc(sizeof(double complex) * ngb * (long)nm * ntq);
  sigc_imag_cz(ngb, nm, ntq, niw, zmel, wc, ldw, wgtim, cz);
  const int ne = sigc_real_cz(ngb, nm, ntq, zmel, wr, ldwr, nw, freq_r, omega, ekc, ef, esmr, nctot, ns2r, wkkr, cz);
  if (ne >= 0) sigc_final(ngb, nm, ntq, cz, zmel, zsec, ldz);
  free(cz);
  return ne;
}

/* wfacx2 / weavx2 (S/wfacx.f90:19-49). */
double orc_wfacx2(double e1, double e2, double ek, double esmr) {
  const double ewidthcut = 1e-6;
  double el = fmin(e1, e2), eh = fmax(e1, e2);
  if (eh - el < ewidthcut) return 0.0;
  if (esmr == 0.0) return (el <= ek && ek < eh) ? 1.0 : 0.0;
  return 0.5 * erfc(-(eh - ek) / sqrt(2.0) / esmr) - 0.5 * erfc(-(el - ek) / sqrt(2.0) / esmr);
}

double orc_weavx2(double e1, double e2, double ek, double esmr) {
  const double pi = 3.1415926535897932;
  double el = fmin(e1, e2), eh = fmax(e1, e2);
  double wtt = 0.5 * erfc(-(eh - ek) / sqrt(2.0) / esmr) - 0.5 * erfc(-(el - ek) / sqrt(2.0) / esmr);
  double sig2 = 2.0 * esmr * esmr;
  return ek + esmr / sqrt(2.0 * pi) * (-exp(-(eh - ek) * (eh - ek) / sig2) + exp(-(el - ek) * (el - ek) / sig2)) / wtt;
}

/* ---------------------------------------------------------------------- */
/* QP energy (S/main_hqpe.sc.f90:152-161); hartree = 2*rydberg(). */
void orc_qp_energy(int ntq, const double* elda, const double* sex, const double* sexcore,
                   const double complex* sec, int ldsec, const double* vxc, double* eqp) {
  const double hartree = 2.0 * 13.605693122994;
  for (int i = 0; i < ntq; ++i) {
    double rsec = creal(sec[i + (long)ldsec * i]) * hartree;
    double eshift = sex[i] * hartree + sexcore[i] * hartree + rsec - vxc[i];
    eqp[i] = elda[i] + eshift;
  }
}

/* Hermitised Sigma - Vxc (S/main_hqpe.sc.f90:175-181). */
void orc_sigma_herm(int ntq, const double complex* sex2, const double complex* sexcore2,
                    const double complex* sec, const double complex* vxcmat, double complex* se) {
  for (int j = 0; j < ntq; ++j)
    for (int i = 0; i < ntq; ++i) {
      long ij = i + (long)ntq * j, ji = j + (long)ntq * i;
      se[ij] = sex2[ij] + sexcore2[ij] + 0.5 * (sec[ij] + conj(sec[ji])) - vxcmat[ij];
    }
}

/* ---------------------------------------------------------------------- */
/* get_zmel_init_gemm's muffin-tin part with core states and the rotated-atom
 * map (S/m_zmel.f90:199-214, 295-305, 323-391), then M->E and conj
 * (:559-560, 590-594).  ppb is the reference's array ppb(nlnmx, nlnmx,
 * mdimx, natom): for atom a, functions 1..nlnmc(a) are core, then
 * nlnmv(a) valence.  Atom ia uses its own eigenvector rows (iasx: the
 * valence counts before it), its own core states (icsx: the core counts
 * before it), and the product basis, ppb block and row offset of atom
 * iap = iatomp(ia) (imdim(iap): the nblocha before iap).  Middle states
 * ns1..ns2 run over core states 1..nctot, then valence bands (band nctot+j
 * is column j - (nm1v - nctot) of cphim, which holds bands nm1v..nm2v).
 *   valence (:349-363): zmelt(ims+i, it, m) = phase(ia) sum_{b,c}
 *       conj(cphim(ias+b, it)) ppb(nc+c, nc+b, i, iap) cphiq(ias+c, m)
 *   core (:374-390):    zmelt(ims+i, it, m) = phase(ia) sum_b
 *       ppb(nc+b, it-ics, i, iap) cphiq(ias+b, m), it in
 *       [max(nm1c, ics+1), min(nm2c, ics+ncorec)], ncorec = nlnmc(ia) if nctot > 0.
 * zmel out: (m2e ? nbb : ngb) x (ns2-ns1+1) x nq.                          */
int orc_zmel_init(int natom, int nlnmx, int mdimx, const int* nlnmc, const int* nlnmv, const int* nblocha,
                  const int* iatomp, const double complex* phase, const double* ppb, int ndima, int ns1, int ns2,
                  int nctot, const double complex* cphim, int nq, const double complex* cphiq,
                  const double complex* m2e, int nbb, int conjg, double complex* zmel) {
  int ngb = 0;
  for (int a = 0; a < natom; ++a) ngb += nblocha[a];
  const int nm1 = ns1, nm2 = ns2, nmt = nm2 - nm1 + 1;
  int nm1c, nm2c, nm1v, nm2v; /* S/m_zmel.f90:199-214 */
  if (nm1 > nctot) { nm1c = 0; nm2c = -1; nm1v = nm1; nm2v = nm2; }
  else if (nm2 > nctot) { nm1c = nm1; nm2c = nctot; nm1v = nctot + 1; nm2v = nm2; }
  else { nm1c = nm1; nm2c = nm2; nm1v = 0; nm2v = -1; }
  const int nmv = nm2v >= nm1v ? nm2v - nm1v + 1 : 0;
  double complex* zmelt = (double complex*)calloc((size_t)ngb * nmt * (nq > 0 ? nq : 1), sizeof(double complex));
  const long pstride1 = nlnmx, pstride2 = (long)nlnmx * nlnmx, pstride3 = (long)nlnmx * nlnmx * mdimx;
  int ias = 1, ics = 0;
  for (int ia = 0; ia < natom; ++ia) {
    const int iap = iatomp ? iatomp[ia] - 1 : ia;
    if (iap < 0 || iap >= natom) { free(zmelt); return 2; }
    const int nc = nlnmc[ia], nv = nlnmv[ia], mdim = nblocha[iap];
    int ims = 0;
    for (int a = 0; a < iap; ++a) ims += nblocha[a];
    const double* P = ppb + pstride3 * iap;
    const int ncorec = nctot > 0 ? nc : 0;
    for (int i = 0; i < mdim; ++i)
      for (int m = 0; m < nq; ++m) {
        /* valence middle states */
        for (int j = 0; j < nmv; ++j) {
          double complex s = 0.0;
          for (int b = 0; b < nv; ++b) {
            double complex t = 0.0;
            for (int c = 0; c < nv; ++c)
              t += P[(nc + c) + pstride1 * (nc + b) + pstride2 * i] * cphiq[(ias - 1 + c) + (long)ndima * m];
            s += conj(cphim[(ias - 1 + b) + (long)ndima * j]) * t;
          }
          const int col = (nm1v + j) - nm1;
          zmelt[ims + i + (long)ngb * (col + (long)nmt * m)] = phase[ia] * s;
        }
        /* core middle states of this atom */
        const int lo = nm1c > ics + 1 ? nm1c : ics + 1, hi = nm2c < ics + ncorec ? nm2c : ics + ncorec;
        for (int it = lo; it <= hi; ++it) {
          double complex s = 0.0;
          for (int b = 0; b < nv; ++b)
            s += P[(nc + b) + pstride1 * (it - ics - 1) + pstride2 * i] * cphiq[(ias - 1 + b) + (long)ndima * m];
          zmelt[ims + i + (long)ngb * ((it - nm1) + (long)nmt * m)] = phase[ia] * s;
        }
      }
    ias += nv;
    ics += nc;
  }
  long nrow = m2e ? nbb : ngb;
  if (m2e)
    orc_zgemm('C', 'N', nbb, nmt * nq, ngb, 1.0, m2e, ngb, zmelt, ngb, 0.0, zmel, nbb);
  else
    memcpy(zmel, zmelt, sizeof(double complex) * (size_t)ngb * nmt * nq);
  if (conjg)
    for (long e = 0; e < nrow * nmt * nq; ++e) zmel[e] = conj(zmel[e]);
  free(zmelt);
  return 0;
}
