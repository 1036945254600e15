// gamma.hip -- the Gamma point of the W build:
//   * the q = 0 body (iq = 1) of WVIllwI / WVRllwR: epsilon restricted to
//     mu, nu >= 2 is inverted, W(1,:) = W(:,1) = 0, W stays in the E basis
//     (S/m_llw.f90:98-178, 299-373 with ix = 1);
//   * the offset points q0i (iq > nqibz): epsilon with the head
//     1 - vcou1 chi(1,1), full inverse, llw = 1/epsinv(1,1) and the wing
//     epsinv(1,2:)/epsinv(1,1) of the static inverse (S/m_llw.f90:183-256,
//     378-440);
//   * getw0, the Gamma-cell average of W - v from llw (S/m_w0w0i.f90:146-311),
//     host FP64;
//   * modifyWV0 on device-resident records: W(1,1,iw) = w0(iw), then E -> M
//     (S/m_w0w0i.f90:89-140).
// The inversions are the batched Gauss-Jordan of kernels.hip (partial
// pivoting on the real axis, where epsilon is not Hermitian); the M <-> E
// transforms are those of the W build in the handle's form (wbuild.hip).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {
namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ inline cd cdiv(cd a, cd b) {
  const double d = b.re * b.re + b.im * b.im;
  return cd{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// dst(i, j, iw) = src(i+1, j+1, iw): the mu, nu >= 2 block, m = n-1
__global__ void block_tail_kernel(const cd* __restrict__ src, int n, int nw, cd* __restrict__ dst) {
  const int m = n - 1;
  const int64_t mm = (int64_t)m * m, tot = mm * nw, nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int iw = (int)(e / mm);
    const int64_t ij = e % mm;
    const int i = (int)(ij % m), j = (int)(ij / m);
    dst[e] = src[nn * iw + (i + 1) + (int64_t)n * (j + 1)];
  }
}

// W(1, :) = W(:, 1) = 0 in each record
__global__ void zero_head_kernel(cd* __restrict__ w, int n, int64_t ldw, int64_t wstride, int nw) {
  const int64_t per = 2 * (int64_t)n - 1, tot = per * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int iw = (int)(e / per);
    const int k = (int)(e % per);
    cd* r = w + wstride * iw;
    if (k < n) r[k] = cd{0.0, 0.0};                   // column 1
    else r[ldw * (k - n + 1)] = cd{0.0, 0.0};          // row 1, columns 2..n
  }
}

// llw(iw) = 1/inv(1,1); wing(j-1) = inv(1,j)/inv(1,1) at wing_iw.  Column j of
// the inverse is stored column perm[j + n*iw] (pivoted) or j.
__global__ void head_kernel(const cd* __restrict__ inv, const int* __restrict__ perm, int n, int nw,
                            cd* __restrict__ llw, int wing_iw, cd* __restrict__ wing) {
  const int64_t nn = (int64_t)n * n;
  const int nwing = (wing && wing_iw >= 0 && wing_iw < nw) ? n - 1 : 0;
  const int tot = nw + nwing;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += gridDim.x * blockDim.x) {
    const int iw = e < nw ? e : wing_iw;
    const cd* a = inv + nn * iw;
    const int* p = perm ? perm + (int64_t)n * iw : nullptr;
    const cd h = a[(int64_t)n * (p ? p[0] : 0)];
    if (e < nw) {
      llw[iw] = cdiv(cd{1.0, 0.0}, h);
    } else {
      const int j = e - nw + 1;
      wing[j - 1] = cdiv(a[(int64_t)n * (p ? p[j] : j)], h);
    }
  }
}

__global__ void set_w0_kernel(cd* __restrict__ w, int64_t wstride, int nw, const cd* __restrict__ w0) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nw; e += gridDim.x * blockDim.x) w[wstride * e] = w0[e];
}

inline int blocks_for(int64_t tot) { return (int)std::max<int64_t>(1, std::min<int64_t>((tot + 255) / 256, 65536)); }

ecalj_status check_common(ecalj_ctx* h, const void* chi, int ngb, int nw, int nspin) {
  if (!h || !chi) return ECALJ_ERR_ARG;
  if (ngb < 2 || nw < 1 || (nspin != 1 && nspin != 2)) return ECALJ_ERR_DIM;
  if (!h->d_vcousq || h->ngb != ngb) return ECALJ_ERR_STATE;
  return ECALJ_OK;
}

// ---- getw0 host helpers
// Gauss-Legendre nodes and weights on [-1, 1] (Newton iteration on P_n).
void gl_rule(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double t = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    x[i] = -t;
    x[n - 1 - i] = t;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - t * t) * dp * dp);
  }
}

// Normalized real harmonics at (cos theta, phi), index L = l*l + l + m:
// m > 0 ~ cos(m phi), m < 0 ~ sin(|m| phi), associated Legendre functions
// without the Condon-Shortley sign (the convention of sylm/sylmnc,
// S/sylm.f90:1-121).
void real_ylm(int lmax, double ct, double ph, double* y) {
  const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
  for (int m = 0; m <= lmax; ++m) {
    double pmm = 1.0;  // P_m^m = (2m-1)!! st^m
    for (int k = 1; k <= m; ++k) pmm *= (2 * k - 1) * st;
    double pl2 = 0.0, pl1 = pmm;
    for (int l = m; l <= lmax; ++l) {
      double plm;
      if (l == m) plm = pmm;
      else if (l == m + 1) plm = ct * (2 * m + 1) * pmm;
      else plm = ((2 * l - 1) * ct * pl1 - (l + m - 1) * pl2) / (l - m);
      if (l > m) {
        pl2 = pl1;
        pl1 = plm;
      }
      double ratio = 1.0;  // (l-m)!/(l+m)!
      for (int k = l - m + 1; k <= l + m; ++k) ratio /= k;
      const double nrm = std::sqrt((2 * l + 1) / (4 * kPi) * ratio);
      const int c = l * l + l;
      if (m == 0) {
        y[c] = nrm * plm;
      } else {
        y[c + m] = std::sqrt(2.0) * nrm * plm * std::cos(m * ph);
        y[c - m] = std::sqrt(2.0) * nrm * plm * std::sin(m * ph);
      }
    }
  }
}

// g[a + n12*(b + n12*c)] = int Y_a Y_b Y_c over the sphere, a, b < (l12+1)^2,
// c < (l3+1)^2, by a product rule exact for the degree 2*l12 + l3.
void gaunt_table(int l12, int l3, std::vector<double>& g) {
  const int n12 = (l12 + 1) * (l12 + 1), n3 = (l3 + 1) * (l3 + 1);
  const int lm = std::max(l12, l3), nl = (lm + 1) * (lm + 1);
  const int deg = 2 * l12 + l3;
  const int nth = deg / 2 + 2, nph = deg + 2;
  std::vector<double> xt, wt, y(nl);
  gl_rule(nth, xt, wt);
  g.assign((size_t)n12 * n12 * n3, 0.0);
  for (int it = 0; it < nth; ++it)
    for (int ip = 0; ip < nph; ++ip) {
      real_ylm(lm, xt[it], 2 * kPi * ip / nph, y.data());
      const double w = wt[it] * 2 * kPi / nph;
      for (int c = 0; c < n3; ++c)
        for (int b = 0; b < n12; ++b) {
          const double wb = w * y[b] * y[c];
          for (int a = 0; a < n12; ++a) g[a + n12 * (b + (size_t)n12 * c)] += wb * y[a];
        }
    }
}

typedef std::complex<double> zc;

// In-place inverse with partial pivoting (small matrices).  false if singular.
bool zinv_small(int n, zc* a) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(a[i + n * k]) > std::abs(a[p + n * k])) p = i;
    if (std::abs(a[p + n * k]) == 0.0) return false;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[k + n * j], a[p + n * j]);
    const zc d = 1.0 / a[k + n * k];
    a[k + n * k] = 1.0;
    for (int j = 0; j < n; ++j) a[k + n * j] *= d;
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      const zc f = a[i + n * k];
      if (f == 0.0) continue;
      a[i + n * k] = 0.0;
      for (int j = 0; j < n; ++j) a[i + n * j] -= f * a[k + n * j];
    }
  }
  for (int k = n - 1; k >= 0; --k)
    if (piv[k] != k)
      for (int i = 0; i < n; ++i) std::swap(a[i + n * k], a[i + n * piv[k]]);
  return true;
}

inline int lof(int lm0) { return (int)std::floor(std::sqrt((double)lm0) + 1e-9); }  // l of a 0-based index

}  // namespace
}  // namespace ecalj

using namespace ecalj;

extern "C" {

ecalj_status ecalj_wc_build_gamma_dev(ecalj_handle h, const void* chi, int ngb, int nw, int nspin,
                                      int is_x0_m_basis, int real_axis, void* wc, int ldw) {
  ECALJ_TRY(check_common(h, chi, ngb, nw, nspin));
  if (!wc) return ECALJ_ERR_ARG;
  if (ldw < ngb) return ECALJ_ERR_DIM;
  const int n = ngb, m = ngb - 1;
  const int64_t nn = (int64_t)n * n;
  cd *eps, *tmp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EPS, sizeof(cd) * nn * nw, (void**)&eps));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_TMP, sizeof(cd) * nn * nw, (void**)&tmp));
  const double fac = (nspin == 1) ? 2.0 : 1.0;  // S/m_llw.f90:88-92, 299-303
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 3, &tev));
  ECALJ_TRY(eps_build(h, chi, n, nw, fac, is_x0_m_basis != 0, 0.0, eps, tmp));
  hipLaunchKernelGGL(block_tail_kernel, dim3(blocks_for((int64_t)m * m * nw)), dim3(256), 0, h->stream, eps, n, nw,
                     tmp);
  ECALJ_HIP(hipGetLastError());
  const int* perm = nullptr;
  ECALJ_TRY(gj_inverse(h, tmp, m, m, (int64_t)m * m, nw, real_axis ? &perm : nullptr));
  cd* w = (cd*)wc;
  ECALJ_TRY(wc_form(h, tmp, perm, h->d_vcousq + 1, m, nw, w + 1 + ldw, ldw, (int64_t)ldw * n));
  hipLaunchKernelGGL(zero_head_kernel, dim3(blocks_for((2 * (int64_t)n - 1) * nw)), dim3(256), 0, h->stream, w, n,
                     (int64_t)ldw, (int64_t)ldw * n, nw);
  ECALJ_HIP(hipGetLastError());
  ECALJ_TRY(timing_end(h, 3, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_wc_build_gamma(ecalj_handle h, const double* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                                  int real_axis, double* wc, int ldw) {
  if (!h || !chi || !wc) return ECALJ_ERR_ARG;
  if (ldw < ngb || ngb < 2 || nw < 1) return ECALJ_ERR_DIM;
  DevBuf dc, dw;
  ECALJ_TRY(dc.put(h, chi, sizeof(cd) * (size_t)ngb * ngb * nw));
  const size_t bw = sizeof(cd) * (size_t)ldw * ngb * nw;
  ECALJ_TRY(dw.put(h, wc, bw));
  ECALJ_TRY(ecalj_wc_build_gamma_dev(h, dc.p, ngb, nw, nspin, is_x0_m_basis, real_axis, dw.p, ldw));
  ECALJ_HIP(hipMemcpyAsync(wc, dw.p, bw, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return check_flag(h);
}

ecalj_status ecalj_eps_head_dev(ecalj_handle h, const void* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                                double vcou1, int real_axis, void* llw, int wing_iw, void* wing) {
  ECALJ_TRY(check_common(h, chi, ngb, nw, nspin));
  if (!llw) return ECALJ_ERR_ARG;
  if (!(vcou1 > 0.0) || !std::isfinite(vcou1)) return ECALJ_ERR_ARG;
  const int n = ngb;
  const int64_t nn = (int64_t)n * n;
  cd *eps, *tmp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EPS, sizeof(cd) * nn * nw, (void**)&eps));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_TMP, sizeof(cd) * nn * nw, (void**)&tmp));
  const double fac = (nspin == 1) ? 2.0 : 1.0;
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 3, &tev));
  ECALJ_TRY(eps_build(h, chi, n, nw, fac, is_x0_m_basis != 0, vcou1, eps, tmp));
  const int* perm = nullptr;
  ECALJ_TRY(gj_inverse(h, eps, n, n, nn, nw, real_axis ? &perm : nullptr));  // zminv (S/m_llw.f90:234, 427)
  const int nwing = (wing && wing_iw >= 0 && wing_iw < nw) ? n - 1 : 0;
  hipLaunchKernelGGL(head_kernel, dim3(blocks_for(nw + nwing)), dim3(256), 0, h->stream, eps, perm, n, nw, (cd*)llw,
                     wing_iw, (cd*)wing);
  ECALJ_HIP(hipGetLastError());
  ECALJ_TRY(timing_end(h, 3, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_eps_head(ecalj_handle h, const double* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                            double vcou1, int real_axis, double* llw, int wing_iw, double* wing) {
  if (!h || !chi || !llw) return ECALJ_ERR_ARG;
  if (ngb < 2 || nw < 1) return ECALJ_ERR_DIM;
  const bool want_wing = wing && wing_iw >= 0 && wing_iw < nw;
  DevBuf dc, dl, dwg;
  ECALJ_TRY(dc.put(h, chi, sizeof(cd) * (size_t)ngb * ngb * nw));
  ECALJ_TRY(dl.alloc(sizeof(cd) * nw));
  ECALJ_TRY(dwg.alloc(sizeof(cd) * (ngb - 1)));
  ECALJ_TRY(ecalj_eps_head_dev(h, dc.p, ngb, nw, nspin, is_x0_m_basis, vcou1, real_axis, dl.p,
                               want_wing ? wing_iw : -1, want_wing ? dwg.p : nullptr));
  ECALJ_HIP(hipMemcpyAsync(llw, dl.p, sizeof(cd) * nw, hipMemcpyDeviceToHost, h->stream));
  if (want_wing) ECALJ_HIP(hipMemcpyAsync(wing, dwg.p, sizeof(cd) * (ngb - 1), hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return check_flag(h);
}

ecalj_status ecalj_wv0_patch_dev(ecalj_handle h, void* w, int ngb, int ldw, int nw, const double* w0,
                                 int is_wc_m_basis) {
  if (!h || !w || !w0) return ECALJ_ERR_ARG;
  if (ngb < 1 || nw < 1 || ldw < ngb) return ECALJ_ERR_DIM;
  if (is_wc_m_basis && (!h->d_zcousq || h->ngb != ngb)) return ECALJ_ERR_STATE;
  const int n = ngb;
  const int64_t nn = (int64_t)n * n, ws = (int64_t)ldw * n;
  std::vector<unsigned char> blob(sizeof(cd) * nw);
  std::memcpy(blob.data(), w0, blob.size());
  unsigned char* d_w0 = nullptr;
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_GAM, blob, &d_w0));
  hipLaunchKernelGGL(set_w0_kernel, dim3(blocks_for(nw)), dim3(256), 0, h->stream, (cd*)w, ws, nw, (const cd*)d_w0);
  ECALJ_HIP(hipGetLastError());
  if (!is_wc_m_basis) return ECALJ_OK;
  cd* tmp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_TMP, sizeof(cd) * nn * nw, (void**)&tmp));
  return w_e2m_inplace(h, (cd*)w, n, nw, (int64_t)ldw, tmp);  // (S/m_w0w0i.f90:130-131)
}

ecalj_status ecalj_wv0_patch(ecalj_handle h, double* w, int ngb, int ldw, int nw, const double* w0,
                             int is_wc_m_basis) {
  if (!h || !w || !w0) return ECALJ_ERR_ARG;
  if (ngb < 1 || nw < 1 || ldw < ngb) return ECALJ_ERR_DIM;
  DevBuf dw;
  const size_t bw = sizeof(cd) * (size_t)ldw * ngb * nw;
  ECALJ_TRY(dw.put(h, w, bw));
  ECALJ_TRY(ecalj_wv0_patch_dev(h, dw.p, ngb, ldw, nw, w0, is_wc_m_basis));
  ECALJ_HIP(hipMemcpyAsync(w, dw.p, bw, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

// getw0 (S/m_w0w0i.f90:146-311)
ecalj_status ecalj_getw0(int ii, int ie, int nq0i, int lmxax, const double* llw, const double* dmlx,
                         const double* epinvq0i, const double* wklm, double wbz, const double* epinv, double* w0,
                         double* llmat) {
  if (!llw || !dmlx || !epinvq0i || !wklm || !w0 || (ii <= 0 && llmat && !epinv)) return ECALJ_ERR_ARG;
  if (ie < ii || nq0i < 1 || lmxax < 0 || lmxax > 12 || !(wbz > 0.0)) return ECALJ_ERR_DIM;
  const int nw = ie - ii + 1, nlk = (lmxax + 1) * (lmxax + 1);
  const zc* L = reinterpret_cast<const zc*>(llw);
  // llw_invr = epinvq0i^-1 llw: coefficients of llw(q) in the invariant tensors (:168-173)
  std::vector<zc> em(nq0i * (size_t)nq0i);
  for (size_t e = 0; e < em.size(); ++e) em[e] = epinvq0i[e];
  if (!zinv_small(nq0i, em.data())) return ECALJ_ERR_SINGULAR;
  std::vector<zc> invr((size_t)nw * nq0i), yl((size_t)nw * 9);
  for (int iw = 0; iw < nw; ++iw) {
    for (int i = 0; i < nq0i; ++i) {
      zc s = 0.0;
      for (int j = 0; j < nq0i; ++j) s += em[i + (size_t)nq0i * j].real() * L[iw + (size_t)nw * j];
      invr[iw + (size_t)nw * i] = s;
    }
    for (int lm = 0; lm < 9; ++lm) {  // the l = 0 and l = 2 components (:174-179)
      if (lm >= 1 && lm <= 3) continue;
      zc s = 0.0;
      for (int i = 0; i < nq0i; ++i) s += dmlx[i + (size_t)nq0i * lm] * invr[iw + (size_t)nw * i];
      yl[lm + 9 * (size_t)iw] = s;
    }
  }
  if (ii <= 0 && llmat) {  // the static 3x3 tensor (:182-188)
    zc* lm = reinterpret_cast<zc*>(llmat);
    for (int e = 0; e < 9; ++e) lm[e] = 0.0;
    for (int i = 0; i < nq0i; ++i)
      for (int e = 0; e < 9; ++e) lm[e] += invr[(size_t)(-ii) + (size_t)nw * i] * epinv[e + 9 * i];
  }
  const int lxx = std::max(lmxax, 2), n12 = (lxx + 1) * (lxx + 1);
  std::vector<double> g;
  gaunt_table(lxx, 2, g);
  std::vector<int> ev;  // even-l indices up to lmxax
  for (int lm = 0; lm < nlk; ++lm)
    if (lof(lm) % 2 == 0) ev.push_back(lm);
  const int ne = (int)ev.size();
  std::vector<zc> a((size_t)ne * ne);
  const double fpi = 4 * kPi;
  zc* out = reinterpret_cast<zc*>(w0);
  for (int iw = 0; iw < nw; ++iw) {
    // cgllw(L1, L2) = sum_K <Y_L2 Y_K Y_L1> llwyl(K), K in {00, 2m} (:246-260);
    // the table holds K in its third index (the coefficient is symmetric)
    for (int c2 = 0; c2 < ne; ++c2)
      for (int c1 = 0; c1 < ne; ++c1) {
        const int l2 = ev[c2], l1 = ev[c1];
        zc s = g[l2 + n12 * (l1 + (size_t)n12 * 0)] * yl[9 * (size_t)iw];
        for (int k = 4; k < 9; ++k) s += g[l2 + n12 * (l1 + (size_t)n12 * k)] * yl[k + 9 * (size_t)iw];
        a[c1 + (size_t)ne * c2] = s;
      }
    // 1 = llw(q) K(q) projected on the even harmonics: K_L = sqrt(4 pi) cgllw^-1(L, 00) (:264-266)
    if (!zinv_small(ne, a.data())) return ECALJ_ERR_SINGULAR;
    zc w = 0.0;
    for (int c1 = 0; c1 < ne; ++c1) {  // (:269-276)
      w += wklm[ev[c1]] * fpi * std::sqrt(fpi) * a[c1] / wbz;
      if (c1 == 0) w -= wklm[0] * std::pow(fpi, 1.5) / wbz;  // the bare Coulomb part
    }
    out[iw] = w;
  }
  return ECALJ_OK;
}

}  // extern "C"
