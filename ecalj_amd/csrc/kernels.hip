// kernels.hip -- the FP64 kernels of the chi0 -> W -> Sigma_c path and their
// launchers.  Every launch checks on the host that its grid and operand
// extents match the buffers it was given before it is issued.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {

// ==========================================================================
// Generic GEMM dispatcher.  Complex operands are complex128 in mode M64 and
// complex64 in M32 / MBF3.
template <int MODE, class T, int BM, int BN>
static ecalj_status launch_gemm(ecalj_ctx* h, const GemmArgs& g0, bool at, bool areal, bool bt) {
  GemmArgs g = g0;
  dim3 grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, g.batch);
  g.xg = g.xt = 0;
  // Many matrices: the XCD order (GemmArgs), when the matrices deal out
  // evenly over the 8 XCDs (a multiple of 8, or enough that the last round
  // is a small share)
  if (g.batch >= 8 && (g.batch % 8 == 0 || g.batch >= 128)) {
    const int64_t xt = (int64_t)grid.x * grid.y, p = 8 * xt * ((g.batch + 7) / 8);
    if (p <= 0x7fffffff) {
      g.xg = (int)grid.x;
      g.xt = (int)xt;
      grid = dim3((unsigned)p, 1, 1);
    }
  }
  if (grid.y > 65535 || grid.z > 65535) return ECALJ_ERR_DIM;
  const int key = (at ? 4 : 0) | (areal ? 2 : 0) | (bt ? 1 : 0);
  switch (key) {
    case 0: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, false, false, false>), grid, dim3(kThreads), 0, h->stream, g); break;
    case 1: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, false, false, true>), grid, dim3(kThreads), 0, h->stream, g); break;
    case 2: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, false, true, false>), grid, dim3(kThreads), 0, h->stream, g); break;
    case 3: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, false, true, true>), grid, dim3(kThreads), 0, h->stream, g); break;
    case 4: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, true, false, false>), grid, dim3(kThreads), 0, h->stream, g); break;
    case 5: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, true, false, true>), grid, dim3(kThreads), 0, h->stream, g); break;
    case 6: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, true, true, false>), grid, dim3(kThreads), 0, h->stream, g); break;
    case 7: hipLaunchKernelGGL((zgemm_kernel<MODE, T, BM, BN, true, true, true>), grid, dim3(kThreads), 0, h->stream, g); break;
  }
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status gemm(ecalj_ctx* h, const GemmCall& c) {
  if (c.M < 0 || c.N < 0 || c.K < 0 || c.batch < 1 || c.div < 1) return ECALJ_ERR_DIM;
  if (c.M == 0 || c.N == 0) return ECALJ_OK;
  GemmArgs g;
  g.M = c.M; g.N = c.N; g.K = c.K; g.batch = c.batch; g.div = c.div;
  g.A = c.A; g.lda = c.lda; g.bA = c.bA; g.csA = c.ca ? -1.0 : 1.0;
  g.B = c.B; g.ldb = c.ldb; g.bB = c.bB; g.csB = c.cb ? -1.0 : 1.0;
  g.C = c.C; g.ldc = c.ldc; g.bC = c.bC;
  g.alr = c.alpha.re; g.ali = c.alpha.im; g.ber = c.beta.re; g.bei = c.beta.im;
  g.conjC = c.conjC; g.epi = c.epi; g.svec = c.svec; g.gj_k0 = c.gj_k0; g.gj_k1 = c.gj_k1;
  const bool at = c.ta, bt = c.tb;
  switch (c.mode) {
    case M64:
      if (c.N <= 32) return launch_gemm<M64, cd, 64, 32>(h, g, at, c.areal, bt);
      return launch_gemm<M64, cd, 64, 64>(h, g, at, c.areal, bt);
    case M32:
      return launch_gemm<M32, cf, 64, 64>(h, g, at, c.areal, bt);
    case MBF3:
      return launch_gemm<MBF3, cf, 64, 64>(h, g, at, c.areal, bt);
  }
  return ECALJ_ERR_ARG;
}

// complex128 <-> complex64 conversion
__global__ void cvt_kernel_d2f(const cd* __restrict__ a, cf* __restrict__ b, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    b[e] = cf{(float)a[e].re, (float)a[e].im};
}
__global__ void cvt_kernel_f2d(const cf* __restrict__ a, cd* __restrict__ b, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    b[e] = cd{(double)a[e].re, (double)a[e].im};
}

ecalj_status convert(ecalj_ctx* h, const void* src, bool src64, void* dst, bool dst64, int64_t n) {
  if (n == 0 || src == dst) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  if (!src64 && dst64)
    hipLaunchKernelGGL(cvt_kernel_f2d, dim3(blocks), dim3(256), 0, h->stream, (const cf*)src, (cd*)dst, n);
  else if (src64 && !dst64)
    hipLaunchKernelGGL(cvt_kernel_d2f, dim3(blocks), dim3(256), 0, h->stream, (const cd*)src, (cf*)dst, n);
  else
    ECALJ_HIP(hipMemcpyAsync(dst, src, n * (src64 ? sizeof(cd) : sizeof(cf)), hipMemcpyDeviceToDevice, h->stream));
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// ==========================================================================
// ppb relayout: P2[(i1 + nv*I) + nv*md*i2] = ppb[i1 + nv*i2 + nv*nv*I]
__global__ void ppb_relayout_kernel(const double* __restrict__ src, double* __restrict__ dst, int nv,
                                    int md) {
  const int64_t n = (int64_t)nv * nv * md;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i1 = e % nv;
    const int64_t t = e / nv;
    const int64_t i2 = t % nv;
    const int64_t I = t / nv;
    dst[(i1 + (int64_t)nv * I) + (int64_t)nv * md * i2] = src[e];
  }
}

ecalj_status ppb_relayout(ecalj_ctx* h, const double* src, double* dst, int nv, int md) {
  const int64_t n = (int64_t)nv * nv * md;
  if (n == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(ppb_relayout_kernel, dim3(blocks), dim3(256), 0, h->stream, src, dst, nv, md);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// ==========================================================================
// x0gemm: gather-Herk per histogram bin.
//   rcxq_slice(i,j) += sum_t conj(z_t(i)) * w_t * z_t(ipr_col-1+j)
// Columns z_t live anywhere in `z` (offset col[t]); the bin's pair list is
// [bstart[l], bstart[l+1]).  Work item = (list l, output slice, tile ti, tj).

template <class T>
struct AccGatherA {  // A(i,t) = conj(z[col[t] + i])
  typedef typename ElemT<T>::V V;
  static constexpr bool kcontig = false;
  const T* z; const int64_t* col; int X, K;
  __device__ void get(int x, int k, V& re, V& im) const {
    re = 0;
    im = 0;
    if (x >= X || k >= K) return;
    ld_c<V>(z, col[k] + x, re, im);
    im = -im;
  }
};

template <class T>
struct AccGatherB {  // B(t,j) = w[t] * z[col[t] + off + j]
  typedef typename ElemT<T>::V V;
  static constexpr bool kcontig = false;
  const T* z; const int64_t* col; const double* w; int64_t off; int X, K;
  __device__ void get(int x, int k, V& re, V& im) const {
    re = 0;
    im = 0;
    if (x >= X || k >= K) return;
    const V wt = (V)w[k];
    ld_c<V>(z, col[k] + off + x, re, im);
    re *= wt;
    im *= wt;
  }
};

struct X0Item { int list, slice, ti, tj; };

template <int MODE, class T, int BT>
__global__ __launch_bounds__(kThreads) void x0_herk_kernel(const T* __restrict__ z,
                                                           const int64_t* __restrict__ col,
                                                           const double* __restrict__ w,
                                                           const int64_t* __restrict__ bstart,
                                                           const X0Item* __restrict__ items,
                                                           T* __restrict__ rcxq, int64_t slice_elems,
                                                           int npr, int npr_col, int ipr_col,
                                                           int mirror) {
  constexpr int BK = ModeBK<MODE>::value;
  __shared__ typename StageOf<MODE, BT, BT, BK>::type s;
  const X0Item it = items[blockIdx.x];
  const int64_t b0 = bstart[it.list];
  const int K = (int)(bstart[it.list + 1] - b0);
  const int m0 = it.ti * BT, n0 = it.tj * BT;
  constexpr int KV = ModeKV<MODE>::value;
  Ld<AccGatherA<T>, BT, BK, true, KV> la{AccGatherA<T>{z, col + b0, npr, K}};
  Ld<AccGatherB<T>, BT, BK, false, KV> lb{AccGatherB<T>{z, col + b0, w + b0, (int64_t)(ipr_col - 1), npr_col, K}};
  Core<MODE, BT, BT, BK, false> core;
  gemm_mainloop<MODE, BT, BT, BK, false>(la, lb, m0, n0, K, s, core);
  T* r = rcxq + slice_elems * it.slice;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool mir = mirror && (it.ti != it.tj);
  core.for_each(wave >> 1, wave & 1, lane, [&](int rr, int cc, double xr, double xi) {
    const int i = m0 + rr, j = n0 + cc;
    if (i >= npr || j >= npr_col) return;
    T* p = r + i + (int64_t)npr * j;
    double orr, oi;
    ld_c2(p, orr, oi);
    st_c(p, orr + xr, oi + xi);
    if (mir) {
      T* q = r + j + (int64_t)npr * i;
      ld_c2(q, orr, oi);
      st_c(q, orr + xr, oi - xi);
    }
  });
}

ecalj_status x0_herk(ecalj_ctx* h, int mode, const void* z, const int64_t* d_col, const double* d_w,
                     const int64_t* d_bstart, const void* d_items, int nitems, void* rcxq,
                     int64_t slice_elems, int npr, int npr_col, int ipr_col, int mirror) {
  if (nitems == 0) return ECALJ_OK;
  const X0Item* it = static_cast<const X0Item*>(d_items);
  if (mode == M64)
    hipLaunchKernelGGL((x0_herk_kernel<M64, cd, 64>), dim3(nitems), dim3(kThreads), 0, h->stream, (const cd*)z, d_col,
                       d_w, d_bstart, it, (cd*)rcxq, slice_elems, npr, npr_col, ipr_col, mirror);
  else if (mode == M32)
    hipLaunchKernelGGL((x0_herk_kernel<M32, cf, 64>), dim3(nitems), dim3(kThreads), 0, h->stream, (const cf*)z, d_col,
                       d_w, d_bstart, it, (cf*)rcxq, slice_elems, npr, npr_col, ipr_col, mirror);
  else
    hipLaunchKernelGGL((x0_herk_kernel<MBF3, cf, 64>), dim3(nitems), dim3(kThreads), 0, h->stream, (const cf*)z, d_col,
                       d_w, d_bstart, it, (cf*)rcxq, slice_elems, npr, npr_col, ipr_col, mirror);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// ==========================================================================
// Hilbert transforms of the binned Im chi (dpsion_chiq):
//   out(e, v) = sum_iw coef(v, iw) rcxq(e, iw)                 (imaginary axis)
//   out(e, v) = i dscale(v0+v) rcxq(e, v0+v) + sum_iw coef(v, iw) rcxq(e, iw)
//                                                              (real axis, dscale != null)
// for the output chunk v in [v0, v0+nv); coef(v, iw) = coef[v + ldc*(iw-1)]
// (chunk base already applied), zero-padded so that NV coefficients can be
// read for every chunk: the FMA loop runs over all NV outputs without a
// per-output test.  FP64 accumulation, complex128 output; rcxq complex128
// or complex64.  With `scale`, rcxq(:, iw) *= scale(iw) is stored back (the
// reference scales rcxq in place once).
template <int NV, class T>
__global__ __launch_bounds__(256) void hilbert_kernel(T* __restrict__ rcxq, int64_t slice, int nwhis,
                                                      int nv, int ldc, const double* __restrict__ coef,
                                                      const double* __restrict__ scale, int v0,
                                                      const double* __restrict__ dscale,
                                                      cd* __restrict__ out, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  double ar[NV], ai[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) { ar[v] = 0.0; ai[v] = 0.0; }
  // four bins per round: their loads are issued before any store-back, so
  // four reads of rcxq per lane are in flight (the loop is latency bound)
  int iw = 1;
  for (; iw + 3 <= nwhis; iw += 4) {
    double xr[4], xi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) ld_c2(rcxq + e + slice * (iw + u), xr[u], xi[u]);
    if (scale) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double s = scale[iw + u - 1];
        st_c(rcxq + e + slice * (iw + u), s * xr[u], s * xi[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double* c = coef + (int64_t)ldc * (iw + u - 1);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        ar[v] += c[v] * xr[u];
        ai[v] += c[v] * xi[u];
      }
    }
  }
  for (; iw <= nwhis; ++iw) {
    double xr, xi;
    ld_c2(rcxq + e + slice * iw, xr, xi);
    if (scale) {
      const double s = scale[iw - 1];
      st_c(rcxq + e + slice * iw, s * xr, s * xi);
    }
    const double* c = coef + (int64_t)ldc * (iw - 1);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      ar[v] += c[v] * xr;
      ai[v] += c[v] * xi;
    }
  }
  if (dscale) {  // i*dscale*rcxq(:, it): slot it of rcxq is the bin centred at freq_r(it); slot 0 the zero frequency
#pragma unroll
    for (int v = 0; v < NV; ++v)
      if (v < nv) {
        double xr, xi;
        ld_c2(rcxq + e + slice * (v0 + v), xr, xi);
        const double s = dscale[v0 + v];
        ar[v] -= s * xi;
        ai[v] += s * xr;
      }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v)
    if (v < nv) out[e + n * v] = cd{ar[v], ai[v]};
}

template <int NV, class T>
static void hilbert_launch(ecalj_ctx* h, T* rcxq, int64_t slice, int nwhis, int nv, int ldc, const double* coef,
                           const double* scale, int v0, const double* dscale, cd* out, int64_t n) {
  const dim3 grid((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((hilbert_kernel<NV, T>), grid, dim3(256), 0, h->stream, rcxq, slice, nwhis, nv, ldc, coef,
                     scale, v0, dscale, out, n);
}

template <class T>
static ecalj_status hilbert_t(ecalj_ctx* h, T* rcxq, int64_t slice, int nwhis, int nv, int ldc, const double* d_coef,
                              const double* d_scale, const double* d_dscale, cd* out, int64_t n) {
  // chunks of at most 32 outputs (64 FP64 accumulators per lane); each chunk
  // re-reads rcxq.  The chunks run last to first, so that the one that stores
  // the scaled rcxq (chunk 0) runs after every other chunk has read it.
  for (int v0 = ((nv - 1) / 32) * 32; v0 >= 0; v0 -= 32) {
    const int c = std::min(32, nv - v0);
    const double* sc = (v0 == 0) ? d_scale : nullptr;
    if (c <= 8)
      hilbert_launch<8, T>(h, rcxq, slice, nwhis, c, ldc, d_coef + v0, sc, v0, d_dscale, out + n * v0, n);
    else if (c <= 16)
      hilbert_launch<16, T>(h, rcxq, slice, nwhis, c, ldc, d_coef + v0, sc, v0, d_dscale, out + n * v0, n);
    else if (c <= 24)  // nf + 1 = 17..24 imaginary frequencies: no FMAs on 8 padded outputs
      hilbert_launch<24, T>(h, rcxq, slice, nwhis, c, ldc, d_coef + v0, sc, v0, d_dscale, out + n * v0, n);
    else
      hilbert_launch<32, T>(h, rcxq, slice, nwhis, c, ldc, d_coef + v0, sc, v0, d_dscale, out + n * v0, n);
    ECALJ_HIP(hipGetLastError());
  }
  return ECALJ_OK;
}

ecalj_status hilbert(ecalj_ctx* h, bool wide, void* rcxq, int64_t slice, int nwhis, int nv, int ldc,
                     const double* d_coef, const double* d_scale, const double* d_dscale, cd* out, int64_t n) {
  if (n == 0 || nv == 0) return ECALJ_OK;
  if (ldc < hilbert_ldc(nv)) return ECALJ_ERR_ARG;
  if (wide) return hilbert_t<cd>(h, (cd*)rcxq, slice, nwhis, nv, ldc, d_coef, d_scale, d_dscale, out, n);
  return hilbert_t<cf>(h, (cf*)rcxq, slice, nwhis, nv, ldc, d_coef, d_scale, d_dscale, out, n);
}

// ==========================================================================
// Composed-weight chi0: B(p,j) = w(v,p) * z(j,p), A(i,p) = conj(z(i,p)).
template <class T>
struct AccWZT {  // B(k=p, j) = w[v + nf*p] * z[j + ni*p]
  typedef typename ElemT<T>::V V;
  static constexpr bool kcontig = false;
  const T* z; const double* w; int64_t ldz; int nf, v, X, K;
  __device__ void get(int x, int k, V& re, V& im) const {
    re = 0;
    im = 0;
    if (x >= X || k >= K) return;
    const V wt = (V)w[v + (int64_t)nf * k];
    ld_c<V>(z, x + ldz * k, re, im);
    re *= wt;
    im *= wt;
  }
};

template <int MODE, class T>
__global__ __launch_bounds__(kThreads) void chi0_freq_kernel(const T* __restrict__ z,
                                                             const double* __restrict__ w, int ni,
                                                             int npairs, int nf, cd* __restrict__ chi) {
  constexpr int BT = 64, BK = ModeBK<MODE>::value;
  typedef typename ElemT<T>::V V;
  __shared__ typename StageOf<MODE, BT, BT, BK>::type s;
  const int v = blockIdx.z;
  const int m0 = blockIdx.x * BT, n0 = blockIdx.y * BT;
  constexpr int KV = ModeKV<MODE>::value;
  Ld<AccN<T, false>, BT, BK, true, KV> la{AccN<T, false>{z, (int64_t)ni, ni, npairs, (V)-1}};  // conj(z(i,p))
  Ld<AccWZT<T>, BT, BK, false, KV> lb{AccWZT<T>{z, w, (int64_t)ni, nf, v, ni, npairs}};
  Core<MODE, BT, BT, BK, false> core;
  gemm_mainloop<MODE, BT, BT, BK, false>(la, lb, m0, n0, npairs, s, core);
  cd* c = chi + (int64_t)ni * ni * v;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  core.for_each(wave >> 1, wave & 1, lane, [&](int rr, int cc, double xr, double xi) {
    const int i = m0 + rr, j = n0 + cc;
    if (i < ni && j < ni) c[i + (int64_t)ni * j] = cd{xr, xi};
  });
}

ecalj_status chi0_freq(ecalj_ctx* h, int mode, const void* z, const double* w, int ni, int npairs, int nf, cd* chi) {
  if (ni == 0 || nf == 0) return ECALJ_OK;
  if (nf > 65535) return ECALJ_ERR_DIM;
  dim3 grid((ni + 63) / 64, (ni + 63) / 64, nf);
  if (mode == M64)
    hipLaunchKernelGGL((chi0_freq_kernel<M64, cd>), grid, dim3(kThreads), 0, h->stream, (const cd*)z, w, ni, npairs, nf, chi);
  else if (mode == M32)
    hipLaunchKernelGGL((chi0_freq_kernel<M32, cf>), grid, dim3(kThreads), 0, h->stream, (const cf*)z, w, ni, npairs, nf, chi);
  else
    hipLaunchKernelGGL((chi0_freq_kernel<MBF3, cf>), grid, dim3(kThreads), 0, h->stream, (const cf*)z, w, ni, npairs, nf, chi);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// ==========================================================================
// W build helpers
// eps = delta - fac * s_i chi_ij s_j (no M->E transform)
__global__ void eps_direct_kernel(const cd* __restrict__ chi, const double* __restrict__ s, double fac,
                                  int n, int nw, cd* __restrict__ eps) {
  const int64_t tot = (int64_t)n * n * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ij = e % ((int64_t)n * n);
    const int i = (int)(ij % n), j = (int)(ij / n);
    const double sc = fac * s[i] * s[j];
    cd x = chi[e];
    eps[e] = cd{(i == j ? 1.0 : 0.0) - sc * x.re, -sc * x.im};
  }
}

// w = s_i inv_ij s_j - delta_ij s_i s_j, written with leading dimension ldw.
// perm (optional): column j of inv is stored as column perm[j + n*iw]
// (the pivoted inverse leaves its columns permuted).
__global__ void wc_form_kernel(const cd* __restrict__ inv, const int* __restrict__ perm,
                               const double* __restrict__ s, int n, int nw, cd* __restrict__ w, int64_t ldw,
                               int64_t wstride) {
  const int64_t tot = (int64_t)n * n * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t nn = (int64_t)n * n;
    const int iw = (int)(e / nn);
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    const double sc = s[i] * s[j];
    const int jc = perm ? perm[j + (int64_t)n * iw] : j;
    cd x = inv[nn * iw + i + (int64_t)n * jc];
    w[(int64_t)iw * wstride + i + ldw * j] = cd{sc * x.re - (i == j ? sc : 0.0), sc * x.im};
  }
}

// eps = delta - s_i chiE_ij s_j from a complex64 chi_E (already scaled), in FP64
// (head > 0: eps(1,1) = 1 - head chiE(1,1), the offset points near Gamma)
__global__ void eps_narrow_kernel(const cf* __restrict__ chi, const double* __restrict__ s, double head, int n,
                                  int nw, cd* __restrict__ eps) {
  const int64_t tot = (int64_t)n * n * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ij = e % ((int64_t)n * n);
    const int i = (int)(ij % n), j = (int)(ij / n);
    const double sc = (head > 0.0 && ij == 0) ? head : s[i] * s[j];
    const cf x = chi[e];
    eps[e] = cd{(i == j ? 1.0 : 0.0) - sc * (double)x.re, -sc * (double)x.im};
  }
}

// wc_form into complex64, leading dimension n (the E -> M input of the mixed form)
__global__ void wc_form32_kernel(const cd* __restrict__ inv, const int* __restrict__ perm,
                                 const double* __restrict__ s, int n, int nw, cf* __restrict__ w) {
  const int64_t nn = (int64_t)n * n, tot = nn * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int iw = (int)(e / nn);
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    const double sc = s[i] * s[j];
    const int jc = perm ? perm[j + (int64_t)n * iw] : j;
    const cd x = inv[nn * iw + i + (int64_t)n * jc];
    w[e] = cf{(float)(sc * x.re - (i == j ? sc : 0.0)), (float)(sc * x.im)};
  }
}

// dst(i, j, iw) (leading dimension ldw, record stride wstride) = src(i, j, iw) (ld n)
__global__ void widen_ld_kernel(const cf* __restrict__ src, int n, int nw, cd* __restrict__ dst, int64_t ldw,
                                int64_t wstride) {
  const int64_t nn = (int64_t)n * n, tot = nn * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int iw = (int)(e / nn);
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    const cf x = src[e];
    dst[wstride * iw + i + ldw * j] = cd{(double)x.re, (double)x.im};
  }
}

ecalj_status eps_narrow(ecalj_ctx* h, const cf* chi, const double* s, double head, int n, int nw, cd* eps) {
  const int64_t tot = (int64_t)n * n * nw;
  if (tot == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(eps_narrow_kernel, dim3(blocks), dim3(256), 0, h->stream, chi, s, head, n, nw, eps);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status wc_form32(ecalj_ctx* h, const cd* inv, const int* perm, const double* s, int n, int nw, cf* w) {
  const int64_t tot = (int64_t)n * n * nw;
  if (tot == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(wc_form32_kernel, dim3(blocks), dim3(256), 0, h->stream, inv, perm, s, n, nw, w);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status widen_ld(ecalj_ctx* h, const cf* src, int n, int nw, cd* dst, int64_t ldw, int64_t wstride) {
  const int64_t tot = (int64_t)n * n * nw;
  if (tot == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(widen_ld_kernel, dim3(blocks), dim3(256), 0, h->stream, src, n, nw, dst, ldw, wstride);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status eps_direct(ecalj_ctx* h, const cd* chi, const double* s, double fac, int n, int nw, cd* eps) {
  const int64_t tot = (int64_t)n * n * nw;
  if (tot == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(eps_direct_kernel, dim3(blocks), dim3(256), 0, h->stream, chi, s, fac, n, nw, eps);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status wc_form(ecalj_ctx* h, const cd* inv, const int* perm, const double* s, int n, int nw, cd* w,
                     int64_t ldw, int64_t wstride) {
  const int64_t tot = (int64_t)n * n * nw;
  if (tot == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(wc_form_kernel, dim3(blocks), dim3(256), 0, h->stream, inv, perm, s, n, nw, w, ldw, wstride);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// --------------------------------------------------------------------------
// Blocked Gauss-Jordan inversion without pivoting, batched.
// Step for pivot block K = [k0, k0+b):
//   (1) gj_diag: P = A[K,K] -> Pinv (LDS elimination); Cp = A[:,K]; A[K,K] = I
//   (2) row panel:  A[K,:] = Pinv * A[K,:]          (A[K,K] becomes Pinv)
//   (3) trailing:   A[i,:] = (j in K ? 0 : A[i,:]) - Cp[i,:] * A[K,:],  i not in K
// Pivot block width.  (64 halves the panel steps, but its 64-step diagonal
// inverse on one block per matrix costs more than the wider updates save.)
constexpr int GJB = 32;
// Outer pivot block.  The trailing update of each outer step is a rank-GJO
// GEMM (K = GJO), so the matrix is read and written n/GJO times instead of
// n/GJB times; the GJO x GJO diagonal block is inverted by the GJB scheme.
constexpr int GJO = 128;

__global__ __launch_bounds__(256) void gj_diag_kernel(cd* __restrict__ a, int n, int64_t lda,
                                                      int64_t astride, int k0, int b,
                                                      cd* __restrict__ pinv, cd* __restrict__ cp,
                                                      int* __restrict__ flag) {
  __shared__ cd p[GJB][GJB + 1];
  const int mat = blockIdx.x;
  cd* A = a + astride * mat;
  cd* P = pinv + (int64_t)GJB * GJB * mat;
  cd* Cp = cp + (int64_t)n * GJB * mat;
  const int tid = threadIdx.x;
  // copy the column panel and the pivot block
  for (int64_t e = tid; e < (int64_t)n * b; e += blockDim.x) {
    const int i = (int)(e % n), l = (int)(e / n);
    Cp[i + (int64_t)n * l] = A[i + lda * (k0 + l)];
  }
  for (int e = tid; e < b * b; e += blockDim.x) {
    const int i = e % b, l = e / b;
    p[i][l] = A[(k0 + i) + lda * (k0 + l)];
  }
  __syncthreads();
  // in-place Gauss-Jordan on the b x b block; each step reads the old values
  // of a thread's elements (b*b <= 4 * blockDim), then writes: two barriers
  //   row kk: x/piv; column kk: -f/piv; pivot: 1/piv; others: y - f * (x/piv)
  constexpr int EPT = GJB * GJB / 256;
  for (int kk = 0; kk < b; ++kk) {
    const cd piv = p[kk][kk];
    if (!(hypot(piv.re, piv.im) > 1e-300)) {
      if (tid == 0) *flag = 1;
    }
    const double mag = piv.re * piv.re + piv.im * piv.im;
    const double dr = piv.re / mag, di = -piv.im / mag;  // 1/piv
    cd nv[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = tid + 256 * q, i = e % GJB, l = e / GJB;
      if (i >= b || l >= b) continue;
      const cd x = p[kk][l], f = p[i][kk];
      const cd xs = cd{x.re * dr - x.im * di, x.re * di + x.im * dr};
      if (i == kk && l == kk)
        nv[q] = cd{dr, di};
      else if (i == kk)
        nv[q] = xs;
      else if (l == kk)
        nv[q] = cd{-(f.re * dr - f.im * di), -(f.re * di + f.im * dr)};
      else
        nv[q] = cd{p[i][l].re - (f.re * xs.re - f.im * xs.im), p[i][l].im - (f.re * xs.im + f.im * xs.re)};
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = tid + 256 * q, i = e % GJB, l = e / GJB;
      if (i < b && l < b) p[i][l] = nv[q];
    }
    __syncthreads();
  }
  for (int e = tid; e < GJB * GJB; e += blockDim.x) {
    const int i = e % GJB, l = e / GJB;
    P[i + GJB * l] = (i < b && l < b) ? p[i][l] : cd{0.0, 0.0};
  }
  for (int e = tid; e < b * b; e += blockDim.x) {
    const int i = e % b, l = e / b;
    A[(k0 + i) + lda * (k0 + l)] = cd{i == l ? 1.0 : 0.0, 0.0};
  }
}

// Row exchanges for pivot block K = [k0, k0+b), b <= GJO (partial pivoting,
// as zgetrf): LU with exchanges of the panel rows k0..n-1 of columns K, on a
// copy in `work` (leading dimension n, n*b per matrix).  Only the exchanges are kept:
// ipiv[k0+kk] = row exchanged with row k0+kk (absolute, 0-based).  The pivot
// is the first row of largest |re|+|im| (izamax).  The in-order elimination
// of the exchanged block then meets exactly these pivots.
__global__ __launch_bounds__(256) void gj_panel_kernel(const cd* __restrict__ a, int n, int64_t lda,
                                                       int64_t astride, int k0, int b, cd* __restrict__ work,
                                                       int* __restrict__ ipiv, int* __restrict__ flag) {
  __shared__ double red_v[4];
  __shared__ int red_i[4];
  __shared__ cd prow[GJO];
  __shared__ int s_p;
  const int mat = blockIdx.x;
  const cd* A = a + astride * mat;
  cd* W = work + (int64_t)n * b * mat;
  int* P = ipiv + (int64_t)n * mat;
  const int m = n - k0;
  const int tid = threadIdx.x;
  for (int64_t e = tid; e < (int64_t)m * b; e += blockDim.x) {
    const int i = (int)(e % m), l = (int)(e / m);
    W[i + (int64_t)n * l] = A[(k0 + i) + lda * (k0 + l)];
  }
  __syncthreads();
  for (int kk = 0; kk < b; ++kk) {
    double bv = -1.0;
    int bi = m;
    for (int i = kk + tid; i < m; i += blockDim.x) {
      const cd x = W[i + (int64_t)n * kk];
      const double v = fabs(x.re) + fabs(x.im);
      if (v > bv) { bv = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { red_v[tid >> 6] = bv; red_i[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      double v = red_v[0];
      int i = red_i[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
        if (red_v[w] > v || (red_v[w] == v && red_i[w] < i)) { v = red_v[w]; i = red_i[w]; }
      if (!(v > 1e-300) || i >= m) {  // zero (or NaN) column: singular
        *flag = 1;
        i = kk;
      }
      s_p = i;
      P[k0 + kk] = k0 + i;
    }
    __syncthreads();
    const int p = s_p;
    if (tid < b) {
      const cd x = W[kk + (int64_t)n * tid], y = W[p + (int64_t)n * tid];
      if (p != kk) {
        W[kk + (int64_t)n * tid] = y;
        W[p + (int64_t)n * tid] = x;
      }
      prow[tid] = (p != kk) ? y : x;
    }
    __syncthreads();
    const cd piv = prow[kk];
    const double mag = piv.re * piv.re + piv.im * piv.im;
    if (mag > 0.0) {
      const double dr = piv.re / mag, di = -piv.im / mag;
      for (int i = kk + 1 + tid; i < m; i += blockDim.x) {
        const cd x = W[i + (int64_t)n * kk];
        const cd l = cd{x.re * dr - x.im * di, x.re * di + x.im * dr};
        W[i + (int64_t)n * kk] = l;
        for (int j = kk + 1; j < b; ++j) {
          const cd u = prow[j];
          cd& y = W[i + (int64_t)n * j];
          y = cd{y.re - (l.re * u.re - l.im * u.im), y.im - (l.re * u.im + l.im * u.re)};
        }
      }
    }
    __syncthreads();
  }
}

// Blocked LU with partial pivoting of the outer panel A[k0:n, K], |K| = b <= GJO,
// on a copy W (leading dimension n, n*GJO per matrix): 32-column sub-panels
// are factored in turn (gj_panel_lu_kernel), their exchanges and the U rows
// carried to the columns on their right (gj_panel_fix_kernel), and the Schur
// complement of those columns updated by a GEMM.  Only the exchanges are
// kept (ipiv[k0+kk], absolute rows, in sequence); they are those of zgetrf on
// the panel, so the in-order elimination of the exchanged block meets the
// partial pivots.
__global__ void gj_wcopy_kernel(const cd* __restrict__ a, int n, int64_t lda, int64_t astride, int k0, int b,
                                int nmat, cd* __restrict__ w) {
  const int m = n - k0;
  const int64_t per = (int64_t)m * b, tot = per * nmat;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int mat = (int)(e / per);
    const int64_t r = e % per;
    const int i = (int)(r % m), c = (int)(r / m);
    w[(int64_t)n * GJO * mat + i + (int64_t)n * c] = a[astride * mat + (k0 + i) + lda * (k0 + c)];
  }
}

// LU with row exchanges of the sub-panel W[r0:m, c0:c0+b), b <= GJB, in place;
// ipiv[k0+r0+kk] = k0 + (exchanged row of W).
__global__ __launch_bounds__(256) void gj_panel_lu_kernel(cd* __restrict__ w, int n, int m, int r0, int c0, int b,
                                                          int k0, int* __restrict__ ipiv, int* __restrict__ flag) {
  __shared__ double red_v[4];
  __shared__ int red_i[4];
  __shared__ cd prow[GJB];
  __shared__ int s_p;
  const int mat = blockIdx.x;
  cd* W = w + (int64_t)n * GJO * mat + r0 + (int64_t)n * c0;
  int* P = ipiv + (int64_t)n * mat;
  const int mm = m - r0;
  const int tid = threadIdx.x;
  for (int kk = 0; kk < b; ++kk) {
    double bv = -1.0;
    int bi = mm;
    for (int i = kk + tid; i < mm; i += blockDim.x) {
      const cd x = W[i + (int64_t)n * kk];
      const double v = fabs(x.re) + fabs(x.im);
      if (v > bv) { bv = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { red_v[tid >> 6] = bv; red_i[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      double v = red_v[0];
      int i = red_i[0];
      for (int wv = 1; wv < (int)(blockDim.x >> 6); ++wv)
        if (red_v[wv] > v || (red_v[wv] == v && red_i[wv] < i)) { v = red_v[wv]; i = red_i[wv]; }
      if (!(v > 1e-300) || i >= mm) {  // zero (or NaN) column: singular
        *flag = 1;
        i = kk;
      }
      s_p = i;
      P[k0 + r0 + kk] = k0 + r0 + i;
    }
    __syncthreads();
    const int p = s_p;
    if (tid < b) {
      const cd x = W[kk + (int64_t)n * tid], y = W[p + (int64_t)n * tid];
      if (p != kk) {
        W[kk + (int64_t)n * tid] = y;
        W[p + (int64_t)n * tid] = x;
      }
      prow[tid] = (p != kk) ? y : x;
    }
    __syncthreads();
    const cd piv = prow[kk];
    const double mag = piv.re * piv.re + piv.im * piv.im;
    if (mag > 0.0) {
      const double dr = piv.re / mag, di = -piv.im / mag;
      for (int i = kk + 1 + tid; i < mm; i += blockDim.x) {
        const cd x = W[i + (int64_t)n * kk];
        const cd l = cd{x.re * dr - x.im * di, x.re * di + x.im * dr};
        W[i + (int64_t)n * kk] = l;
        for (int j = kk + 1; j < b; ++j) {
          const cd u = prow[j];
          cd& y = W[i + (int64_t)n * j];
          y = cd{y.re - (l.re * u.re - l.im * u.im), y.im - (l.re * u.im + l.im * u.re)};
        }
      }
    }
    __syncthreads();
  }
}

// For the panel columns j in [c0+b, bw): the sub-panel's exchanges, then
// U12 = L11^-1 W[r0:r0+b, j] (L11 unit lower triangular).  One thread per (matrix, column).
__global__ void gj_panel_fix_kernel(cd* __restrict__ w, int n, int r0, int c0, int b, int bw, int k0,
                                    const int* __restrict__ ipiv, int nmat) {
  const int ncol = bw - c0 - b;
  const int64_t tot = (int64_t)ncol * nmat;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int mat = (int)(e / ncol), j = c0 + b + (int)(e % ncol);
    cd* Wm = w + (int64_t)n * GJO * mat;
    cd* col = Wm + (int64_t)n * j;
    const int* P = ipiv + (int64_t)n * mat;
    for (int kk = 0; kk < b; ++kk) {
      const int r = r0 + kk, p = P[k0 + r] - k0;
      if (p != r) {
        const cd x = col[r];
        col[r] = col[p];
        col[p] = x;
      }
    }
    const cd* L = Wm + r0 + (int64_t)n * c0;
    for (int i = 1; i < b; ++i) {
      double sr = col[r0 + i].re, si = col[r0 + i].im;
      for (int kk = 0; kk < i; ++kk) {
        const cd l = L[i + (int64_t)n * kk], u = col[r0 + kk];
        sr -= l.re * u.re - l.im * u.im;
        si -= l.re * u.im + l.im * u.re;
      }
      col[r0 + i] = cd{sr, si};
    }
  }
}

// Apply the exchanges of block K to whole rows (all n columns).
__global__ __launch_bounds__(256) void gj_swap_kernel(cd* __restrict__ a, int n, int64_t lda, int64_t astride,
                                                      int k0, int b, const int* __restrict__ ipiv) {
  const int mat = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  cd* col = a + astride * mat + lda * c;
  const int* P = ipiv + (int64_t)n * mat;
  for (int kk = 0; kk < b; ++kk) {
    const int r = k0 + kk, p = P[r];
    if (p != r) {
      const cd x = col[r];
      col[r] = col[p];
      col[p] = x;
    }
  }
}

// Column order of the in-place result: undoing the row exchanges as column
// exchanges in reverse order (j = n-1..0: swap columns j and ipiv[j]) is the
// gather inv(:, j) = a(:, perm[j]) with perm from the same swaps on indices.
__global__ __launch_bounds__(256) void gj_perm_kernel(const int* __restrict__ ipiv, int n, int* __restrict__ perm) {
  extern __shared__ int sh[];
  int* idx = sh;
  int* pv = sh + n;
  const int mat = blockIdx.x;
  const int* P = ipiv + (int64_t)n * mat;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    idx[j] = j;
    pv[j] = P[j];
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int j = n - 1; j >= 0; --j) {
      const int p = pv[j];
      if (p != j) {
        const int t = idx[j];
        idx[j] = idx[p];
        idx[p] = t;
      }
    }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) perm[(int64_t)n * mat + j] = idx[j];
}

// out(:, j) = a(:, perm[j]) per matrix; out has leading dimension n.
__global__ void gj_gather_kernel(const cd* __restrict__ a, int n, int64_t lda, int64_t astride,
                                 const int* __restrict__ perm, int nmat, cd* __restrict__ out) {
  const int64_t nn = (int64_t)n * n, tot = nn * nmat;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int mat = (int)(e / nn);
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    out[e] = a[astride * mat + i + lda * perm[(int64_t)n * mat + j]];
  }
}

__global__ void gj_copy_back_kernel(const cd* __restrict__ src, int n, int nmat, cd* __restrict__ a, int64_t lda,
                                    int64_t astride) {
  const int64_t nn = (int64_t)n * n, tot = nn * nmat;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int mat = (int)(e / nn);
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    a[astride * mat + i + lda * j] = src[e];
  }
}

ecalj_status gj_unscramble(ecalj_ctx* h, cd* a, int n, int64_t lda, int64_t astride, int nmat, const int* perm) {
  const int64_t tot = (int64_t)n * n * nmat;
  if (tot == 0) return ECALJ_OK;
  cd* tmp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_TMP2, sizeof(cd) * tot, (void**)&tmp));
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(gj_gather_kernel, dim3(blocks), dim3(256), 0, h->stream, a, n, lda, astride, perm, nmat, tmp);
  ECALJ_HIP(hipGetLastError());
  hipLaunchKernelGGL(gj_copy_back_kernel, dim3(blocks), dim3(256), 0, h->stream, tmp, n, nmat, a, lda, astride);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// One level: pivot blocks of GJB columns.
static ecalj_status gj_inverse_1(ecalj_ctx* h, cd* a, int n, int64_t lda, int64_t astride, int nmat, int* ipiv) {
  const bool pivot = ipiv != nullptr;
  cd *pinv, *cp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_GJP, sizeof(cd) * GJB * GJB * nmat, (void**)&pinv));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_GJC, sizeof(cd) * (size_t)n * GJB * nmat, (void**)&cp));
  for (int k0 = 0; k0 < n; k0 += GJB) {
    const int b = std::min(GJB, n - k0);
    if (pivot) {  // choose and apply the row exchanges of this block (cp is free until gj_diag)
      hipLaunchKernelGGL(gj_panel_kernel, dim3(nmat), dim3(256), 0, h->stream, a, n, lda, astride, k0, b, cp, ipiv,
                         h->d_flag);
      ECALJ_HIP(hipGetLastError());
      hipLaunchKernelGGL(gj_swap_kernel, dim3((n + 255) / 256, nmat), dim3(256), 0, h->stream, a, n, lda, astride,
                         k0, b, ipiv);
      ECALJ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(gj_diag_kernel, dim3(nmat), dim3(256), 0, h->stream, a, n, lda, astride, k0, b, pinv, cp,
                       h->d_flag);
    ECALJ_HIP(hipGetLastError());
    // (2) row panel: A[K,:] = Pinv (b x b) * A[K,:] (b x n), in place per column tile
    GemmCall r;
    r.M = b; r.N = n; r.K = b;
    r.A = pinv; r.lda = GJB; r.bA = {(int64_t)GJB * GJB, 0};
    r.B = a + k0; r.ldb = lda; r.bB = {astride, 0};
    r.C = a + k0; r.ldc = lda; r.bC = {astride, 0};
    r.batch = nmat; r.div = nmat;
    ECALJ_TRY(gemm(h, r));
    // (3) trailing update over all rows not in K
    GemmCall t;
    t.M = n; t.N = n; t.K = b;
    t.A = cp; t.lda = n; t.bA = {(int64_t)n * GJB, 0};
    t.B = a + k0; t.ldb = lda; t.bB = {astride, 0};
    t.C = a; t.ldc = lda; t.bC = {astride, 0};
    t.batch = nmat; t.div = nmat;
    t.epi = EPI_GJ; t.gj_k0 = k0; t.gj_k1 = k0 + b;
    ECALJ_TRY(gemm(h, t));
  }
  return ECALJ_OK;
}

// Outer step copies: D = A[K,K] (ld GJO), Cp = A[:,K] (ld n), R = A[K,:] with
// R[:,K] = I (ld GJO), for K = [k0, k0+b).
__global__ void gj_outer_copy_kernel(const cd* __restrict__ a, int n, int64_t lda, int64_t astride, int k0, int b,
                                     int nmat, cd* __restrict__ D, cd* __restrict__ Cp, cd* __restrict__ R) {
  const int64_t per = (int64_t)n * b, tot = per * nmat;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int mat = (int)(e / per);
    const int64_t r = e % per;
    const cd* A = a + astride * mat;
    {  // Cp(i, l), i over n
      const int i = (int)(r % n), l = (int)(r / n);
      const cd x = A[i + lda * (k0 + l)];
      Cp[(int64_t)n * b * mat + i + (int64_t)n * l] = x;
      if (i >= k0 && i < k0 + b) D[(int64_t)GJO * GJO * mat + (i - k0) + (int64_t)GJO * l] = x;
    }
    {  // R(l, j), j over n
      const int l = (int)(r % b), j = (int)(r / b);
      const cd x = (j >= k0 && j < k0 + b) ? cd{(j - k0) == l ? 1.0 : 0.0, 0.0} : A[(k0 + l) + lda * j];
      R[(int64_t)GJO * n * mat + l + (int64_t)GJO * j] = x;
    }
  }
}

static size_t round256(size_t b) { return ((b ? b : 256) + 255) & ~(size_t)255; }

// The slot requests of gj_inverse (pivot) + gj_unscramble, in their order.
size_t zminv_scratch_bytes(int64_t n, int64_t nmat) {
  const size_t z = sizeof(cd), m = (size_t)nmat, N = (size_t)n;
  size_t t = round256(sizeof(int) * 2 * N * m);                             // S_GJPIV
  if (n <= GJO) {
    t += round256(z * GJB * GJB * m) + round256(z * N * GJB * m);           // S_GJP, S_GJC
  } else {
    t += round256(z * GJO * GJO * m) + 3 * round256(z * N * GJO * m);       // S_GJD, S_GJOC, S_GJR, S_GJW
    t += round256(z * GJB * GJB * m) + round256(z * (size_t)GJO * GJB * m); // gj_inverse_1 of the diagonal blocks
  }
  return t + round256(z * N * N * m);                                       // S_TMP2 of gj_unscramble
}

ecalj_status gj_inverse(ecalj_ctx* h, cd* a, int n, int64_t lda, int64_t astride, int nmat, const int** perm) {
  if (n == 0 || nmat == 0) return ECALJ_OK;
  if (nmat > 65535) return ECALJ_ERR_DIM;
  const bool pivot = perm != nullptr;
  if (pivot && n > 8192) return ECALJ_ERR_DIM;  // index tables of gj_perm_kernel live in LDS
  int* ipiv = nullptr;
  if (pivot) ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_GJPIV, sizeof(int) * 2 * (size_t)n * nmat, (void**)&ipiv));
  // Two levels above GJO: outer blocks of GJO columns.  With pivoting, the
  // exchanges of an outer block come from a blocked LU of its panel (32-wide
  // sub-panels on a copy), then the exchanged block is eliminated in order.
  if (n <= GJO) {
    ECALJ_TRY(gj_inverse_1(h, a, n, lda, astride, nmat, ipiv));
  } else {
    cd *D, *Cp, *R, *W = nullptr;
    ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_GJD, sizeof(cd) * GJO * GJO * (size_t)nmat, (void**)&D));
    ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_GJOC, sizeof(cd) * (size_t)n * GJO * nmat, (void**)&Cp));
    ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_GJR, sizeof(cd) * (size_t)n * GJO * nmat, (void**)&R));
    if (pivot) ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_GJW, sizeof(cd) * (size_t)n * GJO * nmat, (void**)&W));
    for (int k0 = 0; k0 < n; k0 += GJO) {
      const int b = std::min(GJO, n - k0);
      const int64_t tot = (int64_t)n * b * nmat;
      if (pivot) {
        const int m = n - k0;
        const int64_t tw = (int64_t)m * b * nmat;
        hipLaunchKernelGGL(gj_wcopy_kernel, dim3((unsigned)std::min<int64_t>((tw + 255) / 256, 65536)), dim3(256), 0,
                           h->stream, a, n, lda, astride, k0, b, nmat, W);
        ECALJ_HIP(hipGetLastError());
        for (int c0 = 0; c0 < b; c0 += GJB) {
          const int bs = std::min(GJB, b - c0);
          hipLaunchKernelGGL(gj_panel_lu_kernel, dim3(nmat), dim3(256), 0, h->stream, W, n, m, c0, c0, bs, k0, ipiv,
                             h->d_flag);
          ECALJ_HIP(hipGetLastError());
          const int right = b - c0 - bs;
          if (right <= 0) continue;
          const int64_t tf = (int64_t)right * nmat;
          hipLaunchKernelGGL(gj_panel_fix_kernel, dim3((unsigned)std::min<int64_t>((tf + 255) / 256, 65536)),
                             dim3(256), 0, h->stream, W, n, c0, c0, bs, b, k0, ipiv, nmat);
          ECALJ_HIP(hipGetLastError());
          if (m - c0 - bs <= 0) continue;
          GemmCall u;  // W[c0+bs:m, c0+bs:b] -= W[c0+bs:m, c0:c0+bs] W[c0:c0+bs, c0+bs:b]
          u.M = m - c0 - bs; u.N = right; u.K = bs;
          u.A = W + (c0 + bs) + (int64_t)n * c0; u.lda = n; u.bA = {(int64_t)n * GJO, 0};
          u.B = W + c0 + (int64_t)n * (c0 + bs); u.ldb = n; u.bB = {(int64_t)n * GJO, 0};
          u.C = W + (c0 + bs) + (int64_t)n * (c0 + bs); u.ldc = n; u.bC = {(int64_t)n * GJO, 0};
          u.batch = nmat; u.div = nmat;
          u.alpha = cd{-1.0, 0.0}; u.beta = cd{1.0, 0.0};
          ECALJ_TRY(gemm(h, u));
        }
        hipLaunchKernelGGL(gj_swap_kernel, dim3((n + 255) / 256, nmat), dim3(256), 0, h->stream, a, n, lda, astride,
                           k0, b, ipiv);
        ECALJ_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(gj_outer_copy_kernel, dim3((unsigned)std::min<int64_t>((tot + 255) / 256, 65536)),
                         dim3(256), 0, h->stream, a, n, lda, astride, k0, b, nmat, D, Cp, R);
      ECALJ_HIP(hipGetLastError());
      // D <- D^-1 (in order: the matrices taking this path are safe to eliminate without exchanges)
      ECALJ_TRY(gj_inverse_1(h, D, b, GJO, (int64_t)GJO * GJO, nmat, nullptr));
      // A[K,:] = D^-1 R   (A[K,K] becomes D^-1)
      GemmCall r;
      r.M = b; r.N = n; r.K = b;
      r.A = D; r.lda = GJO; r.bA = {(int64_t)GJO * GJO, 0};
      r.B = R; r.ldb = GJO; r.bB = {(int64_t)GJO * n, 0};
      r.C = a + k0; r.ldc = lda; r.bC = {astride, 0};
      r.batch = nmat; r.div = nmat;
      ECALJ_TRY(gemm(h, r));
      // A[i,:] = (j in K ? 0 : A[i,:]) - Cp[i,:] A[K,:],  i not in K
      GemmCall t;
      t.M = n; t.N = n; t.K = b;
      t.A = Cp; t.lda = n; t.bA = {(int64_t)n * b, 0};
      t.B = a + k0; t.ldb = lda; t.bB = {astride, 0};
      t.C = a; t.ldc = lda; t.bC = {astride, 0};
      t.batch = nmat; t.div = nmat;
      t.epi = EPI_GJ; t.gj_k0 = k0; t.gj_k1 = k0 + b;
      ECALJ_TRY(gemm(h, t));
    }
  }
  if (pivot) {
    int* pm = ipiv + (size_t)n * nmat;
    hipLaunchKernelGGL(gj_perm_kernel, dim3(nmat), dim3(256), sizeof(int) * 2 * (size_t)n, h->stream, ipiv, n, pm);
    ECALJ_HIP(hipGetLastError());
    *perm = pm;
  }
  return ECALJ_OK;
}

// ==========================================================================
// Sigma_c
// Unpack the keepwv upper-triangle form and take the conjugate transpose in
// one pass:  A_v(I,J) = conj(wc_v(J,I)) = (I <= J) ? W_v(I,J) : conj(W_v(J,I)).
// W is complex128 (as the W build writes it); the output is in the contraction's type.
template <class T>
__global__ void sig_unpack_kernel(const cd* __restrict__ w, int64_t ldw, int64_t wstride, int n, int nw,
                                  T* __restrict__ out) {
  const int64_t nn = (int64_t)n * n, tot = nn * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int iw = (int)(e / nn);
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    const cd* W = w + wstride * iw;
    double xr, xi;
    if (i <= j) {
      xr = W[i + ldw * j].re;
      xi = W[i + ldw * j].im;
    } else {
      xr = W[j + ldw * i].re;
      xi = -W[j + ldw * i].im;
    }
    st_c(out + e, xr, xi);
  }
}

// B(kk, c) = wgt[v + (niw+1)*c] * zmel[J + n*c],  v = kk / n, J = kk % n
template <class T>
struct AccSigB {
  typedef typename ElemT<T>::V V;
  static constexpr bool kcontig = true;
  const T* zmel; const double* wgt; int n, nw1, X, K;
  __device__ void get(int x, int kk, V& re, V& im) const {
    re = 0;
    im = 0;
    if (x >= X || kk >= K) return;
    const int v = kk / n;
    const int J = kk - v * n;
    const V g = (V)wgt[v + (int64_t)nw1 * x];
    ld_c<V>(zmel, J + (int64_t)n * x, re, im);
    re *= g;
    im *= g;
  }
};

// czmelwc_b = Wh_b (n x nw1*n) * B_b ; one z-slice per batch (internal k)
template <int MODE, class T, int BM, int BN>
__global__ __launch_bounds__(kThreads) void sig_freq_gemm_kernel(const T* __restrict__ wh, int64_t wh_stride,
                                                                 const T* __restrict__ zmel, int64_t z_stride,
                                                                 const double* __restrict__ wgt, int64_t g_stride,
                                                                 int n, int nw1, int ncol, T* __restrict__ cz,
                                                                 int64_t cz_stride) {
  constexpr int BK = ModeBK<MODE>::value;
  typedef typename ElemT<T>::V V;
  __shared__ typename StageOf<MODE, BM, BN, BK>::type s;
  const int b = blockIdx.z;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int K = n * nw1;
  constexpr int KV = ModeKV<MODE>::value;
  Ld<AccN<T, false>, BM, BK, true, KV> la{AccN<T, false>{wh + wh_stride * b, (int64_t)n, n, K, (V)1}};
  Ld<AccSigB<T>, BN, BK, false, KV> lb{AccSigB<T>{zmel + z_stride * b, wgt + g_stride * b, n, nw1, ncol, K}};
  Core<MODE, BM, BN, BK, false> core;
  gemm_mainloop<MODE, BM, BN, BK, false>(la, lb, m0, n0, K, s, core);
  T* c = cz + cz_stride * b;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  core.for_each(wave >> 1, wave & 1, lane, [&](int rr, int cc, double xr, double xi) {
    const int i = m0 + rr, j = n0 + cc;
    if (i < n && j < ncol) st_c(c + i + (int64_t)n * j, xr, xi);
  });
}

// zsec(m,m') += the split-K partials of each internal k in turn (a fixed-order
// tree per k, FP64), with Im(zsec_mm) <= 0 after every k: the reference clips
// after every (k, band block) call (S/m_sxcf_gemm.f90:599-608).  The partials
// of k are [kstart[k], kstart[k+1]) when kstart is given, else
// [k*chunks, (k+1)*chunks).  One block per output element.
template <class T>
__global__ __launch_bounds__(256) void sig_final_reduce_kernel(const T* __restrict__ part, int nk, int chunks,
                                                               const int* __restrict__ kstart, int ntq,
                                                               cd* __restrict__ zsec, int64_t ldz) {
  __shared__ double sr[256], si[256];
  const int np = ntq * ntq;
  const int p = blockIdx.x;
  const int m = p % ntq, mp = p / ntq;
  cd* z = zsec + m + ldz * mp;
  const cd z0 = *z;
  double zr = z0.re, zi = z0.im;  // running value (thread 0's copy is the one written)
  for (int k = 0; k < nk; ++k) {
    const int q0 = kstart ? kstart[k] : k * chunks, q1 = kstart ? kstart[k + 1] : (k + 1) * chunks;
    double a = 0.0, c = 0.0;
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
      double xr, xi;
      ld_c2(part + (int64_t)np * q + p, xr, xi);
      a += xr;
      c += xi;
    }
    sr[threadIdx.x] = a;
    si[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        sr[threadIdx.x] += sr[threadIdx.x + w];
        si[threadIdx.x] += si[threadIdx.x + w];
      }
      __syncthreads();
    }
    zr += sr[0];
    zi += si[0];
    if (m == mp && zi > 0.0) zi = 0.0;
    __syncthreads();
  }
  if (threadIdx.x == 0) *z = cd{zr, zi};
}

// Split-K partials of the final product (S/m_sxcf_gemm.f90:599):
// part[chunk][b] (ntq x ntq) = cz_b(chunk rows)^H zmel_b(chunk rows) for gb
// internal k-points of R = n*nm rows each (per-k strides zs), chunks of KC rows.
constexpr int SIG_KC = 2048;
static inline int sig_chunks(int64_t R) { return (int)((R + SIG_KC - 1) / SIG_KC); }

template <int MODE, class T>
static ecalj_status sig_final_partials(ecalj_ctx* h, const T* cz, const T* zmel, int64_t R, int64_t zs, int gb,
                                       int ntq, T* part) {
  const int KC = SIG_KC;
  const int chunks = sig_chunks(R);
  const int np = ntq * ntq;
  GemmCall f;
  f.mode = MODE;
  f.ta = true; f.ca = true;
  f.M = ntq; f.N = ntq; f.K = KC;
  f.A = cz; f.lda = R; f.bA = {KC, zs};
  f.B = zmel; f.ldb = R; f.bB = {KC, zs};
  f.C = part; f.ldc = ntq; f.bC = {(int64_t)np, (int64_t)np * chunks};
  f.batch = chunks * gb; f.div = chunks;
  if (R % KC != 0) {
    // last chunk is shorter: run full chunks, then the tail separately
    f.batch = (chunks - 1) * gb;
    f.div = chunks - 1;
    if (chunks > 1) ECALJ_TRY(gemm(h, f));
    GemmCall t = f;
    t.K = (int)(R - (int64_t)KC * (chunks - 1));
    t.A = cz + (int64_t)KC * (chunks - 1);
    t.B = zmel + (int64_t)KC * (chunks - 1);
    t.C = part + (int64_t)np * (chunks - 1);
    t.bA = {0, zs}; t.bB = {0, zs}; t.bC = {0, (int64_t)np * chunks};
    t.batch = gb; t.div = 1;
    ECALJ_TRY(gemm(h, t));
  } else {
    ECALJ_TRY(gemm(h, f));
  }
  return ECALJ_OK;
}

template <int MODE, class T>
static ecalj_status sig_contract_t(ecalj_ctx* h, int nbatch, const T* zmel, int n, int nm, int ntq, int niw,
                                   const cd* wc, int64_t ldw, const double* wgt, cd* zsec, int64_t ldz,
                                   const RealSpec* rs) {
  // real-axis pole lists (host) for all k, applied per group below
  RealPlanHandle* rp = nullptr;
  if (rs) ECALJ_TRY(real_plan_make(h, *rs, nbatch, n, nm, ntq, &rp));
  struct Free {
    RealPlanHandle* p;
    ~Free() { if (p) real_plan_free(p); }
  } free_rp{rp};
  const int nw1 = niw + 1;
  const int ncol = nm * ntq;
  const int64_t nn = (int64_t)n * n;
  T *wh, *cz, *part;
  // W unpacked (or, for bf16x3, split into bf16 planes) per internal k;
  // processed in groups to bound scratch memory
  const size_t wh_bytes_per = (MODE == MBF3) ? sig_bf3_bytes_per_k(n, ncol, nw1) : sizeof(T) * nn * nw1;
  const size_t cz_bytes_per = sizeof(T) * (size_t)n * ncol;
  int group = (int)std::max<size_t>(1, std::min<size_t>(nbatch, ((size_t)8 << 30) / (wh_bytes_per + cz_bytes_per)));
  group = std::min(group, 65535);
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_WH, wh_bytes_per * group, (void**)&wh));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_CZ, cz_bytes_per * group, (void**)&cz));
  // final product: split-K partials of czmelwc^H zmel (K = n*nm per internal k)
  const int64_t R = (int64_t)n * nm;
  const int chunks = sig_chunks(R);
  const int np = ntq * ntq;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_PART, sizeof(T) * (size_t)np * chunks * nbatch, (void**)&part));
  const int64_t wstride = ldw * n * nw1;
  const int64_t tot = nn * nw1;
  const int ublocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  for (int b0 = 0; b0 < nbatch; b0 += group) {
    const int gb = std::min(group, nbatch - b0);
    if constexpr (MODE == MBF3) {
      unsigned short* as = reinterpret_cast<unsigned short*>(wh);
      unsigned short* bs = as + sig_bf3_a_elems_per_k(n, nw1) * gb;
      ECALJ_TRY((h->sigc_fp8 ? sig_f8 : sig_bf3)(h, gb, (const cf*)zmel + (int64_t)n * ncol * b0,
                                                 (int64_t)n * ncol, n, ncol, nw1, wc + wstride * b0, ldw, wstride,
                                                 wgt + (int64_t)nw1 * ncol * b0, (int64_t)nw1 * ncol, (cf*)cz,
                                                 (int64_t)n * ncol, as, bs));
    } else {
      for (int q = 0; q < gb; ++q) {
        hipLaunchKernelGGL((sig_unpack_kernel<T>), dim3(ublocks), dim3(256), 0, h->stream, wc + wstride * (b0 + q),
                           ldw, ldw * n, n, nw1, wh + tot * q);
        ECALJ_HIP(hipGetLastError());
      }
      int tev = -1;
      ECALJ_TRY(timing_begin(h, 0, &tev));
      dim3 grid((n + 63) / 64, (ncol + 63) / 64, gb);
      if (grid.y > 65535) return ECALJ_ERR_DIM;
      hipLaunchKernelGGL((sig_freq_gemm_kernel<MODE, T, 64, 64>), grid, dim3(kThreads), 0, h->stream, wh, tot,
                         zmel + (int64_t)n * ncol * b0, (int64_t)n * ncol, wgt + (int64_t)nw1 * ncol * b0,
                         (int64_t)nw1 * ncol, n, nw1, ncol, cz, (int64_t)n * ncol);
      ECALJ_HIP(hipGetLastError());
      ECALJ_TRY(timing_end(h, 0, tev));
    }
    if (rp)  // czmelwc += real-axis pole columns (S/m_sxcf_gemm.f90:586-589)
      ECALJ_TRY(real_plan_into_cz(h, rp, *rs, MODE == M64 ? M64 : M32, zmel, n, nm, ntq, b0, b0 + gb, cz));
    ECALJ_TRY((sig_final_partials<MODE, T>(h, cz, zmel + (int64_t)n * ncol * b0, R, (int64_t)n * ncol, gb, ntq,
                                            part + (int64_t)np * chunks * b0)));
  }
  hipLaunchKernelGGL((sig_final_reduce_kernel<T>), dim3(np), dim3(256), 0, h->stream, part, nbatch, chunks,
                     (const int*)nullptr, ntq, zsec, ldz);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// Internal k-points with their own middle-state counts nm[k] (host array):
// zmel_k (n x nm[k] x ntq) and wgt_k ((niw+1) x nm[k] x ntq) packed one after
// the other, W_c(k) at a uniform stride.  bf16x3: one grouped persistent
// launch per scratch group, split-K partials per k, one reduction for all k.
// Other forms: the uniform path per k.
ecalj_status sig_contract_grouped(ecalj_ctx* h, int mode, int nk, const int* nm, const void* zmel, int n, int ntq,
                                  int niw, const cd* wc, int64_t ldw, const double* wgt, cd* zsec, int64_t ldz) {
  const int nw1 = niw + 1;
  const int64_t wstride = ldw * n * nw1;
  std::vector<int64_t> zoff(nk + 1, 0);
  for (int k = 0; k < nk; ++k) {
    if (nm[k] < 0) return ECALJ_ERR_DIM;
    zoff[k + 1] = zoff[k] + (int64_t)nm[k] * ntq;  // columns before k
  }
  if (mode != MBF3) {
    const size_t es = (mode == M64) ? sizeof(cd) : sizeof(cf);
    for (int k = 0; k < nk; ++k) {
      if (nm[k] == 0) continue;
      ECALJ_TRY(sig_contract(h, mode, 1, (const char*)zmel + es * (size_t)n * zoff[k], n, nm[k], ntq, niw,
                             wc + wstride * k, ldw, wgt + nw1 * zoff[k], zsec, ldz));
    }
    return ECALJ_OK;
  }
  const cf* z = (const cf*)zmel;
  const int np = ntq * ntq;
  // scratch groups of consecutive k under 8 GB
  std::vector<int> gstart{0};
  size_t acc = 0;
  size_t maxg = 0;
  for (int k = 0; k < nk; ++k) {
    const size_t b = sig_bf3_bytes_per_k(n, std::max(nm[k] * ntq, 1), nw1) + sizeof(cf) * (size_t)n * nm[k] * ntq;
    if (acc > 0 && acc + b > ((size_t)8 << 30)) {
      gstart.push_back(k);
      maxg = std::max(maxg, acc);
      acc = 0;
    }
    acc += b;
  }
  maxg = std::max(maxg, acc);
  gstart.push_back(nk);
  int64_t nparts = 0;
  for (int k = 0; k < nk; ++k) nparts += nm[k] ? sig_chunks((int64_t)n * nm[k]) : 0;
  unsigned char* wh;
  cf* part;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_WH, maxg, (void**)&wh));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_PART, sizeof(cf) * (size_t)np * std::max<int64_t>(nparts, 1),
                        (void**)&part));
  int64_t pofs = 0;
  for (size_t gi = 0; gi + 1 < gstart.size(); ++gi) {
    const int k0 = gstart[gi], k1 = gstart[gi + 1];
    std::vector<int> ncol(k1 - k0);
    size_t a_elems = sig_bf3_a_elems_per_k(n, nw1) * (k1 - k0);
    size_t b_bytes = 0;
    for (int k = k0; k < k1; ++k) {
      ncol[k - k0] = nm[k] * ntq;
      b_bytes += sig_bf3_bytes_per_k(n, std::max(ncol[k - k0], 1), nw1) - 2 * sig_bf3_a_elems_per_k(n, nw1);
    }
    unsigned short* as = reinterpret_cast<unsigned short*>(wh);
    unsigned short* bs = as + a_elems;
    cf* cz = reinterpret_cast<cf*>(reinterpret_cast<unsigned char*>(bs) + b_bytes);
    ECALJ_TRY((h->sigc_fp8 ? sig_f8_grouped : sig_bf3_grouped)(h, k1 - k0, ncol.data(), z + (int64_t)n * zoff[k0],
                                                               n, nw1, wc + wstride * k0, ldw, wstride,
                                                               wgt + nw1 * zoff[k0], cz, as, bs));
    int64_t cofs = 0;
    for (int k = k0; k < k1; ++k) {
      if (nm[k] == 0) continue;
      const int64_t R = (int64_t)n * nm[k];
      ECALJ_TRY((sig_final_partials<MBF3, cf>(h, cz + cofs, z + (int64_t)n * zoff[k], R, 0, 1, ntq,
                                              part + (int64_t)np * pofs)));
      cofs += (int64_t)n * nm[k] * ntq;
      pofs += sig_chunks(R);
    }
  }
  std::vector<unsigned char> ks(sizeof(int) * (nk + 1));
  {
    int* kp = reinterpret_cast<int*>(ks.data());
    kp[0] = 0;
    for (int k = 0; k < nk; ++k) kp[k + 1] = kp[k] + (nm[k] ? sig_chunks((int64_t)n * nm[k]) : 0);
  }
  unsigned char* dks;
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_KST, ks, &dks));
  hipLaunchKernelGGL((sig_final_reduce_kernel<cf>), dim3(np), dim3(256), 0, h->stream, part, nk, 0,
                     (const int*)dks, ntq, zsec, ldz);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status sig_contract(ecalj_ctx* h, int mode, int nbatch, const void* zmel, int n, int nm, int ntq, int niw,
                          const cd* wc, int64_t ldw, const double* wgt, cd* zsec, int64_t ldz, const RealSpec* rs) {
  if (n == 0 || nm * ntq == 0 || nbatch == 0) return ECALJ_OK;
  if (mode == M64)
    return sig_contract_t<M64, cd>(h, nbatch, (const cd*)zmel, n, nm, ntq, niw, wc, ldw, wgt, zsec, ldz, rs);
  if (mode == M32)
    return sig_contract_t<M32, cf>(h, nbatch, (const cf*)zmel, n, nm, ntq, niw, wc, ldw, wgt, zsec, ldz, rs);
  return sig_contract_t<MBF3, cf>(h, nbatch, (const cf*)zmel, n, nm, ntq, niw, wc, ldw, wgt, zsec, ldz, rs);
}

// --------------------------------------------------------------------------
// Imaginary-axis weights (FP64), one thread per (batch, it, itp).
__global__ void sig_weights_kernel(int nbatch, int niw, int nm, int ntq, int ns1, int nctot,
                                   const double* __restrict__ omega, const double* __restrict__ ekc,
                                   double esmr, double ua, const double* __restrict__ wkkr,
                                   const double* __restrict__ x, const double* __restrict__ wt,
                                   const double* __restrict__ expa, double* __restrict__ wgt) {
  const int64_t tot = (int64_t)nbatch * nm * ntq;
  const double pi = 3.14159265358979323846;
  const double rmax = 2.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / ((int64_t)nm * ntq));
    const int64_t r = e % ((int64_t)nm * ntq);
    const int itl = (int)(r % nm);     // 0-based middle state
    const int itp = (int)(r / nm);
    const int it = ns1 + itl;          // Fortran index
    const double we = 0.5 * (omega[(int64_t)ntq * b + itp] - ekc[(int64_t)nm * b + itl]);
    const double aw = fabs(ua * we);
    const double aw2 = aw * aw;
    const double sgn = signbit(we) ? -1.0 : 1.0;
    const double sig = 0.5 * esmr;
    const double sig2 = 2.0 * sig * sig;
    const double wk = wkkr[b];
    double* out = wgt + (int64_t)(niw + 1) * e;
    if (it <= nctot) {
      double s = 0.0;
      for (int i = 0; i < niw; ++i) {
        const double cons = 1.0 / (we * we * x[i] * x[i] + (1.0 - x[i]) * (1.0 - x[i]));
        const double w1 = we * cons * wt[i] * (-1.0 / pi);
        out[1 + i] = wk * w1;
        s += w1 * expa[i];
      }
      const double w0 = (fabs(we) < rmax / ua) ? (-s - 0.5 * sgn * exp(we * we * ua * ua) * erfc(ua * fabs(we))) : 0.0;
      out[0] = wk * w0;
    } else {
      double s = 0.0;
      for (int i = 0; i < niw; ++i) {
        const double omd = 1.0 / x[i] - 1.0;
        const double omd2w = omd * omd + we * we;
        const double rr = omd2w / sig2;
        double cons;
        if (rr > 5e-3) {
          cons = (1.0 - exp(-rr)) / omd2w;
        } else {
          const double s2 = sig2, s3 = s2 * s2 * s2, s4 = s3 * s2, s5 = s4 * s2, s6 = s5 * s2;
          cons = 1.0 / s2 - omd2w / (s2 * s2) / 2.0 + omd2w * omd2w / s3 / 6.0 - omd2w * omd2w * omd2w / s4 / 24.0 +
                 omd2w * omd2w * omd2w * omd2w / s5 / 120.0 - omd2w * omd2w * omd2w * omd2w * omd2w / s6 / 720.0;
        }
        out[1 + i] = wk * (-we * cons * wt[i] / (x[i] * x[i]) / pi);
        s += cons * expa[i] * wt[i] / (x[i] * x[i]);
      }
      out[0] = wk * (we * s / pi + sgn * 0.5 * exp(aw2) * (erfc(sqrt(aw2 + we * we / sig2)) - erfc(aw)));
    }
  }
}

ecalj_status sig_weights(ecalj_ctx* h, int nbatch, int niw, int nm, int ntq, int ns1, int nctot,
                         const double* omega, const double* ekc, double esmr, double ua, const double* wkkr,
                         const double* x, const double* wt, const double* expa, double* wgt) {
  const int64_t tot = (int64_t)nbatch * nm * ntq;
  if (tot == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(sig_weights_kernel, dim3(blocks), dim3(256), 0, h->stream, nbatch, niw, nm, ntq, ns1, nctot,
                     omega, ekc, esmr, ua, wkkr, x, wt, expa, wgt);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// --------------------------------------------------------------------------
// Device utilities
__global__ void zero_kernel(uint4* __restrict__ p, int64_t n16, unsigned char* __restrict__ tail, int ntail) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n16; e += stride) p[e] = uint4{0, 0, 0, 0};
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

ecalj_status dev_zero(ecalj_ctx* h, void* p, size_t bytes) {
  if (bytes == 0) return ECALJ_OK;
  if (reinterpret_cast<uintptr_t>(p) % 16 != 0) return ECALJ_ERR_ARG;
  const int64_t n16 = bytes / 16;
  const int ntail = (int)(bytes % 16);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n16 + 255) / 256, 16384));
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, h->stream, (uint4*)p, n16,
                     (unsigned char*)p + n16 * 16, ntail);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// dst = src, 16-byte vectors (both 16-byte aligned, bytes a multiple of 16)
__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n16; e += stride) dst[e] = src[e];
}

ecalj_status dev_copy(ecalj_ctx* h, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return ECALJ_OK;
  if (reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16 || bytes % 16)
    return ECALJ_ERR_ARG;
  const int64_t n16 = bytes / 16;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n16 + 255) / 256, 16384));
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, h->stream, (const uint4*)src, (uint4*)dst, n16);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// out[k][m](I,J) = d_k(I) * a[m](I,J) * conj(d_k(J))
__global__ void phase_sim_kernel(const cd* __restrict__ a, int n, int nmat, const cd* __restrict__ d, int nk,
                                 cd* __restrict__ out) {
  const int64_t nn = (int64_t)n * n, per = nn * nmat, tot = per * nk;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(e / per);
    const int64_t r = e % per;
    const int64_t ij = r % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    const cd x = a[r], di = d[(int64_t)n * k + i], dj = d[(int64_t)n * k + j];
    // di * x
    const double tr = di.re * x.re - di.im * x.im, ti = di.re * x.im + di.im * x.re;
    // * conj(dj)
    out[e] = cd{tr * dj.re + ti * dj.im, ti * dj.re - tr * dj.im};
  }
}

ecalj_status phase_similarity(ecalj_ctx* h, const cd* a, int n, int nmat, const cd* d, int nk, cd* out) {
  const int64_t tot = (int64_t)n * n * nmat * nk;
  if (tot == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(phase_sim_kernel, dim3(blocks), dim3(256), 0, h->stream, a, n, nmat, d, nk, out);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

}  // namespace ecalj
