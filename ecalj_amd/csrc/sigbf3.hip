// sigbf3.hip -- Sigma_c frequency contraction in the bf16x3 form.
//
//   czmelwc(I,c) = sum_v wgt(v,c) * sum_J Wt_v(I,J) zmel(J,c),
//   Wt_v(I,J) = (I <= J) ? W_v(I,J) : conj(W_v(J,I))   (keepwv upper-triangle
//   convention, S/m_sxcf_gemm.f90:468-475, after the conjugate transpose of :492)
//
// The weight depends only on (v, c), so it is applied to the per-frequency
// partial sums in FP32 instead of to every B element: zmel is split into
// bf16 hi/lo planes once, W_v once, both zero-padded (rows to MP, k to NP,
// columns to NCP), and the main loop only moves 16-byte chunks into LDS and
// issues MFMAs.  Plane p: 0 = re hi, 1 = re lo, 2 = im hi, 3 = im lo.
#include <algorithm>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {

// ---------------------------------------------------------------- pre-split
// As[b][((p*nw1 + v)*MP + I)*NP + J] from W (complex128, ldw, per-batch stride w_stride)
constexpr int SPT = 32;
__global__ __launch_bounds__(256) void split_w_kernel(const cd* __restrict__ w, int64_t ldw, int64_t w_stride,
                                                      int n, int nw1, int MP, int NP,
                                                      unsigned short* __restrict__ as, int64_t as_stride) {
  __shared__ cd up[SPT][SPT + 1];  // up[jj][ii] = W(I0+ii, J0+jj)
  __shared__ cd lo[SPT][SPT + 1];  // lo[ii][jj] = W(J0+jj, I0+ii)
  const int J0 = blockIdx.x * SPT, I0 = blockIdx.y * SPT;
  const int v = blockIdx.z % nw1, b = blockIdx.z / nw1;
  const cd* W = w + w_stride * b + (int64_t)ldw * n * v;
  const int tx = threadIdx.x % SPT, ty = threadIdx.x / SPT;  // 32 x 8
  for (int r = ty; r < SPT; r += 8) {
    const int I = I0 + tx, J = J0 + r;
    up[r][tx] = (I < n && J < n) ? W[I + ldw * J] : cd{0.0, 0.0};
    const int J2 = J0 + tx, I2 = I0 + r;
    lo[r][tx] = (I2 < n && J2 < n) ? W[J2 + ldw * I2] : cd{0.0, 0.0};
  }
  __syncthreads();
  unsigned short* A = as + as_stride * b;
  const int64_t pstride = (int64_t)nw1 * MP * NP;
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  if (threadIdx.x < SPT * (SPT / 8)) {
    const int r = threadIdx.x / (SPT / 8), jg = 8 * (threadIdx.x % (SPT / 8));
    const int I = I0 + r;
    us8v rh, rl, ih, il;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int jj = jg + e, J = J0 + jj;
      double xr = 0.0, xi = 0.0;
      if (I < n && J < n) {
        if (I <= J) {
          xr = up[jj][r].re;
          xi = up[jj][r].im;
        } else {
          xr = lo[r][jj].re;
          xi = -lo[r][jj].im;
        }
      }
      unsigned short h, l;
      split_bf16((float)xr, h, l);
      rh[e] = h;
      rl[e] = l;
      split_bf16((float)xi, h, l);
      ih[e] = h;
      il[e] = l;
    }
    const int64_t o = ((int64_t)v * MP + I) * NP + J0 + jg;
    *reinterpret_cast<us8v*>(A + o) = rh;
    *reinterpret_cast<us8v*>(A + o + pstride) = rl;
    *reinterpret_cast<us8v*>(A + o + 2 * pstride) = ih;
    *reinterpret_cast<us8v*>(A + o + 3 * pstride) = il;
  }
}

// Bs[b][(p*NCP + c)*NP + J] from zmel (complex64, n x ncol per batch)
__global__ void split_z_kernel(const cf* __restrict__ z, int64_t z_stride, int n, int ncol, int NCP, int NP,
                               unsigned short* __restrict__ bs, int64_t bs_stride, int nb) {
  const int64_t per = (int64_t)NCP * NP, tot = per * nb;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e % per;
    const int c = (int)(r / NP), J = (int)(r % NP);
    float xr = 0.f, xi = 0.f;
    if (c < ncol && J < n) {
      const cf x = z[z_stride * b + J + (int64_t)n * c];
      xr = x.re;
      xi = x.im;
    }
    unsigned short* B = bs + bs_stride * b;
    unsigned short h, l;
    split_bf16(xr, h, l);
    B[(0 * (int64_t)NCP + c) * NP + J] = h;
    B[(1 * (int64_t)NCP + c) * NP + J] = l;
    split_bf16(xi, h, l);
    B[(2 * (int64_t)NCP + c) * NP + J] = h;
    B[(3 * (int64_t)NCP + c) * NP + J] = l;
  }
}

// ---------------------------------------------------------------- main kernel
constexpr int SBM = 128, SBN = 64, SBK = 64;

__global__ __launch_bounds__(kThreads, 1) void sig_bf3_kernel(const unsigned short* __restrict__ as,
                                                              int64_t as_stride,
                                                              const unsigned short* __restrict__ bs,
                                                              int64_t bs_stride, const double* __restrict__ wgt,
                                                              int64_t g_stride, int n, int nw1, int ncol, int MP,
                                                              int NP, int NCP, cf* __restrict__ cz,
                                                              int64_t cz_stride) {
  typedef StageBF3<SBM, SBN, SBK> S;
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  __shared__ S s;
  const int b = blockIdx.z;
  const int I0 = blockIdx.x * SBM, c0 = blockIdx.y * SBN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, wm = wave >> 1, wn = wave & 1;
  const unsigned short* A = as + as_stride * b;
  const unsigned short* B = bs + bs_stride * b;
  const int64_t apl = (int64_t)nw1 * MP * NP;  // A plane stride
  const int64_t bpl = (int64_t)NCP * NP;       // B plane stride
  const int NS = NP / SBK;                     // k-stages per frequency
  const int nst = NS * nw1;

  Core<MBF3, SBM, SBN, SBK, false> core;
  constexpr int TM = SBM / 32, TN = SBN / 32;
  f4 tr[TM][TN], ti[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      tr[i][j] = f4{0, 0, 0, 0};
      ti[i][j] = f4{0, 0, 0, 0};
    }
  core.zero();

  // register staging of one stage: A and B chunks of 16 bytes (8 bf16 of one row)
  constexpr int PARTS = SBK / 8;                        // 16-byte chunks per row and plane
  constexpr int NA = 4 * SBM * PARTS / kThreads;        // A chunks per thread
  constexpr int NB = 4 * SBN * PARTS / kThreads;        // B chunks per thread
  constexpr int APL = SBM * PARTS, BPL = SBN * PARTS;   // chunks per plane
  us8v ra[NA], rb[NB];
  auto fetch = [&](int st) {
    const int v = st / NS, ks = st - v * NS;
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int cid = t + kThreads * q;
      const int p = cid / APL, rem = cid % APL, row = rem / PARTS, part = rem % PARTS;
      const unsigned short* src = A + p * apl + ((int64_t)v * MP + I0 + row) * NP + ks * SBK + 8 * part;
      ra[q] = *reinterpret_cast<const us8v*>(src);
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int cid = t + kThreads * q;
      const int p = cid / BPL, rem = cid % BPL, col = rem / PARTS, part = rem % PARTS;
      const unsigned short* src = B + p * bpl + (int64_t)(c0 + col) * NP + ks * SBK + 8 * part;
      rb[q] = *reinterpret_cast<const us8v*>(src);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int cid = t + kThreads * q;
      const int p = cid / APL, rem = cid % APL, row = rem / PARTS, part = rem % PARTS;
      unsigned short* dst = (p == 0 ? s.arh : p == 1 ? s.arl : p == 2 ? s.aih : s.ail) + row * S::LDK + 8 * part;
      *reinterpret_cast<us8v*>(dst) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int cid = t + kThreads * q;
      const int p = cid / BPL, rem = cid % BPL, col = rem / PARTS, part = rem % PARTS;
      unsigned short* dst = (p == 0 ? s.brh : p == 1 ? s.brl : p == 2 ? s.bih : s.bil) + col * S::LDK + 8 * part;
      *reinterpret_cast<us8v*>(dst) = rb[q];
    }
  };

  const double* g = wgt + g_stride * b;
  fetch(0);
  for (int st = 0; st < nst; ++st) {
    __syncthreads();
    store();
    __syncthreads();
    if (st + 1 < nst) fetch(st + 1);
    core.compute(s, wm, wn, lane);
    const int v = st / NS;
    if (st - v * NS == NS - 1) {  // end of frequency v: tot += wgt(v,c) * acc
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = c0 + wn * TN * 16 + j * 16 + (lane & 15);
        const float wv = (c < ncol) ? (float)g[v + (int64_t)nw1 * c] : 0.f;
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          tr[i][j] += wv * core.cr[i][j];
          ti[i][j] += wv * core.ci[i][j];
        }
      }
      core.zero();
    }
  }
  cf* C = cz + cz_stride * b;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int I = I0 + wm * TM * 16 + i * 16 + acc_row_f32(lane, e);
        const int c = c0 + wn * TN * 16 + j * 16 + (lane & 15);
        if (I < n && c < ncol) C[I + (int64_t)n * c] = cf{tr[i][j][e], ti[i][j][e]};
      }
}

// ---------------------------------------------------------------- launcher
// Computes czmelwc for nb internal k-points; returns ECALJ_OK or a status.
ecalj_status sig_bf3(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, int nw1,
                     const cd* wc, int64_t ldw, int64_t w_stride, const double* wgt, int64_t g_stride, cf* cz,
                     int64_t cz_stride, unsigned short* as, unsigned short* bs) {
  const int MP = (n + SBM - 1) / SBM * SBM;
  const int NP = (n + SBK - 1) / SBK * SBK;
  const int NCP = (ncol + SBN - 1) / SBN * SBN;
  const int64_t as_stride = 4LL * nw1 * MP * NP;
  const int64_t bs_stride = 4LL * NCP * NP;
  // W planes (rows beyond n come out zero: the split kernel covers all of MP x NP)
  dim3 gw(NP / SPT, MP / SPT, nw1 * nb);
  if (gw.z > 65535) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL(split_w_kernel, gw, dim3(256), 0, h->stream, wc, ldw, w_stride, n, nw1, MP, NP, as, as_stride);
  ECALJ_HIP(hipGetLastError());
  const int64_t totz = (int64_t)NCP * NP * nb;
  const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
  hipLaunchKernelGGL(split_z_kernel, dim3(bz), dim3(256), 0, h->stream, zmel, z_stride, n, ncol, NCP, NP, bs,
                     bs_stride, nb);
  ECALJ_HIP(hipGetLastError());
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 0, &tev));
  dim3 grid(MP / SBM, NCP / SBN, nb);
  if (grid.y > 65535 || grid.z > 65535) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL(sig_bf3_kernel, grid, dim3(kThreads), 0, h->stream, as, as_stride, bs, bs_stride, wgt, g_stride,
                     n, nw1, ncol, MP, NP, NCP, cz, cz_stride);
  ECALJ_HIP(hipGetLastError());
  ECALJ_TRY(timing_end(h, 0, tev));
  return ECALJ_OK;
}

size_t sig_bf3_a_elems_per_k(int n, int nw1) {
  const size_t MP = (n + SBM - 1) / SBM * SBM, NP = (n + SBK - 1) / SBK * SBK;
  return 4 * (size_t)nw1 * MP * NP;
}

size_t sig_bf3_bytes_per_k(int n, int ncol, int nw1) {
  const size_t MP = (n + SBM - 1) / SBM * SBM, NP = (n + SBK - 1) / SBK * SBK, NCP = (ncol + SBN - 1) / SBN * SBN;
  return 2 * 4 * ((size_t)nw1 * MP * NP + NCP * NP);
}

}  // namespace ecalj
