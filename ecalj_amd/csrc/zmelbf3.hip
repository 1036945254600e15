// zmelbf3.hip -- the two zmel GEMMs (S/m_zmel.f90:323-415) in the bf16x3 form on the block tile of sigtile.hpp.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "sigtile.hpp"

namespace ecalj {

// ================================================================ zmel contraction on the bf16x3 core
// The second GEMM of zmel (S/m_zmel.f90:323-415 in the occupied- or end-side
// order): per batch b, C_b(I, c) = s(alpha sum_k T_b(k, I) X_b(k, c)), k
// contiguous in both operands (T stored T[k + lda I], X stored X[k + ldb c]),
// s = conjugation when conjC.  Both operands are split once into the
// pre-tiled planes (alpha and the conjugations folded in), then the Sigma_c
// tile code runs per (row tile, column tile, batch) with unit weights.

// rows of a K-contiguous complex64 source -> planes; one thread per (b, ks, row, chunk), chunk fastest
__global__ __launch_bounds__(256) void zsplit_kc_kernel(const cf* __restrict__ src, int64_t ld, int64_t s0,
                                                        int64_t s1, int div, int rows, int K, float ar, float ai,
                                                        int conj, int RP, int NS, unsigned short* __restrict__ dst,
                                                        int64_t dstride, int nb) {
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  const int64_t per = (int64_t)NS * RP * 4, tot = per * nb;
  const int64_t pst = (int64_t)RP * QBK;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(e % 4);
    const int row = (int)((e / 4) % RP);
    const int ks = (int)((e / (4 * (int64_t)RP)) % NS);
    const int b = (int)(e / per);
    const cf* base = src + (int64_t)(b % div) * s0 + (int64_t)(b / div) * s1 + ld * row;
    us8v rh, rl, ih, il;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kk = ks * QBK + 8 * ch + q;
      float xr = 0.f, xi = 0.f;
      if (row < rows && kk < K) {
        const cf x = base[kk];
        xr = ar * x.re - ai * x.im;
        xi = ar * x.im + ai * x.re;
        if (conj) xi = -xi;
      }
      unsigned short h, l;
      split_bf16(xr, h, l);
      rh[q] = h;
      rl[q] = l;
      split_bf16(xi, h, l);
      ih[q] = h;
      il[q] = l;
    }
    unsigned short* D = dst + dstride * b + (int64_t)ks * 4 * pst + (int64_t)row * QBK + 8 * chunk_slot(row, ch);
    *reinterpret_cast<us8v*>(D) = rh;
    *reinterpret_cast<us8v*>(D + pst) = rl;
    *reinterpret_cast<us8v*>(D + 2 * pst) = ih;
    *reinterpret_cast<us8v*>(D + 3 * pst) = il;
  }
}

static __global__ __launch_bounds__(256) void ones_kernel(double* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 1.0;
}

__global__ __launch_bounds__(512, 1) void zmel_bf3_kernel(const unsigned short* __restrict__ as, int64_t ape,
                                                          const unsigned short* __restrict__ bs, int64_t bpe,
                                                          int bdiv, const double* __restrict__ ones, int M, int N,
                                                          int MP, int NS, int NCP, cf* __restrict__ C, int64_t ldc,
                                                          int64_t c_s0, int64_t c_s1, int div) {
  __shared__ SigSmem<128, 2> s;
  const int b = blockIdx.z, r = blockIdx.x, c = blockIdx.y;
  const unsigned short* Bk = bs + (bdiv ? bpe * (b / bdiv) : 0);
  cf* Cb = C + (int64_t)(b % div) * c_s0 + (int64_t)(b / div) * c_s1;
  sig_tile<128, 2, 8>(s, as + ape * b, Bk, ones, M, 1, 0, 1, N, MP, NS, NCP, r * QBM, c * 128, 0, Cb, 0, ldc);
}

ecalj_status zmel_z_bf3(ecalj_ctx* h, int M, int N, int K, const cf* A, int64_t lda, int64_t a_s0, int64_t a_s1,
                        const cf* B, int64_t ldb, int64_t b_s1, bool conjB, cd alpha, bool conjC, cf* C,
                        int64_t ldc, int64_t c_s0, int64_t c_s1, int batch, int div) {
  if (M < 1 || N < 1 || batch < 1) return ECALJ_OK;
  if (K < 1 || div < 1 || batch > 65535) return ECALJ_ERR_DIM;
  const int MP = pad_to(M, QBM), NCP = pad_to(N, 128), NS = pad_to(K, QBK) / QBK;
  const int nbb = b_s1 ? (batch + div - 1) / div : 1;  // distinct B operands
  const int64_t ape = 4LL * NS * MP * QBK, bpe = 4LL * NS * NCP * QBK;
  unsigned short *as, *bs;
  double* ones;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_ZA, sizeof(unsigned short) * (size_t)ape * batch, (void**)&as));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_ZB, sizeof(unsigned short) * (size_t)bpe * nbb + sizeof(double) * NCP,
                        (void**)&bs));
  ones = reinterpret_cast<double*>(bs + (size_t)bpe * nbb);
  hipLaunchKernelGGL(ones_kernel, dim3((NCP + 255) / 256), dim3(256), 0, h->stream, ones, NCP);
  ECALJ_HIP(hipGetLastError());
  // conj(alpha T) on A; on B its own conjugation combined with the output's
  {
    const int64_t tot = (int64_t)NS * MP * 4 * batch;
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536 * 4);
    hipLaunchKernelGGL(zsplit_kc_kernel, dim3(blocks), dim3(256), 0, h->stream, A, lda, a_s0, a_s1, div, M, K,
                       (float)alpha.re, (float)alpha.im, conjC ? 1 : 0, MP, NS, as, ape, batch);
    ECALJ_HIP(hipGetLastError());
  }
  {
    const int64_t tot = (int64_t)NS * NCP * 4 * nbb;
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536 * 4);
    hipLaunchKernelGGL(zsplit_kc_kernel, dim3(blocks), dim3(256), 0, h->stream, B, ldb, (int64_t)0, b_s1, 1, N, K,
                       1.f, 0.f, (conjB != conjC) ? 1 : 0, NCP, NS, bs, bpe, nbb);
    ECALJ_HIP(hipGetLastError());
  }
  dim3 grid(MP / QBM, NCP / 128, batch);
  hipLaunchKernelGGL(zmel_bf3_kernel, grid, dim3(512), 0, h->stream, as, ape, bs, bpe, b_s1 ? div : 0, ones, M, N,
                     MP, NS, NCP, C, ldc, c_s0, c_s1, div);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}


// ================================================================ zmel first contraction (real ppb)
// T(r, c) = sum_k P(r, k) opB(X)(k, c): P real (ppb blocks, P[r + lda k]),
// X complex64 K-contiguous (X[k + ldb c], conjugated when conjB), per batch b
// X_b = X + b x_s1, T_b = C + b c_s1; P shared by all batches.

// real rows -> planes (imaginary planes zero); one thread per (ks, chunk, row), row fastest
__global__ __launch_bounds__(256) void rsplit_kernel(const double* __restrict__ a, int64_t lda, int rows, int K,
                                                     int RP, int NS, unsigned short* __restrict__ dst) {
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  const int64_t tot = (int64_t)NS * 4 * RP;
  const int64_t pst = (int64_t)RP * QBK;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(e % RP);
    const int ch = (int)((e / RP) % 4), ks = (int)(e / (4 * (int64_t)RP));
    us8v rh, rl, z;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kk = ks * QBK + 8 * ch + q;
      const float x = (row < rows && kk < K) ? (float)a[row + lda * kk] : 0.f;
      unsigned short h, l;
      split_bf16(x, h, l);
      rh[q] = h;
      rl[q] = l;
      z[q] = 0;
    }
    unsigned short* D = dst + (int64_t)ks * 4 * pst + (int64_t)row * QBK + 8 * chunk_slot(row, ch);
    *reinterpret_cast<us8v*>(D) = rh;
    *reinterpret_cast<us8v*>(D + pst) = rl;
    *reinterpret_cast<us8v*>(D + 2 * pst) = z;
    *reinterpret_cast<us8v*>(D + 3 * pst) = z;
  }
}

__global__ __launch_bounds__(512, 1) void zmel_p_kernel(const unsigned short* __restrict__ as,
                                                        const unsigned short* __restrict__ bs, int64_t bpe,
                                                        const double* __restrict__ ones, int M, int N, int MP, int NS,
                                                        int NCP, cf* __restrict__ C, int64_t ldc, int64_t c_s1) {
  __shared__ SigSmem<128, 2> s;
  const int b = blockIdx.z, r = blockIdx.x, c = blockIdx.y;
  sig_tile<128, 2, 8, true>(s, as, bs + bpe * b, ones, M, 1, 0, 1, N, MP, NS, NCP, r * QBM, c * 128, 0,
                            C + c_s1 * b, 0, ldc);
}

ecalj_status zmel_p_bf3(ecalj_ctx* h, int M, int N, int K, const double* P, int64_t lda, const cf* X,
                        int64_t ldb, int64_t x_s1, bool conjB, cf* C, int64_t ldc, int64_t c_s1, int batch) {
  if (M < 1 || N < 1 || batch < 1) return ECALJ_OK;
  if (K < 1 || batch > 65535) return ECALJ_ERR_DIM;
  const int MP = pad_to(M, QBM), NCP = pad_to(N, 128), NS = pad_to(K, QBK) / QBK;
  if ((int64_t)MP / QBM > 0x7fffffff) return ECALJ_ERR_DIM;
  const int64_t ape = 4LL * NS * MP * QBK, bpe = 4LL * NS * NCP * QBK;
  unsigned short *as, *bs;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_ZPA, sizeof(unsigned short) * (size_t)ape, (void**)&as));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_ZPB, sizeof(unsigned short) * (size_t)bpe * batch + sizeof(double) * NCP,
                        (void**)&bs));
  double* ones = reinterpret_cast<double*>(bs + (size_t)bpe * batch);
  hipLaunchKernelGGL(ones_kernel, dim3((NCP + 255) / 256), dim3(256), 0, h->stream, ones, NCP);
  ECALJ_HIP(hipGetLastError());
  {
    const int64_t tot = (int64_t)NS * 4 * MP;
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536 * 4);
    hipLaunchKernelGGL(rsplit_kernel, dim3(blocks), dim3(256), 0, h->stream, P, lda, M, K, MP, NS, as);
    ECALJ_HIP(hipGetLastError());
  }
  {
    const int64_t tot = (int64_t)NS * NCP * 4 * batch;
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536 * 4);
    hipLaunchKernelGGL(zsplit_kc_kernel, dim3(blocks), dim3(256), 0, h->stream, X, ldb, (int64_t)0, x_s1, 1, N, K,
                       1.f, 0.f, conjB ? 1 : 0, NCP, NS, bs, bpe, batch);
    ECALJ_HIP(hipGetLastError());
  }
  dim3 grid(MP / QBM, NCP / 128, batch);
  hipLaunchKernelGGL(zmel_p_kernel, grid, dim3(512), 0, h->stream, as, bs, bpe, ones, M, N, MP, NS, NCP, C, ldc,
                     c_s1);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

}  // namespace ecalj
