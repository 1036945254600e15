// gemm.hpp -- complex GEMM engine on CDNA4 MFMA (gfx950), FP64 path.
//
// One workgroup = 256 threads = 4 wavefronts in a 2x2 grid; each wave owns a
// (BM/2) x (BN/2) block of the output made of 16x16 MFMA tiles.  A and B
// stages go global -> registers -> LDS (re and im in separate k-major
// planes), with the next stage's global loads issued before the current
// stage's MFMAs.  Complex products use four real MFMAs
// (re += ar*br - ai*bi, im += ar*bi + ai*br); a real A operand uses two.
//
// Operand loaders and epilogues are small structs so that the gathers and
// scalings of the ecalj kernels (weighted band-pair columns for chi0,
// frequency-weighted zmel for Sigma_c, the epsilon and Gauss-Jordan updates)
// fold into the GEMM instead of running as separate passes.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace ecalj {

constexpr int kThreads = 256;

template <int BM, int BN, int BK>
struct StageF64 {
  static constexpr int LDA = BM + 4;
  static constexpr int LDB = BN + 4;
  double ar[BK * LDA];
  double ai[BK * LDA];
  double br[BK * LDB];
  double bi[BK * LDB];
};

// --------------------------------------------------------------------------
// Operand loaders.  Each owns a fixed thread -> element mapping chosen so
// that consecutive lanes touch consecutive addresses of the contiguous
// dimension.  fetch() reads one BK-deep stage into registers (zero outside
// the matrix); store() writes them to the LDS planes.

// A(i,k) = a[i + lda*k]  (column-major, rows contiguous)
template <int BM, int BK, bool REAL>
struct LdA_N {
  static constexpr int R = BM * BK / kThreads;
  static constexpr int KSTEP = kThreads / BM;
  const void* a; int64_t lda; int M, K; double cs;
  struct Regs { double re[R]; double im[R]; };
  __device__ void fetch(int m0, int k0, Regs& r) const {
    const int i = m0 + (threadIdx.x % BM);
    const int kb = k0 + threadIdx.x / BM;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = kb + q * KSTEP;
      const bool ok = (i < M) && (k < K);
      if (REAL) {
        r.re[q] = ok ? static_cast<const double*>(a)[i + lda * k] : 0.0;
        r.im[q] = 0.0;
      } else {
        cd v = ok ? static_cast<const cd*>(a)[i + lda * k] : cd{0.0, 0.0};
        r.re[q] = v.re;
        r.im[q] = cs * v.im;
      }
    }
  }
  template <class S>
  __device__ void store(S& s, const Regs& r) const {
    const int i = threadIdx.x % BM;
    const int kb = threadIdx.x / BM;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      s.ar[(kb + q * KSTEP) * S::LDA + i] = r.re[q];
      s.ai[(kb + q * KSTEP) * S::LDA + i] = r.im[q];
    }
  }
};

// A(i,k) = a[k + lda*i]  (op(A) = A^T or A^H of a column-major matrix)
template <int BM, int BK, bool REAL>
struct LdA_T {
  static constexpr int R = BM * BK / kThreads;
  static constexpr int ISTEP = kThreads / BK;
  const void* a; int64_t lda; int M, K; double cs;
  struct Regs { double re[R]; double im[R]; };
  __device__ void fetch(int m0, int k0, Regs& r) const {
    const int k = k0 + (threadIdx.x % BK);
    const int ib = m0 + threadIdx.x / BK;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = ib + q * ISTEP;
      const bool ok = (i < M) && (k < K);
      if (REAL) {
        r.re[q] = ok ? static_cast<const double*>(a)[k + lda * i] : 0.0;
        r.im[q] = 0.0;
      } else {
        cd v = ok ? static_cast<const cd*>(a)[k + lda * i] : cd{0.0, 0.0};
        r.re[q] = v.re;
        r.im[q] = cs * v.im;
      }
    }
  }
  template <class S>
  __device__ void store(S& s, const Regs& r) const {
    const int k = threadIdx.x % BK;
    const int ib = threadIdx.x / BK;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      s.ar[k * S::LDA + ib + q * ISTEP] = r.re[q];
      s.ai[k * S::LDA + ib + q * ISTEP] = r.im[q];
    }
  }
};

// B(k,j) = b[k + ldb*j]  (column-major, k contiguous)
template <int BN, int BK>
struct LdB_N {
  static constexpr int R = BN * BK / kThreads;
  static constexpr int JSTEP = kThreads / BK;
  const cd* b; int64_t ldb; int N, K; double cs;
  struct Regs { double re[R]; double im[R]; };
  __device__ void fetch(int n0, int k0, Regs& r) const {
    const int k = k0 + (threadIdx.x % BK);
    const int jb = n0 + threadIdx.x / BK;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int j = jb + q * JSTEP;
      cd v = (j < N && k < K) ? b[k + ldb * j] : cd{0.0, 0.0};
      r.re[q] = v.re;
      r.im[q] = cs * v.im;
    }
  }
  template <class S>
  __device__ void store(S& s, const Regs& r) const {
    const int k = threadIdx.x % BK;
    const int jb = threadIdx.x / BK;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      s.br[k * S::LDB + jb + q * JSTEP] = r.re[q];
      s.bi[k * S::LDB + jb + q * JSTEP] = r.im[q];
    }
  }
};

// B(k,j) = b[j + ldb*k]  (op(B) = B^T or B^H, j contiguous)
template <int BN, int BK>
struct LdB_T {
  static constexpr int R = BN * BK / kThreads;
  static constexpr int KSTEP = kThreads / BN;
  const cd* b; int64_t ldb; int N, K; double cs;
  struct Regs { double re[R]; double im[R]; };
  __device__ void fetch(int n0, int k0, Regs& r) const {
    const int j = n0 + (threadIdx.x % BN);
    const int kb = k0 + threadIdx.x / BN;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = kb + q * KSTEP;
      cd v = (j < N && k < K) ? b[j + ldb * k] : cd{0.0, 0.0};
      r.re[q] = v.re;
      r.im[q] = cs * v.im;
    }
  }
  template <class S>
  __device__ void store(S& s, const Regs& r) const {
    const int j = threadIdx.x % BN;
    const int kb = threadIdx.x / BN;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      s.br[(kb + q * KSTEP) * S::LDB + j] = r.re[q];
      s.bi[(kb + q * KSTEP) * S::LDB + j] = r.im[q];
    }
  }
};

// --------------------------------------------------------------------------
// MFMA core: wave (wm, wn) of the 2x2 grid accumulates TM x TN tiles.
template <int BM, int BN, int BK, bool AREAL>
struct CoreF64 {
  static constexpr int TM = BM / 32;
  static constexpr int TN = BN / 32;
  d4 cr[TM][TN];
  d4 ci[TM][TN];

  __device__ void zero() {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        cr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
        ci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      }
  }

  template <class S>
  __device__ void compute(const S& s, int wm, int wn, int lane) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int k = kk + (lane >> 4);
      double ar[TM], ai[TM], an[TM], br[TN], bi[TN];
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        const int i = wm * TM * 16 + t * 16 + (lane & 15);
        ar[t] = s.ar[k * S::LDA + i];
        if (!AREAL) {
          ai[t] = s.ai[k * S::LDA + i];
          an[t] = -ai[t];
        }
      }
#pragma unroll
      for (int t = 0; t < TN; ++t) {
        const int j = wn * TN * 16 + t * 16 + (lane & 15);
        br[t] = s.br[k * S::LDB + j];
        bi[t] = s.bi[k * S::LDB + j];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], ci[i][j], 0, 0, 0);
          if (!AREAL) {
            cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(an[i], bi[j], cr[i][j], 0, 0, 0);
            ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], ci[i][j], 0, 0, 0);
          }
        }
    }
  }

  // Visit every accumulator element with its (row, col) inside the block tile.
  template <class F>
  __device__ void for_each(int wm, int wn, int lane, F&& f) const {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = wm * TM * 16 + i * 16 + acc_row_f64(lane, v);
          const int c = wn * TN * 16 + j * 16 + acc_col(lane, v);
          f(r, c, cr[i][j][v], ci[i][j][v]);
        }
  }
};

// Main loop shared by every GEMM-shaped kernel.
template <int BM, int BN, int BK, bool AREAL, class LA, class LB>
__device__ inline void gemm_mainloop(const LA& la, const LB& lb, int m0, int n0, int K,
                                     StageF64<BM, BN, BK>& s, CoreF64<BM, BN, BK, AREAL>& core) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  typename LA::Regs ra;
  typename LB::Regs rb;
  core.zero();
  if (K <= 0) return;
  la.fetch(m0, 0, ra);
  lb.fetch(n0, 0, rb);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    la.store(s, ra);
    lb.store(s, rb);
    __syncthreads();
    if (k0 + BK < K) {
      la.fetch(m0, k0 + BK, ra);
      lb.fetch(n0, k0 + BK, rb);
    }
    core.compute(s, wm, wn, lane);
  }
}

// --------------------------------------------------------------------------
// Generic batched GEMM: C = alpha * op(A) op(B) + beta * C, with epilogue
// variants.  op(A) is N (AT=false) or T/C (AT=true, conj via conjA).

struct GemmArgs {
  int M, N, K, batch, div;
  const void* A; int64_t lda; Batch2 bA; double csA;
  const cd* B; int64_t ldb; Batch2 bB; double csB;
  cd* C; int64_t ldc; Batch2 bC;
  double alr, ali, ber, bei;
  int conjC;
  int epi;
  const double* svec;  // EPI_EPS: eps = delta - alpha * s_i * acc * s_j
  int gj_k0, gj_k1;    // EPI_GJ: rows [k0,k1) skipped; cols [k0,k1) start from 0
};

template <int BM, int BN, bool AT, bool AREAL, bool BT>
__global__ __launch_bounds__(kThreads) void zgemm_f64_kernel(GemmArgs g) {
  constexpr int BK = 16;
  __shared__ StageF64<BM, BN, BK> s;
  const int b = blockIdx.z;
  const int m0 = blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  const char* Ab = static_cast<const char*>(g.A) + g.bA.off(b, g.div) * (AREAL ? 8 : 16);
  const cd* Bb = g.B + g.bB.off(b, g.div);
  cd* Cb = g.C + g.bC.off(b, g.div);
  CoreF64<BM, BN, BK, AREAL> core;
  using LA = typename std::conditional<AT, LdA_T<BM, BK, AREAL>, LdA_N<BM, BK, AREAL>>::type;
  using LB = typename std::conditional<BT, LdB_T<BN, BK>, LdB_N<BN, BK>>::type;
  LA la{Ab, g.lda, g.M, g.K, g.csA};
  LB lb{Bb, g.ldb, g.N, g.K, g.csB};
  gemm_mainloop<BM, BN, BK, AREAL>(la, lb, m0, n0, g.K, s, core);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  core.for_each(wave >> 1, wave & 1, lane, [&](int r, int c, double xr, double xi) {
    const int i = m0 + r, j = n0 + c;
    if (i >= g.M || j >= g.N) return;
    cd* p = Cb + i + g.ldc * j;
    double vr = g.alr * xr - g.ali * xi;
    double vi = g.alr * xi + g.ali * xr;
    if (g.epi == EPI_STORE) {
      if (g.ber != 0.0 || g.bei != 0.0) {
        cd o = *p;
        vr += g.ber * o.re - g.bei * o.im;
        vi += g.ber * o.im + g.bei * o.re;
      }
      if (g.conjC) vi = -vi;
      *p = cd{vr, vi};
    } else if (g.epi == EPI_EPS) {
      const double sc = g.svec[i] * g.svec[j];
      *p = cd{(i == j ? 1.0 : 0.0) - sc * vr, -sc * vi};
    } else {  // EPI_GJ
      if (i >= g.gj_k0 && i < g.gj_k1) return;
      cd o = (j >= g.gj_k0 && j < g.gj_k1) ? cd{0.0, 0.0} : *p;
      *p = cd{o.re - vr, o.im - vi};
    }
  });
}

}  // namespace ecalj
