// gemm.hpp -- complex GEMM engine on CDNA4 MFMA (gfx950).
//
// One workgroup = 256 threads = 4 wavefronts in a 2x2 grid; each wave owns a
// (BM/2) x (BN/2) block of the output made of 16x16 MFMA tiles.  A and B
// stages go global -> registers -> LDS, with the next stage's global loads
// issued before the current stage's MFMAs.
//
// Three arithmetic modes share the loaders and epilogues:
//   M64  : complex128 data, FP64 MFMA 16x16x4, four real MFMAs per complex
//          product (re += ar*br - ai*bi, im += ar*bi + ai*br);
//   M32  : complex64 data, FP32 MFMA 16x16x4, same scheme, FP32 accumulation;
//   MBF3 : complex64 data split into bf16 hi + lo parts when written to LDS;
//          each real product uses hi*hi + hi*lo + lo*hi (the lo*lo term,
//          2^-16 relative, is dropped) on bf16 MFMA 16x16x32 with FP32
//          accumulation: 12 MFMAs per complex 16x16x32 block.
// A real A operand (ppb) skips the imaginary products.
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
enum Mode { M64 = 0, M32 = 1, MBF3 = 2 };

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------- elements
template <class T> struct ElemT;
template <> struct ElemT<cd> { typedef double V; };
template <> struct ElemT<cf> { typedef float V; };

template <class V, class T>
__device__ inline void ld_c(const T* p, int64_t idx, V& re, V& im) {
  const T v = p[idx];
  re = (V)v.re;
  im = (V)v.im;
}

// round-to-nearest-even float -> bf16 bits (finite inputs)
__device__ inline unsigned short bf16_bits(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_float(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }
__device__ inline void split_bf16(float x, unsigned short& hi, unsigned short& lo) {
  hi = bf16_bits(x);
  lo = bf16_bits(x - bf16_float(hi));
}

// ---------------------------------------------------------------- LDS stages
// Planar (M64, M32): separate re/im planes, k-major: plane[k*LD + m].
template <int BM, int BN, int BK, class V>
struct StagePlanar {
  static constexpr int LDA = BM + 4;
  static constexpr int LDB = BN + 4;
  V ar[BK * LDA];
  V ai[BK * LDA];
  V br[BK * LDB];
  V bi[BK * LDB];
  __device__ void putA(int i, int k, V re, V im) {
    ar[k * LDA + i] = re;
    ai[k * LDA + i] = im;
  }
  __device__ void putB(int k, int j, V re, V im) {
    br[k * LDB + j] = re;
    bi[k * LDB + j] = im;
  }
};

// bf16 hi/lo planes, row-major with k contiguous: plane[m*LDK + k].
template <int BM, int BN, int BK>
struct StageBF3 {
  static constexpr int LDK = BK + 8;
  alignas(16) unsigned short arh[BM * LDK];
  alignas(16) unsigned short arl[BM * LDK];
  alignas(16) unsigned short aih[BM * LDK];
  alignas(16) unsigned short ail[BM * LDK];
  alignas(16) unsigned short brh[BN * LDK];
  alignas(16) unsigned short brl[BN * LDK];
  alignas(16) unsigned short bih[BN * LDK];
  alignas(16) unsigned short bil[BN * LDK];
  __device__ void putA(int i, int k, float re, float im) {
    unsigned short h, l;
    split_bf16(re, h, l);
    arh[i * LDK + k] = h;
    arl[i * LDK + k] = l;
    split_bf16(im, h, l);
    aih[i * LDK + k] = h;
    ail[i * LDK + k] = l;
  }
  __device__ void putB(int k, int j, float re, float im) {
    unsigned short h, l;
    split_bf16(re, h, l);
    brh[j * LDK + k] = h;
    brl[j * LDK + k] = l;
    split_bf16(im, h, l);
    bih[j * LDK + k] = h;
    bil[j * LDK + k] = l;
  }
  // 8 consecutive k of row x (A) or column x (B), written as 16-byte vectors
  template <bool IS_A, class V>
  __device__ void put8(int x, int k, const V* re, const V* im) {
    typedef unsigned short us8v __attribute__((ext_vector_type(8)));
    us8v rh, rl, ih, il;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      unsigned short h, l;
      split_bf16((float)re[e], h, l);
      rh[e] = h;
      rl[e] = l;
      split_bf16((float)im[e], h, l);
      ih[e] = h;
      il[e] = l;
    }
    const int o = x * LDK + k;
    if (IS_A) {
      *reinterpret_cast<us8v*>(arh + o) = rh;
      *reinterpret_cast<us8v*>(arl + o) = rl;
      *reinterpret_cast<us8v*>(aih + o) = ih;
      *reinterpret_cast<us8v*>(ail + o) = il;
    } else {
      *reinterpret_cast<us8v*>(brh + o) = rh;
      *reinterpret_cast<us8v*>(brl + o) = rl;
      *reinterpret_cast<us8v*>(bih + o) = ih;
      *reinterpret_cast<us8v*>(bil + o) = il;
    }
  }
};

template <int MODE, int BM, int BN, int BK> struct StageOf;
template <int BM, int BN, int BK> struct StageOf<M64, BM, BN, BK> { typedef StagePlanar<BM, BN, BK, double> type; };
template <int BM, int BN, int BK> struct StageOf<M32, BM, BN, BK> { typedef StagePlanar<BM, BN, BK, float> type; };
template <int BM, int BN, int BK> struct StageOf<MBF3, BM, BN, BK> { typedef StageBF3<BM, BN, BK> type; };

template <int MODE> struct ModeBK { static constexpr int value = (MODE == MBF3 || MODE == M64) ? 32 : 16; };

// --------------------------------------------------------------------------
// Operand loaders.  An accessor gives element (x, k) of an operand -- x is
// the row of A or the column of B -- already conjugated / weighted / gathered
// as the kernel needs, zero outside the matrix, and says which index is
// contiguous in memory.  The generic loader Ld maps threads to elements so
// that consecutive lanes touch consecutive addresses, reads one BK-deep stage
// into registers (fetch) and hands it to the LDS stage (store).  With KV = 8
// (bf16 stages) every thread owns runs of 8 consecutive k of one row, which
// the stage splits and writes as 16-byte vectors.

// A(x,k) = a[x + ld*k]  (x contiguous)
template <class T, bool REAL>
struct AccN {
  typedef typename ElemT<T>::V V;
  static constexpr bool kcontig = false;
  const void* a; int64_t ld; int X, K; V cs;
  __device__ void get(int x, int k, V& re, V& im) const {
    re = 0;
    im = 0;
    if (x >= X || k >= K) return;
    if (REAL) {
      re = (V) static_cast<const double*>(a)[x + ld * k];
    } else {
      ld_c<V>(static_cast<const T*>(a), x + ld * k, re, im);
      im *= cs;
    }
  }
};

// A(x,k) = a[k + ld*x]  (k contiguous)
template <class T, bool REAL>
struct AccT {
  typedef typename ElemT<T>::V V;
  static constexpr bool kcontig = true;
  const void* a; int64_t ld; int X, K; V cs;
  __device__ void get(int x, int k, V& re, V& im) const {
    re = 0;
    im = 0;
    if (x >= X || k >= K) return;
    if (REAL) {
      re = (V) static_cast<const double*>(a)[k + ld * x];
    } else {
      ld_c<V>(static_cast<const T*>(a), k + ld * x, re, im);
      im *= cs;
    }
  }
};

template <class Acc, int BX, int BK, bool IS_A, int KV>
struct Ld {
  typedef typename Acc::V V;
  static constexpr int R = BX * BK / kThreads;
  static_assert(R % KV == 0, "stage too small for the k-vector mapping");
  Acc acc;
  struct Regs { V re[R]; V im[R]; };

  __device__ static void coords(int q, int& x, int& k) {
    const int t = threadIdx.x;
    if (KV == 1) {
      if (!Acc::kcontig) {
        x = t % BX;
        k = t / BX + q * (kThreads / BX);
      } else {
        k = t % BK;
        x = t / BK + q * (kThreads / BK);
      }
    } else {
      const int g = q / KV, e = q % KV;
      if (!Acc::kcontig) {
        x = t % BX;
        k = KV * (t / BX + g * (kThreads / BX)) + e;
      } else {
        constexpr int NKV = BK / KV;
        x = t / NKV + g * (kThreads / NKV);
        k = KV * (t % NKV) + e;
      }
    }
  }
  __device__ void fetch(int x0, int k0, Regs& r) const {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int x, k;
      coords(q, x, k);
      acc.get(x0 + x, k0 + k, r.re[q], r.im[q]);
    }
  }
  template <class S>
  __device__ void store(S& s, const Regs& r) const {
    if constexpr (KV == 1) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        int x, k;
        coords(q, x, k);
        if (IS_A)
          s.putA(x, k, r.re[q], r.im[q]);
        else
          s.putB(k, x, r.re[q], r.im[q]);
      }
    } else {
#pragma unroll
      for (int g = 0; g < R / KV; ++g) {
        int x, k;
        coords(g * KV, x, k);
        s.template put8<IS_A>(x, k, r.re + g * KV, r.im + g * KV);
      }
    }
  }
};

template <int MODE> struct ModeKV { static constexpr int value = (MODE == MBF3) ? 8 : 1; };

// --------------------------------------------------------------------------
// MFMA cores: wave (wm, wn) of the 2x2 grid accumulates TM x TN tiles.
template <int MODE, int BM, int BN, int BK, bool AREAL> struct Core;

// Planar cores (FP64 and FP32 16x16x4)
template <int BM, int BN, int BK, bool AREAL, class V, class ACC>
struct CorePlanar {
  static constexpr int TM = BM / 32;
  static constexpr int TN = BN / 32;
  ACC cr[TM][TN];
  ACC ci[TM][TN];

  __device__ void zero() {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        cr[i][j] = ACC{0, 0, 0, 0};
        ci[i][j] = ACC{0, 0, 0, 0};
      }
  }

  __device__ static ACC mfma(V a, V b, ACC c) {
    if constexpr (std::is_same<V, double>::value)
      return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    else
      return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }

  template <class S>
  __device__ void compute(const S& s, int wm, int wn, int lane) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int k = kk + (lane >> 4);
      V ar[TM], ai[TM], an[TM], br[TN], bi[TN];
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
          cr[i][j] = mfma(ar[i], br[j], cr[i][j]);
          ci[i][j] = mfma(ar[i], bi[j], ci[i][j]);
          if (!AREAL) {
            cr[i][j] = mfma(an[i], bi[j], cr[i][j]);
            ci[i][j] = mfma(ai[i], br[j], ci[i][j]);
          }
        }
    }
  }

  template <class F>
  __device__ void for_each(int wm, int wn, int lane, F&& f) const {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = std::is_same<V, double>::value ? acc_row_f64(lane, v) : acc_row_f32(lane, v);
          const int r = wm * TM * 16 + i * 16 + rr;
          const int c = wn * TN * 16 + j * 16 + acc_col(lane, v);
          f(r, c, (double)cr[i][j][v], (double)ci[i][j][v]);
        }
  }
};

template <int BM, int BN, int BK, bool AREAL>
struct Core<M64, BM, BN, BK, AREAL> : CorePlanar<BM, BN, BK, AREAL, double, d4> {};
template <int BM, int BN, int BK, bool AREAL>
struct Core<M32, BM, BN, BK, AREAL> : CorePlanar<BM, BN, BK, AREAL, float, f4> {};

// bf16x3 core (bf16 MFMA 16x16x32, FP32 accumulation)
template <int BM, int BN, int BK, bool AREAL>
struct Core<MBF3, BM, BN, BK, AREAL> {
  static constexpr int TM = BM / 32;
  static constexpr int TN = BN / 32;
  f4 cr[TM][TN];
  f4 ci[TM][TN];

  __device__ void zero() {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        cr[i][j] = f4{0, 0, 0, 0};
        ci[i][j] = f4{0, 0, 0, 0};
      }
  }

  __device__ static f4 mfma(const us8& a, const us8& b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0,
                                                   0);
  }
  __device__ static us8 ld8(const unsigned short* p) { return *reinterpret_cast<const us8*>(p); }
  __device__ static us8 neg8(const us8& x) { return x ^ (unsigned short)0x8000; }

  template <class S>
  __device__ void compute(const S& s, int wm, int wn, int lane) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      const int ko = kk + 8 * (lane >> 4);
      us8 brh[TN], brl[TN], bih[TN], bil[TN];
      us8 arh[TM], arl[TM], aih[TM], ail[TM], nih[TM], nil[TM];
#pragma unroll
      for (int t = 0; t < TN; ++t) {
        const int j = wn * TN * 16 + t * 16 + (lane & 15);
        brh[t] = ld8(s.brh + j * S::LDK + ko);
        brl[t] = ld8(s.brl + j * S::LDK + ko);
        bih[t] = ld8(s.bih + j * S::LDK + ko);
        bil[t] = ld8(s.bil + j * S::LDK + ko);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int ir = wm * TM * 16 + i * 16 + (lane & 15);
        arh[i] = ld8(s.arh + ir * S::LDK + ko);
        arl[i] = ld8(s.arl + ir * S::LDK + ko);
        if (!AREAL) {
          aih[i] = ld8(s.aih + ir * S::LDK + ko);
          ail[i] = ld8(s.ail + ir * S::LDK + ko);
          nih[i] = neg8(aih[i]);
          nil[i] = neg8(ail[i]);
        }
      }
      // each product term sweeps all TM x TN tiles: consecutive MFMAs write
      // different accumulators
#define ECALJ_BF3_TERM(ACC, AF, BF)                   \
  _Pragma("unroll") for (int i = 0; i < TM; ++i)      \
  _Pragma("unroll") for (int j = 0; j < TN; ++j)      \
    ACC[i][j] = mfma(AF[i], BF[j], ACC[i][j]);
      ECALJ_BF3_TERM(cr, arh, brh)
      ECALJ_BF3_TERM(ci, arh, bih)
      ECALJ_BF3_TERM(cr, arh, brl)
      ECALJ_BF3_TERM(ci, arh, bil)
      ECALJ_BF3_TERM(cr, arl, brh)
      ECALJ_BF3_TERM(ci, arl, bih)
      if (!AREAL) {
        ECALJ_BF3_TERM(cr, nih, bih)
        ECALJ_BF3_TERM(ci, aih, brh)
        ECALJ_BF3_TERM(cr, nih, bil)
        ECALJ_BF3_TERM(ci, aih, brl)
        ECALJ_BF3_TERM(cr, nil, bih)
        ECALJ_BF3_TERM(ci, ail, brh)
      }
#undef ECALJ_BF3_TERM
    }
  }

  template <class F>
  __device__ void for_each(int wm, int wn, int lane, F&& f) const {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = wm * TM * 16 + i * 16 + acc_row_f32(lane, v);
          const int c = wn * TN * 16 + j * 16 + acc_col(lane, v);
          f(r, c, (double)cr[i][j][v], (double)ci[i][j][v]);
        }
  }
};

// Main loop shared by every GEMM-shaped kernel.
template <int MODE, int BM, int BN, int BK, bool AREAL, class LA, class LB>
__device__ inline void gemm_mainloop(const LA& la, const LB& lb, int m0, int n0, int K,
                                     typename StageOf<MODE, BM, BN, BK>::type& s,
                                     Core<MODE, BM, BN, BK, AREAL>& core) {
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

// Store / load an output element as complex T.
__device__ inline void st_c(cd* p, double re, double im) { *p = cd{re, im}; }
__device__ inline void st_c(cf* p, double re, double im) { *p = cf{(float)re, (float)im}; }
__device__ inline void ld_c2(const cd* p, double& re, double& im) { re = p->re; im = p->im; }
__device__ inline void ld_c2(const cf* p, double& re, double& im) { re = p->re; im = p->im; }

// --------------------------------------------------------------------------
// Generic batched GEMM: C = alpha * op(A) op(B) + beta * C, with epilogue
// variants.  op(A) is N (AT=false) or T (AT=true); conjugation via csA/csB.
struct GemmArgs {
  int M, N, K, batch, div;
  const void* A; int64_t lda; Batch2 bA; double csA;
  const void* B; int64_t ldb; Batch2 bB; double csB;
  void* C; int64_t ldc; Batch2 bC;
  double alr, ali, ber, bei;
  int conjC;
  int epi;
  const double* svec;  // EPI_EPS: eps = delta - alpha * s_i * acc * s_j
  int gj_k0, gj_k1;    // EPI_GJ: rows [k0,k1) skipped; cols [k0,k1) start from 0
  // XCD order (xt > 0): a 1-D grid whose block p the dispatcher deals to XCD
  // p % 8; XCD x's j-th block takes tile j % xt (xg row tiles per column
  // tile) of matrix (j / xt) * 8 + x, so every tile of one matrix runs on one
  // XCD and its A and B panels are fetched into one L2 instead of all eight
  int xg, xt;
};

template <int MODE, class T, int BM, int BN, bool AT, bool AREAL, bool BT>
__global__ __launch_bounds__(kThreads) void zgemm_kernel(GemmArgs g) {
  constexpr int BK = ModeBK<MODE>::value;
  typedef typename StageOf<MODE, BM, BN, BK>::type Stage;
  typedef typename ElemT<T>::V V;
  __shared__ Stage s;
  int b, mt, nt;
  if (g.xt == 0) {
    b = blockIdx.z;
    mt = blockIdx.x;
    nt = blockIdx.y;
  } else {
    const int p = blockIdx.x, j = p >> 3, t = j % g.xt;
    b = (j / g.xt) * 8 + (p & 7);
    if (b >= g.batch) return;
    mt = t % g.xg;
    nt = t / g.xg;
  }
  const int m0 = mt * BM;
  const int n0 = nt * BN;
  const char* Ab = static_cast<const char*>(g.A) + g.bA.off(b, g.div) * (AREAL ? 8 : (int)sizeof(T));
  const T* Bb = static_cast<const T*>(g.B) + g.bB.off(b, g.div);
  T* Cb = static_cast<T*>(g.C) + g.bC.off(b, g.div);
  Core<MODE, BM, BN, BK, AREAL> core;
  constexpr int KV = ModeKV<MODE>::value;
  // op(A)(i,k): stored A (N: A[i + lda k]) or A^T (T: A[k + lda i]);
  // op(B)(k,j): stored B (N: B[k + ldb j], k contiguous) or B^T (T: B[j + ldb k])
  using AccA = typename std::conditional<AT, AccT<T, AREAL>, AccN<T, AREAL>>::type;
  using AccB = typename std::conditional<BT, AccN<T, false>, AccT<T, false>>::type;
  Ld<AccA, BM, BK, true, KV> la{AccA{Ab, g.lda, g.M, g.K, (V)g.csA}};
  Ld<AccB, BN, BK, false, KV> lb{AccB{Bb, g.ldb, g.N, g.K, (V)g.csB}};
  gemm_mainloop<MODE, BM, BN, BK, AREAL>(la, lb, m0, n0, g.K, s, core);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  core.for_each(wave >> 1, wave & 1, lane, [&](int r, int c, double xr, double xi) {
    const int i = m0 + r, j = n0 + c;
    if (i >= g.M || j >= g.N) return;
    T* p = Cb + i + g.ldc * j;
    double vr = g.alr * xr - g.ali * xi;
    double vi = g.alr * xi + g.ali * xr;
    if (g.epi == EPI_STORE) {
      if (g.ber != 0.0 || g.bei != 0.0) {
        double orr, oi;
        ld_c2(p, orr, oi);
        vr += g.ber * orr - g.bei * oi;
        vi += g.ber * oi + g.bei * orr;
      }
      if (g.conjC) vi = -vi;
      st_c(p, vr, vi);
    } else if (g.epi == EPI_EPS) {
      const double sc = g.svec[i] * g.svec[j];
      st_c(p, (i == j ? 1.0 : 0.0) - sc * vr, -sc * vi);
    } else {  // EPI_GJ
      if (i >= g.gj_k0 && i < g.gj_k1) return;
      double orr = 0.0, oi = 0.0;
      if (!(j >= g.gj_k0 && j < g.gj_k1)) ld_c2(p, orr, oi);
      st_c(p, orr - vr, oi - vi);
    }
  });
}

}  // namespace ecalj
