// sig3m.hip -- the bf16x3 Sigma_c frequency contraction, frequency-inner with
// three real products per complex product.
//
//   czmelwc(I,c) = sum_v sum_J Wt_v(I,J) [wgt(v,c) zmel(J,c)]
//
// (the same sum as sigbf3.hip's, with the weight applied to the zmel column
// before the product, as the FP64 form does).  Loop order: for each 32-wide
// k slice of one block's columns, zmel stays in registers (FP32) while the
// W_v stages of all frequencies stream through a three-stage LDS ring; the
// weighted column wgt(v,c) zmel(J,c) is formed and split into bf16 hi/lo in
// registers per frequency, and every frequency accumulates into the same
// MFMA registers.  Per stage a 128x128 block moves only the W tile (48 KB)
// from L2, half of what a stage of the frequency-outer kernel moves.
//
// Complex products use three real ones (the 3M form of zgemm3m, which the
// reference's CPU path calls, S/m_blas.f90:225):
//   P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi);
//   re = P1 - P2, im = P3 - P1 - P2,
// each P in bf16x3 (hi*hi + hi*lo + lo*hi, FP32 accumulation): 9 MFMAs per
// 16x16 tile and k slice instead of 12.  W is split into six planes
// (re, im, re + im; hi and lo each), pre-tiled as in sigbf3.hip (rows of 64
// bytes per k slice, chunk order chunk_slot).
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "bf16split.hpp"
#include "sigtile.hpp"

namespace ecalj {

constexpr int Q3PL = 6;    // W planes: re hi, re lo, im hi, im lo, (re+im) hi, (re+im) lo
constexpr int Q3R = 3;     // W stages in the LDS ring (two in flight)
constexpr int Q3NW = 24;   // frequencies per launch (weight table in LDS)
constexpr int Q3BN = 128;  // block columns

struct alignas(16) Sig3Smem {
  unsigned short a[Q3R][Q3PL][QBM * QBK];  // 8 KB per plane, 48 KB per stage
  float wt[Q3NW][Q3BN];                    // per-frequency weights of the block's columns
  unsigned pf[512];                        // landing area of the 32x32 form's L2 prefetch (never read)
};

// ---------------------------------------------------------------- operand preparation
// A[b][(((v*NS + ks)*6 + p)*MP + I)*32 + slot] from W (complex128, ldw, per-batch stride w_stride),
// keepwv convention Wt(I,J) = (I <= J) ? W(I,J) : conj(W(J,I)); zero beyond n.
constexpr int S3T = 32;
// One 32 x 32 tile pair per block: tile (a, b), a <= b, of the upper
// triangle of W is read once and gives two output tiles, (row tile a, k
// slice b) from W itself and (row tile b, k slice a) from its conjugate
// transpose, so every element of W is read from HBM once per frequency
// (half of reading both triangles) and each output tile written once.
__global__ __launch_bounds__(256) void split_w6_kernel(const cd* __restrict__ w, int64_t ldw, int64_t w_stride, int n,
                                                       int nw1, int MP, int NS, unsigned short* __restrict__ as,
                                                       int64_t as_stride) {
  __shared__ cd up[S3T][S3T + 1];  // up[jj][ii] = W(I0+ii, J0+jj)
  const int64_t tri = blockIdx.x;
  int b = (int)((sqrt(8.0 * (double)tri + 1.0) - 1.0) * 0.5);
  while ((int64_t)(b + 1) * (b + 2) / 2 <= tri) ++b;
  while ((int64_t)b * (b + 1) / 2 > tri) --b;
  const int a = (int)(tri - (int64_t)b * (b + 1) / 2);
  const int I0 = a * S3T, J0 = b * S3T;
  const int v = blockIdx.y % nw1, bk = blockIdx.y / nw1;
  const cd* W = w + w_stride * bk + (int64_t)ldw * n * v;
  const int tx = threadIdx.x % S3T, ty = threadIdx.x / S3T;
  for (int r = ty; r < S3T; r += 8) {
    const int I = I0 + tx, J = J0 + r;
    up[r][tx] = (I < n && J < n) ? W[I + ldw * J] : cd{0.0, 0.0};
  }
  __syncthreads();
  typedef unsigned short us4v __attribute__((ext_vector_type(4)));
  // thread: row r of the output tile, 4 k values (half hf of chunk ch)
  const int r = threadIdx.x >> 3, ch = (threadIdx.x >> 1) & 3, hf = threadIdx.x & 1;
  const int64_t pst = (int64_t)MP * QBK;
  unsigned short* Ab = as + as_stride * bk + (int64_t)v * NS * Q3PL * pst;
  auto put = [&](int I, int ks, const double (&xr)[4], const double (&xi)[4]) {
    us4v p[Q3PL];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      unsigned short h, l;
      split_bf16((float)xr[e], h, l);
      p[0][e] = h;
      p[1][e] = l;
      split_bf16((float)xi[e], h, l);
      p[2][e] = h;
      p[3][e] = l;
      split_bf16((float)(xr[e] + xi[e]), h, l);
      p[4][e] = h;
      p[5][e] = l;
    }
    unsigned short* A = Ab + (int64_t)ks * Q3PL * pst + (int64_t)I * QBK + 8 * chunk_slot(I, ch) + 4 * hf;
#pragma unroll
    for (int q = 0; q < Q3PL; ++q) *reinterpret_cast<us4v*>(A + q * pst) = p[q];
  };
  double xr[4], xi[4];
  // (row tile a, k slice b): Wt(I, J) = W(I, J) for I <= J, conj(W(J, I)) below the diagonal
  if (a < MP / S3T && b < NS) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int jj = 8 * ch + 4 * hf + e;
      if (a == b && r > jj) {
        xr[e] = up[r][jj].re;
        xi[e] = -up[r][jj].im;
      } else {
        xr[e] = up[jj][r].re;
        xi[e] = up[jj][r].im;
      }
    }
    put(I0 + r, b, xr, xi);
  }
  // (row tile b, k slice a), a < b: every element below the diagonal, conj(W(K, I))
  if (a < b && b < MP / S3T && a < NS) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int jj = 8 * ch + 4 * hf + e;
      xr[e] = up[r][jj].re;
      xi[e] = -up[r][jj].im;
    }
    put(J0 + r, a, xr, xi);
  }
}

// Launch of split_w6_kernel over nb k-points: the tile pairs of the larger
// of the row-tile and k-slice counts.
static ecalj_status split_w6_launch(ecalj_ctx* h, const cd* wc, int64_t ldw, int64_t w_stride, int n, int nw1, int nb,
                                    int MP, int NS, unsigned short* as, int64_t as_stride) {
  const int64_t T = std::max(MP / S3T, NS), npair = T * (T + 1) / 2;
  if (npair > 0x7fffffff || (int64_t)nw1 * nb > 65535) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL(split_w6_kernel, dim3((unsigned)npair, nw1 * nb), dim3(256), 0, h->stream, wc, ldw, w_stride, n,
                     nw1, MP, NS, as, as_stride);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// zp[b][c*NP + J] = zmel_b(J, c) (complex64), zero for J >= n or c >= ncol; c < NCP, J < NP
__global__ __launch_bounds__(256) void pad_z_kernel(const cf* __restrict__ z, int64_t z_stride, int n, int ncol,
                                                    int NP, int NCP, cf* __restrict__ zp, int64_t zp_stride, int nb) {
  const int64_t per = (int64_t)NCP * NP, tot = per * nb;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e % per;
    const int c = (int)(r / NP), J = (int)(r % NP);
    zp[zp_stride * b + r] = (c < ncol && J < n) ? z[z_stride * b + J + (int64_t)n * c] : cf{0.f, 0.f};
  }
}

// ---------------------------------------------------------------- block tile
// Rows I0..I0+128 of one internal k, columns c0..c0+128; 8 waves in 2 x 4,
// wave tile 64 x 32 (4 x 2 MFMA tiles).  Ak: the k's W planes; zp: its padded
// zmel (column c at zp + c*NP).  Stage st = ks*nwg + vl carries the W tile of
// frequency v0+vl and k slice ks.
// WM: wave rows; the 8 waves form WM x 8/WM, each (128/WM) rows x 16*WM columns
// (WM = 2: 4 x 2 MFMA tiles per wave; WM = 1: 8 x 1, so one split of the
// weighted columns feeds twice the MFMAs and each wave reads all 128 rows).
template <int WM>
__device__ __forceinline__ void sig3m_tile(Sig3Smem& s, const unsigned short* __restrict__ Ak,
                                           const cf* __restrict__ zp, const double* __restrict__ g, int n, int nw1,
                                           int v0, int nwg, int ncol, int MP, int NS, int NP, int I0, int c0,
                                           int accumulate, cf* __restrict__ C, int64_t ldc) {
  static_assert(WM == 1 || WM == 2, "one or two wave rows");
  constexpr int NT = 512, WN = 8 / WM, TM = 8 / WM, TN = WM;
  constexpr int APP = QBM * QBK / 512;        // 1 KB runs per plane
  constexpr int ARW = Q3PL * APP / 8;         // runs per wave and stage
  static_assert(ARW * 8 == Q3PL * APP, "runs split evenly over the waves");
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), wm = wave / WN, wn = wave % WN;
  const int64_t apst = (int64_t)MP * QBK;
  // uniform base (scalar registers) + the lane's 16 bytes of each 1 KB run, so
  // the stage loads take the scalar-base form with no per-load address VALU
  const unsigned short* A = Ak + (int64_t)I0 * QBK;
  const int lane8 = lane * 8;
  const int nst = NS * nwg;

  for (int e = t; e < nwg * Q3BN; e += NT) {
    const int v = e / Q3BN, c = c0 + e % Q3BN;
    s.wt[v][e % Q3BN] = (c < ncol) ? (float)g[(v0 + v) + (int64_t)nw1 * c] : 0.f;
  }
  __syncthreads();

  auto issue = [&](int st, int buf) {
    const int ks = st / nwg, v = v0 + st % nwg;
    const unsigned short* a = A + ((int64_t)v * NS + ks) * Q3PL * apst;
#pragma unroll
    for (int q = 0; q < ARW; ++q) {
      const int run = ARW * wave + q, p = run / APP, sub = run % APP;
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(a + (p * apst + sub * 512) + lane8),
                                       (lds_ptr_t)(&s.a[buf][p][sub * 512]), 16, 0, 0);
    }
  };

  // the lane's zmel: column c0 + 32 wn + 16 j + (lane & 15), k values 8 (lane >> 4) .. +7 of the slice,
  // as (re, im) pairs: zv[j][q] = (re, im) of k = 2q and 2q+1
  f4 zv[TN][4];
  const cf* zcol[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j)
    zcol[j] = zp + (int64_t)(c0 + wn * 16 * TN + j * 16 + (lane & 15)) * NP + 8 * (lane >> 4);
  auto zload = [&](int ks) {
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) zv[j][q] = *reinterpret_cast<const f4*>(zcol[j] + (int64_t)ks * QBK + 2 * q);
  };

  f4 p1[TM][TN], p2[TM][TN], p3[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) p1[i][j] = p2[i][j] = p3[i][j] = f4{0, 0, 0, 0};
  const int slot = chunk_slot(lane & 15, lane >> 4);
  const int aoff = (wm * TM * 16 + (lane & 15)) * QBK + 8 * slot;

  // issue order: zmel of slice 0, then stages 0 .. Q3R-2; in the loop the zmel of
  // slice ks+1 is loaded in the last frequency of slice ks, before that
  // iteration's stage.  Every iteration issues one stage: past the last one it
  // loads the last stage again into the free buffer, so that exactly the ARW
  // loads of one stage follow any other load, and waiting for all but ARW
  // loads covers the current stage and the current slice's zmel.
  zload(0);
  __builtin_amdgcn_sched_barrier(0);
  for (int q = 0; q < Q3R - 1; ++q) issue(min(q, nst - 1), q);
  // weighted zmel columns of frequency vl, split into bf16 hi/lo: re, im, re + im
  u4q brh[TN], brl[TN], bih[TN], bil[TN], bsh[TN], bsl[TN];
  // the residual of the split: v_dot2c in the 1 x 8 layout (fewest VALU
  // instructions); unpack and subtract in the 2 x 4 layout, where the split of
  // one wave then overlaps the MFMAs of the other wave on the SIMD, which
  // v_dot2c does not (C3 step: 422 against 435 ms)
  const SplitMasks mk = split_masks();
  auto split2 = [&](float x0, float x1) { return WM == 1 ? split2_bf16(x0, x1, mk) : split2_bf16_alu(x0, x1); };
  auto make_b = [&](int vl) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const float w = s.wt[vl][wn * 16 * TN + j * 16 + (lane & 15)];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const f4 z = zv[j][q];
        const float r0 = w * z[0], i0 = w * z[1], r1 = w * z[2], i1 = w * z[3];
        const float s0 = r0 + i0, s1 = r1 + i1;
        {
          const HiLo2 t = split2(r0, r1);
          brh[j][q] = t.hi;
          brl[j][q] = t.lo;
        }
        {
          const HiLo2 t = split2(i0, i1);
          bih[j][q] = t.hi;
          bil[j][q] = t.lo;
        }
        {
          const HiLo2 t = split2(s0, s1);
          bsh[j][q] = t.hi;
          bsl[j][q] = t.lo;
        }
      }
    }
  };

  // One frequency of one k slice.  TOP: form this frequency's weighted columns
  // after the barrier; END: form the next frequency's after the MFMAs (then
  // the next step has them already).  ZL: the last frequency of the slice,
  // which also loads the next slice's zmel (the last slice's again at the end,
  // so the step is straight-line code; scheduling barriers keep the loads
  // after the last read of zv and before the stage's loads).  Waves 0-3 form
  // every frequency's columns at the top; waves 4-7 form those of the slice's
  // first frequency at the top and every later one at the end of the step
  // before.  Wave w runs on SIMD w % 4, so between two barriers one wave of a
  // SIMD splits operands on the VALU while the other issues MFMAs.
  auto substep = [&](int ks, int vl, int st, auto top, auto end, auto zl) {
    constexpr bool TOP = decltype(top)::value, END = decltype(end)::value, ZL = decltype(zl)::value;
    const int buf = st % Q3R;
    ECALJ_WAIT_VMCNT(ARW);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    // W fragments of row tile i (planes: re hi, re lo, im hi, im lo,
    // (re + im) hi, (re + im) lo).  The LDS stage loads in flight make the
    // compiler drain every outstanding LDS read (lgkmcnt(0)) before the first
    // use of any of them, so each group of reads is issued just after the
    // MFMA that waits for the previous group, and the rest of the group's
    // MFMAs cover it.  Tile 0's first reads go before the stage's global loads.
    //  WM = 1 (TN = 1, registers to spare): two register sets; tile i+1's six
    //    planes are read after tile i's first MFMA.
    //  WM = 2: one set; the lo planes of tile i after its first MFMA, the hi
    //    planes of tile i+1 (in the registers just released) after its first
    //    lo MFMA.
    us8q fa[WM == 1 ? 2 : 1][Q3PL];
    auto read = [&](int i, int set, int p0, int dp) {
      const int o = aoff + i * 16 * QBK;
#pragma unroll
      for (int p = p0; p < Q3PL; p += dp) fa[set][p] = *reinterpret_cast<const us8q*>(&s.a[buf][p][o]);
    };
    if constexpr (TOP) make_b(vl);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (WM == 1)
      read(0, 0, 0, 1);
    else
      read(0, 0, 0, 2);
    if constexpr (ZL) {
      __builtin_amdgcn_sched_barrier(0);
      zload(min(ks + 1, NS - 1));
      __builtin_amdgcn_sched_barrier(0);
    }
    // the buffer of stage st+Q3R-1 was last read in iteration st-1, which every wave has left
    issue(min(st + Q3R - 1, nst - 1), (st + Q3R - 1) % Q3R);

    auto mma = [&](f4& acc, const us8q& a, const u4q& b) {
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), acc, 0,
                                                    0, 0);
    };
    if constexpr (WM == 1) {
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int c = i & 1;
        const us8q* a = fa[c];
        __builtin_amdgcn_sched_barrier(0);
        mma(p1[i][0], a[0], brh[0]);
        __builtin_amdgcn_sched_barrier(0);
        if (i + 1 < TM) read(i + 1, c ^ 1, 0, 1);
        __builtin_amdgcn_sched_barrier(0);
        mma(p2[i][0], a[2], bih[0]);
        mma(p3[i][0], a[4], bsh[0]);
        mma(p1[i][0], a[0], brl[0]);
        mma(p2[i][0], a[2], bil[0]);
        mma(p3[i][0], a[4], bsl[0]);
        mma(p1[i][0], a[1], brh[0]);
        mma(p2[i][0], a[3], bih[0]);
        mma(p3[i][0], a[5], bsh[0]);
      }
    } else {
      us8q* a = fa[0];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        __builtin_amdgcn_sched_barrier(0);
        mma(p1[i][0], a[0], brh[0]);
        __builtin_amdgcn_sched_barrier(0);
        read(i, 0, 1, 2);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 1; j < TN; ++j) mma(p1[i][j], a[0], brh[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j) mma(p2[i][j], a[2], bih[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j) mma(p3[i][j], a[4], bsh[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j) mma(p1[i][j], a[0], brl[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j) mma(p2[i][j], a[2], bil[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j) mma(p3[i][j], a[4], bsl[j]);
        __builtin_amdgcn_sched_barrier(0);
        mma(p1[i][0], a[1], brh[0]);
        __builtin_amdgcn_sched_barrier(0);
        if (i + 1 < TM) read(i + 1, 0, 0, 2);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 1; j < TN; ++j) mma(p1[i][j], a[1], brh[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j) mma(p2[i][j], a[3], bih[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j) mma(p3[i][j], a[5], bsh[j]);
      }
    }
    if constexpr (END) {  // after the MFMAs in program order too: one set of operand registers
      __builtin_amdgcn_sched_barrier(0);
      make_b(vl + 1);
    }
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  int st = 0, ks = 0;
  if (wave < 4 || nwg == 1) {
    do {  // do-while (NS, nwg >= 1): the prologue's loads stay in the block that precedes the loop
      for (int vl = 0; vl < nwg - 1; ++vl, ++st) substep(ks, vl, st, T_{}, F_{}, F_{});
      substep(ks, nwg - 1, st, T_{}, F_{}, T_{});
      ++st;
    } while (++ks < NS);
  } else {
    do {
      substep(ks, 0, st++, T_{}, T_{}, F_{});
      for (int vl = 1; vl < nwg - 1; ++vl, ++st) substep(ks, vl, st, F_{}, T_{}, F_{});
      substep(ks, nwg - 1, st++, F_{}, F_{}, T_{});
    } while (++ks < NS);
  }
  ECALJ_WAIT_VMCNT(0);  // the repeated last stages have landed before the block's LDS is released
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int I = I0 + wm * TM * 16 + i * 16 + acc_row_f32(lane, e);
        const int c = c0 + wn * TN * 16 + j * 16 + (lane & 15);
        if (I < n && c < ncol) {
          const float re = p1[i][j][e] - p2[i][j][e];
          const float im = p3[i][j][e] - p1[i][j][e] - p2[i][j][e];
          cf* p = C + I + ldc * c;
          if (accumulate) {
            const cf o = *p;
            *p = cf{o.re + re, o.im + im};
          } else {
            *p = cf{re, im};
          }
        }
      }
}

// Uniform batch.  G == 0: block (x, y, z) = (column tile, row tile, internal k):
// consecutive blocks share their W stream.  G > 0, XCD-aware: a 1-D grid
// whose blocks the dispatcher deals round-robin to the 8 XCDs (block p on XCD
// p % 8).  XCD x's j-th block belongs to super-tile (j / G)*8 + x of G column
// tiles of one (k, row tile), so the G blocks that read the same W stages run
// together on one XCD and share them through its L2.
template <int WM>
__global__ __launch_bounds__(512, 1) void sig3m_kernel(const unsigned short* __restrict__ as, int64_t as_stride,
                                                       const cf* __restrict__ zp, int64_t zp_stride,
                                                       const double* __restrict__ wgt, int64_t g_stride, int n,
                                                       int nw1, int v0, int nwg, int ncol, int MP, int NS, int NP,
                                                       int NCP, int accumulate, cf* __restrict__ cz,
                                                       int64_t cz_stride, int G, int nk) {
  __shared__ Sig3Smem s;
  int b, r, c;
  if (G == 0) {
    c = blockIdx.x;
    r = blockIdx.y;
    b = blockIdx.z;
  } else {
    const int R = MP / QBM, Ct = NCP / Q3BN, NG = (Ct + G - 1) / G;
    const int p = blockIdx.x, x = p & 7, j = p >> 3;
    const int sigma = (j / G) * 8 + x, loc = j % G;
    b = sigma / (R * NG);
    const int rem = sigma % (R * NG);
    r = rem / NG;
    c = (rem % NG) * G + loc;
    if (b >= nk || c >= Ct) return;
  }
  sig3m_tile<WM>(s, as + as_stride * b, zp + zp_stride * b, wgt + g_stride * b, n, nw1, v0, nwg, ncol, MP, NS, NP,
             r * QBM, c * Q3BN, accumulate, cz + cz_stride * b, (int64_t)n);
}

// Grouped, persistent: internal k-points with different column counts.  The
// tile list holds (k, row tile, column tile) in the order the blocks take
// them: block p takes entries p, p + grid, ... (grid a multiple of 8), and
// the list gives the blocks of one XCD in one round consecutive tiles of one
// (k, row tile), which then share their W stages in that XCD's L2.
struct Sig3Group {
  int64_t zoff;   // padded zmel of k (complex64 elements)
  int64_t goff;   // weights of k (elements)
  int64_t czoff;  // output of k (elements)
  int ncol, pad;
};
template <int WM>
__global__ __launch_bounds__(512, 1) void sig3m_grouped_kernel(const int* __restrict__ tiles, int ntiles,
                                                               const Sig3Group* __restrict__ grp,
                                                               const unsigned short* __restrict__ as,
                                                               int64_t as_stride, const cf* __restrict__ zp,
                                                               const double* __restrict__ wgt, int n, int nw1,
                                                               int v0, int nwg, int MP, int NS, int NP,
                                                               int accumulate, cf* __restrict__ cz) {
  __shared__ Sig3Smem s;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int k = tiles[3 * t];
    if (k < 0) continue;  // a hole of the XCD order
    const int it = tiles[3 * t + 1], jt = tiles[3 * t + 2];
    const Sig3Group gk = grp[k];
    __syncthreads();  // the previous tile's LDS reads are done before this tile's loads land
    sig3m_tile<WM>(s, as + as_stride * k, zp + gk.zoff, wgt + gk.goff, n, nw1, v0, nwg, gk.ncol, MP, NS, NP,
                   it * QBM, jt * Q3BN, accumulate, cz + gk.czoff, (int64_t)n);
  }
}

// ---------------------------------------------------------------- three products, 32x32 MFMA, one wave per SIMD
// The frequency-inner three-product contraction on v_mfma_f32_32x32x16_bf16:
// a block of 128 rows x 128 columns, 4 waves side by side, each all 128 rows
// x 32 columns (4 row tiles of 32).  A wave reads each W fragment once for
// 32 columns (half the LDS reads of the 16-column layout) and splits its 32
// columns once per block (half the splits of the 2 x 4 layout).  With one
// wave per SIMD the VALU split of the next frequency's weighted columns runs
// between this frequency's MFMAs, in the same wave, while each 32x32x16 MFMA
// occupies the matrix core for 32 cycles.  Operand layout of the 32x32x16
// instruction: lane l holds A(row l % 32, k 8 (l / 32) .. +7) and B(k the
// same, column l % 32); accumulator element e of lane l is C(row 8 (e / 4) +
// 4 (l / 32) + e % 4, column l % 32).
constexpr int Q32NT = 256;  // threads of the 32x32 form

typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void sig3m32_tile(Sig3Smem& s, const unsigned short* __restrict__ Ak,
                                             const cf* __restrict__ zp, const double* __restrict__ g, int n,
                                             int nw1, int v0, int nwg, int ncol, int MP, int NS, int NP, int I0,
                                             int c0, int accumulate, cf* __restrict__ C, int64_t ldc) {
  constexpr int TM = 4;                       // row tiles of 32
  constexpr int APP = QBM * QBK / 512;        // 1 KB runs per plane
  constexpr int ARW = Q3PL * APP / 4;         // runs per wave and stage
  static_assert(ARW * 4 == Q3PL * APP, "runs split evenly over the waves");
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t apst = (int64_t)MP * QBK;
  const unsigned short* A = Ak + (int64_t)I0 * QBK;
  const int lane8 = lane * 8;
  const int nst = NS * nwg;

  for (int e = t; e < nwg * Q3BN; e += Q32NT) {
    const int v = e / Q3BN, c = c0 + e % Q3BN;
    s.wt[v][e % Q3BN] = (c < ncol) ? (float)g[(v0 + v) + (int64_t)nw1 * c] : 0.f;
  }
  __syncthreads();

  auto issue = [&](int st, int buf) {
    const int ks = st / nwg, v = v0 + st % nwg;
    const unsigned short* a = A + ((int64_t)v * NS + ks) * Q3PL * apst;
#pragma unroll
    for (int q = 0; q < ARW; ++q) {
      const int run = ARW * wave + q, p = run / APP, sub = run % APP;
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(a + (p * apst + sub * 512) + lane8),
                                       (lds_ptr_t)(&s.a[buf][p][sub * 512]), 16, 0, 0);
    }
  };

  // the lane's zmel: column c0 + 32 wave + (lane & 31), k = 16 h + 8 (lane >> 5) + 2q, +1 of the slice
  const int col = wave * 32 + (lane & 31);
  const cf* zcol = zp + (int64_t)(c0 + col) * NP + 8 * (lane >> 5);
  f4 zc[2][4], zn[2][4];  // this slice's, the next slice's
  auto zload = [&](f4 (&z)[2][4], int ks) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < 4; ++q) z[h][q] = *reinterpret_cast<const f4*>(zcol + (int64_t)ks * QBK + 16 * h + 2 * q);
  };

  // weighted zmel columns, split: [plane][half] with planes re hi, re lo, im hi, im lo, (re+im) hi, (re+im) lo
  typedef u4q BSet[6][2];
  BSet b0, b1;
  // one eighth (k pair q of half h) of the split of frequency vl from zc
  // (the empty asm statements pin each result where it is formed: the
  // computation is otherwise sunk to its first use, the next step's MFMAs)
  auto split_part = [&](BSet& b, float w, int h, int q) {
    const f4 z = zc[h][q];
    const float r0 = w * z[0], i0 = w * z[1], r1 = w * z[2], i1 = w * z[3];
    const float s0 = r0 + i0, s1 = r1 + i1;
    HiLo2 x[3] = {split2_bf16_alu(r0, r1), split2_bf16_alu(i0, i1), split2_bf16_alu(s0, s1)};
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      asm volatile("" : "+v"(x[p].hi), "+v"(x[p].lo));
      b[2 * p][h][q] = x[p].hi;
      b[2 * p + 1][h][q] = x[p].lo;
    }
  };

  f16v p1[TM], p2[TM], p3[TM];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 16; ++e) p1[i][e] = p2[i][e] = p3[i][e] = 0.f;
  // fragment offsets in a plane: row 32 i + (lane & 31), logical chunk 2 h + (lane >> 5)
  int aoff[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) aoff[h] = (lane & 31) * QBK + 8 * chunk_slot(lane & 31, 2 * h + (lane >> 5));

  // prologue: slice 0's zmel (and slice 1's, which one frequency per slice
  // needs at the first step), the first stages, frequency 0's split
  zload(zc, 0);
  zload(zn, min(1, NS - 1));
  __builtin_amdgcn_sched_barrier(0);
  for (int q = 0; q < Q3R - 1; ++q) issue(min(q, nst - 1), q);
  {
    const float w = s.wt[0][col];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < 4; ++q) split_part(b0, w, h, q);
  }

  auto mma = [&](f16v& acc, const us8q& a, const u4q& b) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), acc, 0, 0,
                                                  0);
  };
  // One step: frequency vl of slice ks from stage st, with the weighted
  // columns in bc; the next step's are formed into bn between the MFMAs.
  // The LDS stage loads in flight make the compiler drain every outstanding
  // LDS read (lgkmcnt(0)) before the first use of any, so each group of
  // fragment reads is issued just after the MFMA that waits for the previous
  // group: half 1 of row tile i after the first MFMA of half 0, half 0 of
  // tile i+1 after the first MFMA of half 1.
  // L2 prefetch of the stage one step beyond the LDS ring: the ring holds
  // two stages in flight, which at this kernel's step time does not cover
  // the fabric latency.  Wave w touches the 64 128-byte lines of planes w and
  // w + 4 (the last plane again when there is no plane w + 4), one line per
  // lane, with dword loads through the LDS-direct path into a 2 KB scratch
  // area that nothing reads: no register waits for them, and the stage wait
  // counts them with the stage loads.
  constexpr int NPF = 2;
  auto prefetch = [&](int stp) {
    const int ks = stp / nwg, v = v0 + stp % nwg;
    const unsigned short* a = A + ((int64_t)v * NS + ks) * Q3PL * apst;
#pragma unroll
    for (int q = 0; q < NPF; ++q) {
      const int p = min(wave + 4 * q, Q3PL - 1);
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(a + p * apst + lane * 64), (lds_ptr_t)(&s.pf[(q * 4 + wave) * 64]),
                                       4, 0, 0);
    }
  };
  prefetch(min(Q3R - 1, nst - 1));  // every step's stage wait counts ARW stage loads + NPF prefetches last
  auto step = [&](int st, BSet& bc, BSet& bn) {
    const int ks = st / nwg, vl = st - ks * nwg, buf = st % Q3R;
    ECALJ_WAIT_VMCNT(ARW + NPF);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    // The split at the last frequency of slice ks is that of slice ks+1's
    // first, from slice ks+1's zmel: moved into zc here from zn, which the
    // first frequency of the slice loaded (one frequency per slice: the
    // slice after, as the move comes first) before the step's stage loads,
    // so the next step's stage wait covers it.
    if (vl == nwg - 1) {
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int q = 0; q < 4; ++q) zc[h][q] = zn[h][q];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (vl == 0) zload(zn, min(ks + (nwg == 1 ? 2 : 1), NS - 1));
    us8q fr[2][6];  // [half][plane]
    auto read = [&](int i, int h) {
      const int o = aoff[h] + i * 32 * QBK;
#pragma unroll
      for (int p = 0; p < 6; ++p) fr[h][p] = *reinterpret_cast<const us8q*>(&s.a[buf][p][o]);
    };
    __builtin_amdgcn_sched_barrier(0);
    read(0, 0);
    const float wn = s.wt[vl + 1 < nwg ? vl + 1 : 0][col];
    __builtin_amdgcn_sched_barrier(0);
    // the buffer of stage st+Q3R-1 was last read in step st-1, which every wave has left
    issue(min(st + Q3R - 1, nst - 1), (st + Q3R - 1) % Q3R);
    prefetch(min(st + Q3R, nst - 1));
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const us8q* a = fr[h];
        __builtin_amdgcn_sched_barrier(0);
        mma(p1[i], a[0], bc[0][h]);
        __builtin_amdgcn_sched_barrier(0);
        if (h == 0)
          read(i, 1);
        else if (i + 1 < TM)
          read(i + 1, 0);
        __builtin_amdgcn_sched_barrier(0);
        mma(p2[i], a[2], bc[2][h]);
        mma(p3[i], a[4], bc[4][h]);
        mma(p1[i], a[0], bc[1][h]);
        mma(p2[i], a[2], bc[3][h]);
        mma(p3[i], a[4], bc[5][h]);
        mma(p1[i], a[1], bc[0][h]);
        mma(p2[i], a[3], bc[2][h]);
        mma(p3[i], a[5], bc[4][h]);
        // one eighth of the next step's split between these MFMAs
        split_part(bn, wn, (2 * i + h) >> 2, (2 * i + h) & 3);
      }
    }
  };
  int st = 0;
  for (; st + 1 < nst; st += 2) {
    step(st, b0, b1);
    step(st + 1, b1, b0);
  }
  if (st < nst) step(st, b0, b1);
  ECALJ_WAIT_VMCNT(0);  // the repeated last stages have landed before the block's LDS is released
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int I = I0 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
      const int c = c0 + col;
      if (I < n && c < ncol) {
        const float re = p1[i][e] - p2[i][e];
        const float im = p3[i][e] - p1[i][e] - p2[i][e];
        cf* p = C + I + ldc * c;
        if (accumulate) {
          const cf o = *p;
          *p = cf{o.re + re, o.im + im};
        } else {
          *p = cf{re, im};
        }
      }
    }
}

// Uniform batch, block order as sig3m_kernel.
__global__ __launch_bounds__(Q32NT, 1) void sig3m32_kernel(const unsigned short* __restrict__ as, int64_t as_stride,
                                                           const cf* __restrict__ zp, int64_t zp_stride,
                                                           const double* __restrict__ wgt, int64_t g_stride, int n,
                                                           int nw1, int v0, int nwg, int ncol, int MP, int NS,
                                                           int NP, int NCP, int accumulate, cf* __restrict__ cz,
                                                           int64_t cz_stride, int G, int nk) {
  __shared__ Sig3Smem s;
  int b, r, c;
  if (G == 0) {
    c = blockIdx.x;
    r = blockIdx.y;
    b = blockIdx.z;
  } else {
    const int R = MP / QBM, Ct = NCP / Q3BN, NG = (Ct + G - 1) / G;
    const int p = blockIdx.x, x = p & 7, j = p >> 3;
    const int sigma = (j / G) * 8 + x, loc = j % G;
    b = sigma / (R * NG);
    const int rem = sigma % (R * NG);
    r = rem / NG;
    c = (rem % NG) * G + loc;
    if (b >= nk || c >= Ct) return;
  }
  sig3m32_tile(s, as + as_stride * b, zp + zp_stride * b, wgt + g_stride * b, n, nw1, v0, nwg, ncol, MP, NS, NP,
               r * QBM, c * Q3BN, accumulate, cz + cz_stride * b, (int64_t)n);
}

// ---------------------------------------------------------------- four-product, waves side by side
// The same frequency-inner loop with the four-product complex form on the
// plain W planes of sigbf3.hip (re hi, re lo, im hi, im lo): a block of
// 128 rows x 128 columns, 8 waves side by side, each all 128 rows x 16
// columns (8 MFMA tiles), so the weighted columns one wave splits serve 8
// row tiles (96 products per split of 8 complex values) and no column is
// split twice.  Four W stages in the ring (three in flight).
constexpr int QV4BN = 128;
constexpr int QV4R = 4;
constexpr int QV4NW = 24;

struct alignas(16) SigV4Smem {
  unsigned short a[QV4R][4][QBM * QBK];  // 8 KB per plane, 32 KB per stage
  float wt[QV4NW][QV4BN];
};

// Ak: the k's W planes, 4-plane layout [((v*NS + ks)*4 + p)*MP + I)*32 + slot];
// zp: its padded zmel (column c at zp + c*NP).
__device__ __forceinline__ void sigv4_tile(SigV4Smem& s, const unsigned short* __restrict__ Ak,
                                           const cf* __restrict__ zp, const double* __restrict__ g, int n, int nw1,
                                           int v0, int nwg, int ncol, int MP, int NS, int NP, int I0, int c0,
                                           int accumulate, cf* __restrict__ C, int64_t ldc) {
  constexpr int NT = 512, TM = 8, TN = 1;
  constexpr int APP = QBM * QBK / 512;   // 1 KB runs per plane
  constexpr int ARW = 4 * APP / 8;       // runs per wave and stage
  static_assert(ARW * 8 == 4 * APP, "runs split evenly over the waves");
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t apst = (int64_t)MP * QBK;
  // uniform base (scalar registers) + the lane's 16 bytes of each 1 KB run, so
  // the stage loads take the scalar-base form with no per-load address VALU
  const unsigned short* A = Ak + (int64_t)I0 * QBK;
  const int lane8 = lane * 8;
  const int nst = NS * nwg;

  for (int e = t; e < nwg * QV4BN; e += NT) {
    const int v = e / QV4BN, c = c0 + e % QV4BN;
    s.wt[v][e % QV4BN] = (c < ncol) ? (float)g[(v0 + v) + (int64_t)nw1 * c] : 0.f;
  }
  __syncthreads();

  auto issue = [&](int st, int buf) {
    const int ks = st / nwg, v = v0 + st % nwg;
    const unsigned short* a = A + ((int64_t)v * NS + ks) * 4 * apst;
#pragma unroll
    for (int q = 0; q < ARW; ++q) {
      const int run = ARW * wave + q, p = run / APP, sub = run % APP;
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(a + (p * apst + sub * 512) + lane8),
                                       (lds_ptr_t)(&s.a[buf][p][sub * 512]), 16, 0, 0);
    }
  };

  f4 zv[TN][4];
  const cf* zcol[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) zcol[j] = zp + (int64_t)(c0 + wave * 16 + j * 16 + (lane & 15)) * NP + 8 * (lane >> 4);
  auto zload = [&](int ks) {
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) zv[j][q] = *reinterpret_cast<const f4*>(zcol[j] + (int64_t)ks * QBK + 2 * q);
  };

  f4 cr[TM][TN], ci[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) cr[i][j] = ci[i][j] = f4{0, 0, 0, 0};
  const int aoff = (lane & 15) * QBK + 8 * chunk_slot(lane & 15, lane >> 4);

  zload(0);
  __builtin_amdgcn_sched_barrier(0);
  for (int q = 0; q < QV4R - 1; ++q) issue(min(q, nst - 1), q);

  // weighted zmel columns (re, im; -im for the real part), split into bf16 hi/lo
  u4q brh[TN], brl[TN], bih[TN], bil[TN], nbh[TN], nbl[TN];
  const SplitMasks mk = split_masks();
  auto make_b = [&](int vl) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const float w = s.wt[vl][wave * 16 + j * 16 + (lane & 15)];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const f4 z = zv[j][q];
        const float r0 = w * z[0], i0 = w * z[1], r1 = w * z[2], i1 = w * z[3];
        {
          const HiLo2 t = split2_bf16(r0, r1, mk);
          brh[j][q] = t.hi;
          brl[j][q] = t.lo;
        }
        {
          const HiLo2 t = split2_bf16(i0, i1, mk);
          bih[j][q] = t.hi;
          bil[j][q] = t.lo;
        }
      }
      nbh[j] = bih[j] ^ 0x80008000u;
      nbl[j] = bil[j] ^ 0x80008000u;
    }
  };

  // TOP / END / ZL as in sig3m_tile.  The zmel of slice ks+1 is loaded after
  // stage st+Q3R-2 = st+2 and before stage st+3 of the slice's last step st, so
  // the next step (the first of slice ks+1) waits for all but one stage's loads.
  auto substep = [&](int ks, int vl, int st, auto top, auto end, auto zl) {
    constexpr bool TOP = decltype(top)::value, END = decltype(end)::value, ZL = decltype(zl)::value;
    const int buf = st % QV4R;
    if (vl == 0 && ks > 0) {
      ECALJ_WAIT_VMCNT(ARW);
    } else {
      ECALJ_WAIT_VMCNT(2 * ARW);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if constexpr (TOP) make_b(vl);
    if constexpr (ZL) {
      __builtin_amdgcn_sched_barrier(0);
      zload(min(ks + 1, NS - 1));
      __builtin_amdgcn_sched_barrier(0);
    }
    // the buffer of stage st+QV4R-1 was last read in iteration st-1, which every wave has left
    issue(min(st + QV4R - 1, nst - 1), (st + QV4R - 1) % QV4R);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int o = aoff + i * 16 * QBK;
      const us8q arh = *reinterpret_cast<const us8q*>(&s.a[buf][0][o]);
      const us8q arl = *reinterpret_cast<const us8q*>(&s.a[buf][1][o]);
      const us8q aih = *reinterpret_cast<const us8q*>(&s.a[buf][2][o]);
      const us8q ail = *reinterpret_cast<const us8q*>(&s.a[buf][3][o]);
#define ECALJ_V4_TERM(ACC, AF, BF)                                                                      \
  _Pragma("unroll") for (int j = 0; j < TN; ++j) ACC[i][j] =                                            \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, AF), __builtin_bit_cast(bf8, BF[j]), \
                                              ACC[i][j], 0, 0, 0);
      // re = ar br - ai bi, im = ar bi + ai br; hi*hi + hi*lo + lo*hi of each
      ECALJ_V4_TERM(cr, arh, brh)
      ECALJ_V4_TERM(ci, arh, bih)
      ECALJ_V4_TERM(cr, arh, brl)
      ECALJ_V4_TERM(ci, arh, bil)
      ECALJ_V4_TERM(cr, arl, brh)
      ECALJ_V4_TERM(ci, arl, bih)
      ECALJ_V4_TERM(cr, aih, nbh)
      ECALJ_V4_TERM(ci, aih, brh)
      ECALJ_V4_TERM(cr, aih, nbl)
      ECALJ_V4_TERM(ci, aih, brl)
      ECALJ_V4_TERM(cr, ail, nbh)
      ECALJ_V4_TERM(ci, ail, brh)
#undef ECALJ_V4_TERM
    }
    if constexpr (END) {
      __builtin_amdgcn_sched_barrier(0);
      make_b(vl + 1);
    }
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  int st = 0, ks = 0;
  if (wave < 4 || nwg == 1) {
    do {
      for (int vl = 0; vl < nwg - 1; ++vl, ++st) substep(ks, vl, st, T_{}, F_{}, F_{});
      substep(ks, nwg - 1, st, T_{}, F_{}, T_{});
      ++st;
    } while (++ks < NS);
  } else {
    do {
      substep(ks, 0, st++, T_{}, T_{}, F_{});
      for (int vl = 1; vl < nwg - 1; ++vl, ++st) substep(ks, vl, st, F_{}, T_{}, F_{});
      substep(ks, nwg - 1, st++, F_{}, F_{}, T_{});
    } while (++ks < NS);
  }
  ECALJ_WAIT_VMCNT(0);  // the repeated last stages have landed before the block's LDS is released
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int I = I0 + i * 16 + acc_row_f32(lane, e);
        const int c = c0 + wave * 16 + j * 16 + (lane & 15);
        if (I < n && c < ncol) {
          cf* p = C + I + ldc * c;
          if (accumulate) {
            const cf o = *p;
            *p = cf{o.re + cr[i][j][e], o.im + ci[i][j][e]};
          } else {
            *p = cf{cr[i][j][e], ci[i][j][e]};
          }
        }
      }
}

// Uniform batch, as sig3m_kernel (G == 0: block (column tile, row tile, k);
// G > 0: XCD-aware super-tiles of G column tiles of one (k, row tile)).
__global__ __launch_bounds__(512, 1) void sigv4_kernel(const unsigned short* __restrict__ as, int64_t as_stride,
                                                       const cf* __restrict__ zp, int64_t zp_stride,
                                                       const double* __restrict__ wgt, int64_t g_stride, int n,
                                                       int nw1, int v0, int nwg, int ncol, int MP, int NS, int NP,
                                                       int NCP, int accumulate, cf* __restrict__ cz,
                                                       int64_t cz_stride, int G, int nk) {
  __shared__ SigV4Smem s;
  int b, r, c;
  if (G == 0) {
    c = blockIdx.x;
    r = blockIdx.y;
    b = blockIdx.z;
  } else {
    const int R = MP / QBM, Ct = NCP / QV4BN, NG = (Ct + G - 1) / G;
    const int p = blockIdx.x, x = p & 7, j = p >> 3;
    const int sigma = (j / G) * 8 + x, loc = j % G;
    b = sigma / (R * NG);
    const int rem = sigma % (R * NG);
    r = rem / NG;
    c = (rem % NG) * G + loc;
    if (b >= nk || c >= Ct) return;
  }
  sigv4_tile(s, as + as_stride * b, zp + zp_stride * b, wgt + g_stride * b, n, nw1, v0, nwg, ncol, MP, NS, NP,
             r * QBM, c * QV4BN, accumulate, cz + cz_stride * b, (int64_t)n);
}

// ---------------------------------------------------------------- launcher
size_t sig3m_a_elems_per_k(int n, int nw1) {
  return (size_t)Q3PL * QBK * nw1 * pad_to(n, QBM) * (pad_to(n, QBK) / QBK);
}

ecalj_status sig3m_prep(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, int nw1,
                        const cd* wc, int64_t ldw, int64_t w_stride, unsigned short* as, cf* zp) {
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NCP = pad_to(ncol, QBN_MAX), NS = NP / QBK;
  const int64_t as_stride = (int64_t)sig3m_a_elems_per_k(n, nw1);
  ECALJ_TRY(split_w6_launch(h, wc, ldw, w_stride, n, nw1, nb, MP, NS, as, as_stride));
  const int64_t totz = (int64_t)NCP * NP * nb;
  const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
  hipLaunchKernelGGL(pad_z_kernel, dim3(bz), dim3(256), 0, h->stream, zmel, z_stride, n, ncol, NP, NCP, zp,
                     (int64_t)NCP * NP, nb);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// Column tiles per super-tile of the XCD-aware order: the column tiles of one
// (k, row tile) split evenly into groups of at most 32 (one XCD's CUs).
static int sig3m_group(int Ct) {
  const int ng = (Ct + 31) / 32;
  return (Ct + ng - 1) / ng;
}

ecalj_status sig3m_run(ecalj_ctx* h, bool xcd, int wm, int nb, int n, int ncol, int nw1, const unsigned short* as,
                       const cf* zp, const double* wgt, int64_t g_stride, cf* cz, int64_t cz_stride) {
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NCP = pad_to(ncol, QBN_MAX), NS = NP / QBK;
  const int R = MP / QBM, Ct = NCP / Q3BN;
  const int64_t as_stride = (int64_t)sig3m_a_elems_per_k(n, nw1);
  dim3 grid(Ct, R, nb);
  int G = 0;
  if (xcd) {
    G = sig3m_group(Ct);
    const int64_t NG = (Ct + G - 1) / G;
    const int64_t blocks = 8LL * G * (((int64_t)nb * R * NG + 7) / 8);
    if (blocks > 0x7fffffff) return ECALJ_ERR_DIM;
    grid = dim3((unsigned)blocks, 1, 1);
  } else if (grid.y > 65535 || grid.z > 65535) {
    return ECALJ_ERR_DIM;
  }
  for (int v0 = 0; v0 < nw1; v0 += Q3NW) {
    const int nwg = std::min(Q3NW, nw1 - v0);
    if (wm == 0)
      hipLaunchKernelGGL(sig3m32_kernel, grid, dim3(Q32NT), 0, h->stream, as, as_stride, zp, (int64_t)NCP * NP, wgt,
                         g_stride, n, nw1, v0, nwg, ncol, MP, NS, NP, NCP, v0 > 0 ? 1 : 0, cz, cz_stride, G, nb);
    else if (wm == 1)
      hipLaunchKernelGGL(sig3m_kernel<1>, grid, dim3(512), 0, h->stream, as, as_stride, zp, (int64_t)NCP * NP, wgt,
                         g_stride, n, nw1, v0, nwg, ncol, MP, NS, NP, NCP, v0 > 0 ? 1 : 0, cz, cz_stride, G, nb);
    else
      hipLaunchKernelGGL(sig3m_kernel<2>, grid, dim3(512), 0, h->stream, as, as_stride, zp, (int64_t)NCP * NP, wgt,
                         g_stride, n, nw1, v0, nwg, ncol, MP, NS, NP, NCP, v0 > 0 ? 1 : 0, cz, cz_stride, G, nb);
    ECALJ_HIP(hipGetLastError());
  }
  return ECALJ_OK;
}

// The four-product frequency-inner form: W planes as split by sigbf3.hip
// (4-plane layout, as_stride elements per k), padded zmel as for sig3m.
ecalj_status sigv4_run(ecalj_ctx* h, bool xcd, int nb, int n, int ncol, int nw1, const unsigned short* as,
                       int64_t as_stride, const cf* zp, const double* wgt, int64_t g_stride, cf* cz,
                       int64_t cz_stride) {
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NCP = pad_to(ncol, QBN_MAX), NS = NP / QBK;
  const int R = MP / QBM, Ct = NCP / QV4BN;
  dim3 grid(Ct, R, nb);
  int G = 0;
  if (xcd) {
    G = sig3m_group(Ct);
    const int64_t NG = (Ct + G - 1) / G;
    const int64_t blocks = 8LL * G * (((int64_t)nb * R * NG + 7) / 8);
    if (blocks > 0x7fffffff) return ECALJ_ERR_DIM;
    grid = dim3((unsigned)blocks, 1, 1);
  } else if (grid.y > 65535 || grid.z > 65535) {
    return ECALJ_ERR_DIM;
  }
  for (int v0 = 0; v0 < nw1; v0 += QV4NW) {
    const int nwg = std::min(QV4NW, nw1 - v0);
    hipLaunchKernelGGL(sigv4_kernel, grid, dim3(512), 0, h->stream, as, as_stride, zp, (int64_t)NCP * NP, wgt,
                       g_stride, n, nw1, v0, nwg, ncol, MP, NS, NP, NCP, v0 > 0 ? 1 : 0, cz, cz_stride, G, nb);
    ECALJ_HIP(hipGetLastError());
  }
  return ECALJ_OK;
}

// Padded complex64 zmel alone (the four-product form's B operand)
ecalj_status sig_pad_z(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, cf* zp) {
  const int NP = pad_to(n, QBK), NCP = pad_to(ncol, QBN_MAX);
  const int64_t totz = (int64_t)NCP * NP * nb;
  const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
  hipLaunchKernelGGL(pad_z_kernel, dim3(bz), dim3(256), 0, h->stream, zmel, z_stride, n, ncol, NP, NCP, zp,
                     (int64_t)NCP * NP, nb);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// Grouped form: the operands of nk k-points (zmel packed per k, n x ncol[k];
// the padded zmel of k at zp + zoff[k] with pad_to(max(ncol[k], 1), 128)
// columns), then the grouped contraction.
ecalj_status sig3m_grouped_prep(ecalj_ctx* h, int nk, const int* ncol, const cf* zmel, int n, int nw1, const cd* wc,
                                int64_t ldw, int64_t w_stride, unsigned short* as, cf* zp) {
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NS = NP / QBK;
  const int64_t as_stride = (int64_t)sig3m_a_elems_per_k(n, nw1);
  ECALJ_TRY(split_w6_launch(h, wc, ldw, w_stride, n, nw1, nk, MP, NS, as, as_stride));
  int64_t zo = 0, co = 0;
  for (int k = 0; k < nk; ++k) {
    const int ncp = pad_to(std::max(ncol[k], 1), QBN_MAX);
    if (ncol[k] > 0) {
      const int64_t totz = (int64_t)ncp * NP;
      const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
      hipLaunchKernelGGL(pad_z_kernel, dim3(bz), dim3(256), 0, h->stream, zmel + co, (int64_t)0, n, ncol[k], NP, ncp,
                         zp + zo, (int64_t)0, 1);
      ECALJ_HIP(hipGetLastError());
    }
    zo += (int64_t)ncp * NP;
    co += (int64_t)n * ncol[k];
  }
  return ECALJ_OK;
}

ecalj_status sig3m_grouped_run(ecalj_ctx* h, int wm, int nk, const int* ncol, int n, int nw1,
                               const unsigned short* as, const cf* zp, const double* wgt, cf* cz) {
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NS = NP / QBK, R = MP / QBM;
  const int64_t as_stride = (int64_t)sig3m_a_elems_per_k(n, nw1);
  std::vector<Sig3Group> grp(nk);
  std::vector<int> seq;  // (k, row tile, column tile), column tile fastest
  int64_t zo = 0, go = 0, co = 0;
  for (int k = 0; k < nk; ++k) {
    grp[k] = Sig3Group{zo, go, co, ncol[k], 0};
    const int ncp = pad_to(std::max(ncol[k], 1), QBN_MAX);
    zo += (int64_t)ncp * NP;
    go += (int64_t)nw1 * ncol[k];
    co += (int64_t)n * ncol[k];
    for (int it = 0; it < R; ++it)
      for (int jt = 0; jt < (ncol[k] + Q3BN - 1) / Q3BN; ++jt) {
        seq.push_back(k);
        seq.push_back(it);
        seq.push_back(jt);
      }
  }
  const int nseq = (int)(seq.size() / 3);
  if (nseq == 0) return ECALJ_OK;
  // grid: a multiple of 8 blocks, at most one per CU; per round each XCD's
  // grid/8 blocks take the next grid/8 entries of seq
  const int grid = std::max(8, std::min(h->ncu, (nseq + 7) / 8 * 8) / 8 * 8);
  const int per = grid / 8, rounds = (nseq + grid - 1) / grid;
  std::vector<int> tiles((size_t)3 * rounds * grid, -1);
  for (int q = 0; q < nseq; ++q) {
    const int m = q / grid, r = q % grid, x = r / per, b = r % per;  // round m, XCD x, its b-th block
    const int pos = m * grid + b * 8 + x;
    std::memcpy(&tiles[3 * (size_t)pos], &seq[3 * (size_t)q], 3 * sizeof(int));
  }
  const int ntiles = rounds * grid;
  std::vector<unsigned char> blob(sizeof(Sig3Group) * nk + sizeof(int) * tiles.size());
  std::memcpy(blob.data(), grp.data(), sizeof(Sig3Group) * nk);
  std::memcpy(blob.data() + sizeof(Sig3Group) * nk, tiles.data(), sizeof(int) * tiles.size());
  unsigned char* d;
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_GTAB, blob, &d));
  const Sig3Group* dg = reinterpret_cast<const Sig3Group*>(d);
  const int* dt = reinterpret_cast<const int*>(d + sizeof(Sig3Group) * nk);
  for (int v0 = 0; v0 < nw1; v0 += Q3NW) {
    const int nwg = std::min(Q3NW, nw1 - v0);
    if (wm == 1)
      hipLaunchKernelGGL(sig3m_grouped_kernel<1>, dim3(grid), dim3(512), 0, h->stream, dt, ntiles, dg, as, as_stride,
                         zp, wgt, n, nw1, v0, nwg, MP, NS, NP, v0 > 0 ? 1 : 0, cz);
    else
      hipLaunchKernelGGL(sig3m_grouped_kernel<2>, dim3(grid), dim3(512), 0, h->stream, dt, ntiles, dg, as, as_stride,
                         zp, wgt, n, nw1, v0, nwg, MP, NS, NP, v0 > 0 ? 1 : 0, cz);
    ECALJ_HIP(hipGetLastError());
  }
  return ECALJ_OK;
}

}  // namespace ecalj
