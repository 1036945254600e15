// sigtile.hpp -- the block tile shared by the bf16x3 contractions: the
// Sigma_c frequency contraction (sigbf3.hip), the chi0 bins (x0bf3.hip) and
// the zmel GEMMs (zmelbf3.hip).  Operands are pre-split bf16 hi/lo planes,
// pre-tiled per 32-wide k slice (rows of 64 bytes, the four 16-byte chunks
// of row r stored in the order chunk ^ (((r >> 1) ^ (r >> 2)) & 3)), moved
// global -> LDS with direct-to-LDS loads and multiplied on the bf16 MFMA.
#pragma once

#include <algorithm>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {

constexpr int QBM = 128, QBK = 32;  // block rows, k per stage
constexpr int QNW = 32;                                // frequencies per launch (weight table in LDS)
constexpr int F8_ROWB = 260;  // fp8 cross-term form: bytes per operand row and 32-wide k slice

__device__ __host__ inline int chunk_slot(int row, int chunk) { return chunk ^ (((row >> 1) ^ (row >> 2)) & 3); }

// ---------------------------------------------------------------- block tile
typedef __attribute__((address_space(3))) void* lds_ptr_t;
typedef __attribute__((address_space(1))) void* gbl_ptr_t;
typedef unsigned short us8q __attribute__((ext_vector_type(8)));
typedef unsigned int u4q __attribute__((ext_vector_type(4)));

template <int BN, int NSTG>
struct alignas(16) SigSmem {
  unsigned short a[NSTG][4][QBM * QBK];  // 8 KB per plane
  unsigned short b[NSTG][4][BN * QBK];   // BN/16 KB per plane
  float wt[QNW][BN];                     // per-frequency weights of the block's columns
};

// s_waitcnt with only vmcnt constrained (gfx9 encoding: vmcnt[3:0] + vmcnt[5:4] at [15:14],
// expcnt [6:4] and lgkmcnt [11:8] left at their maximum)
#define ECALJ_WAIT_VMCNT(N) __builtin_amdgcn_s_waitcnt(((N) & 0xF) | (((N) >> 4) << 14) | (0x7 << 4) | (0xF << 8))

// One block tile of czmelwc: rows I0..I0+QBM of one internal k, columns
// c0..c0+BN; NW waves in 2 x NW/2, wave tile 64 x 2BN/NW; NSTG stages of QBK
// in LDS, NSTG-1 of them in flight.  A and B point at the k's plane sets
// (A: 4 planes of MP x QBK per (v, ks); B: 4 planes of NCP x QBK per ks).
// NW = 8 puts two waves on every SIMD: one wave alone issues a 16x16x32
// bf16 MFMA only every ~30 cycles, two interleave to ~20.
// AREAL: the A operand is real (its imaginary planes are zero and are not
// multiplied): 6 MFMAs per 16x16 tile and k slice instead of 12.
template <int BN, int NSTG, int NW, bool AREAL = false>
__device__ __forceinline__ void sig_tile(SigSmem<BN, NSTG>& s, const unsigned short* __restrict__ Ak,
                                         const unsigned short* __restrict__ Bk, const double* __restrict__ g,
                                         int n, int nw1, int v0, int nwg, int ncol, int MP, int NS, int NCP,
                                         int I0, int c0, int accumulate, cf* __restrict__ C, int mirror = 0,
                                         int64_t ldc = 0) {
  static_assert(NSTG == 2 || NSTG == 3, "two or three stages");
  static_assert(NW == 4 || NW == 8, "four or eight waves");
  constexpr int NT = NW * 64, WN = NW / 2;
  constexpr int TM = 4, TN = BN / (16 * WN);          // 16x16 tiles per wave
  constexpr int ARW = 4 * QBM * QBK * 2 / 1024 / NW;  // 1 KB A runs per wave and stage
  constexpr int BRW = 4 * BN * QBK * 2 / 1024 / NW;   // 1 KB B runs per wave and stage
  static_assert(TN >= 1 && ARW * NW == 4 * QBM * QBK * 2 / 1024 && BRW * NW == 4 * BN * QBK * 2 / 1024,
                "runs split evenly over the waves");
  constexpr int APP = QBM * QBK / 512, BPP = BN * QBK / 512;  // runs per plane
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), wm = wave / WN, wn = wave % WN;
  const int64_t apst = (int64_t)MP * QBK, bpst = (int64_t)NCP * QBK;
  // uniform bases (scalar registers) + the lane's 16 bytes of each 1 KB run, so
  // the stage loads take the scalar-base form with no per-load address VALU
  const unsigned short* A = Ak + (int64_t)I0 * QBK;
  const unsigned short* B = Bk + (int64_t)c0 * QBK;
  const int lane8 = lane * 8;
  const int nst = nwg * NS;

  // weight table: wt[v][c] = wgt(v0+v, c0+c) (0 beyond ncol)
  for (int e = t; e < nwg * BN; e += NT) {
    const int v = e / BN, c = c0 + e % BN;
    s.wt[v][e % BN] = (c < ncol) ? (float)g[(v0 + v) + (int64_t)nw1 * c] : 0.f;
  }
  __syncthreads();

  auto issue = [&](int st, int buf) {
    const int v = v0 + st / NS, ks = st % NS;
    const unsigned short* a = A + ((int64_t)v * NS + ks) * 4 * apst;
    const unsigned short* bb = B + (int64_t)ks * 4 * bpst;
#pragma unroll
    for (int q = 0; q < ARW; ++q) {
      const int run = ARW * wave + q, p = run / APP, sub = run % APP;
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(a + (p * apst + sub * 512) + lane8),
                                       (lds_ptr_t)(&s.a[buf][p][sub * 512]), 16, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < BRW; ++q) {
      const int run = BRW * wave + q, p = run / BPP, sub = run % BPP;
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(bb + (p * bpst + sub * 512) + lane8),
                                       (lds_ptr_t)(&s.b[buf][p][sub * 512]), 16, 0, 0);
    }
  };

  f4 cr[TM][TN], ci[TM][TN], tr[TM][TN], ti[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      cr[i][j] = ci[i][j] = tr[i][j] = ti[i][j] = f4{0, 0, 0, 0};
    }
  // operand offsets in a plane: row r, logical chunk lane>>4, stored at chunk_slot(r, lane>>4);
  // rows are 16-aligned tile rows + (lane & 15), so the slot depends on the lane only
  const int slot = chunk_slot(lane & 15, lane >> 4);
  const int aoff = (wm * TM * 16 + (lane & 15)) * QBK + 8 * slot;
  const int boff = (wn * TN * 16 + (lane & 15)) * QBK + 8 * slot;

  for (int q = 0; q < NSTG - 1; ++q)
    if (q < nst) issue(q, q);
  // frequency loop outside, k-slice loop inside: the inner loop touches only the
  // per-frequency accumulators (MFMA-resident), the weighted totals only at
  // the end of each frequency
  int st = 0;
  for (int vl = 0; vl < nwg; ++vl) {
    for (int ks = 0; ks < NS; ++ks, ++st) {
      const int buf = st % NSTG;
      // stage st landed: the loads of later stages already issued may still be in flight
      if (NSTG == 3 && st + 1 < nst) {
        ECALJ_WAIT_VMCNT(ARW + BRW);
      } else {
        ECALJ_WAIT_VMCNT(0);
      }
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      // the buffer of stage st+NSTG-1 was last read in iteration st-1, which every wave has left
      if (st + NSTG - 1 < nst) issue(st + NSTG - 1, (st + NSTG - 1) % NSTG);

      // B fragments as 32-bit words: the sign flips are then one XOR per word
      u4q brh[TN], brl[TN], bih[TN], bil[TN], nbh[TN], nbl[TN];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int o = boff + j * 16 * QBK;
        brh[j] = *reinterpret_cast<const u4q*>(&s.b[buf][0][o]);
        brl[j] = *reinterpret_cast<const u4q*>(&s.b[buf][1][o]);
        bih[j] = *reinterpret_cast<const u4q*>(&s.b[buf][2][o]);
        bil[j] = *reinterpret_cast<const u4q*>(&s.b[buf][3][o]);
      }
      // A fragments of row tile i (planes re hi, re lo, im hi, im lo) in two
      // register sets.  The LDS stage loads in flight make the compiler drain
      // every outstanding LDS read (lgkmcnt(0)) before the first use of any of
      // them, so tile i+1's reads are issued just after tile i's first MFMA
      // (which waits for tile i's) and the tile's other MFMAs cover them.
      us8q fa[2][4];
      auto read_a = [&](int i, int set) {
        const int o = aoff + i * 16 * QBK;
        fa[set][0] = *reinterpret_cast<const us8q*>(&s.a[buf][0][o]);
        fa[set][1] = *reinterpret_cast<const us8q*>(&s.a[buf][1][o]);
        if constexpr (!AREAL) {
          fa[set][2] = *reinterpret_cast<const us8q*>(&s.a[buf][2][o]);
          fa[set][3] = *reinterpret_cast<const us8q*>(&s.a[buf][3][o]);
        }
      };
      read_a(0, 0);
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        // -b: flip the sign bit of every bf16 (two per 32-bit word)
        nbh[j] = bih[j] ^ 0x80008000u;
        nbl[j] = bil[j] ^ 0x80008000u;
      }
      // one A row tile at a time: its 4 planes meet the TN column tiles
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const us8q* a = fa[i & 1];
        const us8q arh = a[0], arl = a[1];
#define ECALJ_Q_TERM(ACC, AF, BF, J0)                                                                 \
  _Pragma("unroll") for (int j = J0; j < TN; ++j) ACC[i][j] =                                         \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, AF), __builtin_bit_cast(bf8, BF[j]), \
                                              ACC[i][j], 0, 0, 0);
        // re = ar br - ai bi, im = ar bi + ai br; hi*hi + hi*lo + lo*hi of each
        __builtin_amdgcn_sched_barrier(0);
        cr[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, arh), __builtin_bit_cast(bf8, brh[0]),
                                                           cr[i][0], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        if (i + 1 < TM) read_a(i + 1, (i + 1) & 1);
        __builtin_amdgcn_sched_barrier(0);
        ECALJ_Q_TERM(cr, arh, brh, 1)
        ECALJ_Q_TERM(ci, arh, bih, 0)
        ECALJ_Q_TERM(cr, arh, brl, 0)
        ECALJ_Q_TERM(ci, arh, bil, 0)
        ECALJ_Q_TERM(cr, arl, brh, 0)
        ECALJ_Q_TERM(ci, arl, bih, 0)
        if constexpr (!AREAL) {
          const us8q aih = a[2], ail = a[3];
          ECALJ_Q_TERM(cr, aih, nbh, 0)
          ECALJ_Q_TERM(ci, aih, brh, 0)
          ECALJ_Q_TERM(cr, aih, nbl, 0)
          ECALJ_Q_TERM(ci, aih, brl, 0)
          ECALJ_Q_TERM(cr, ail, nbh, 0)
          ECALJ_Q_TERM(ci, ail, brh, 0)
        }
#undef ECALJ_Q_TERM
      }
    }
    // end of frequency: tot += wgt(v,c) * acc
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const float wv = s.wt[vl][wn * TN * 16 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        tr[i][j] += wv * cr[i][j];
        ti[i][j] += wv * ci[i][j];
        cr[i][j] = f4{0, 0, 0, 0};
        ci[i][j] = f4{0, 0, 0, 0};
      }
    }
  }
  const int64_t ld = ldc ? ldc : n;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int I = I0 + wm * TM * 16 + i * 16 + acc_row_f32(lane, e);
        const int c = c0 + wn * TN * 16 + j * 16 + (lane & 15);
        if (I < n && c < ncol) {
          cf* p = C + I + ld * c;
          if (accumulate) {
            const cf o = *p;
            *p = cf{o.re + tr[i][j][e], o.im + ti[i][j][e]};
          } else {
            *p = cf{tr[i][j][e], ti[i][j][e]};
          }
          if (mirror) {  // Hermitian output, off-diagonal tile: C(c, I) += conj
            cf* q = C + c + ld * I;
            const cf o = *q;
            *q = cf{o.re + tr[i][j][e], o.im - ti[i][j][e]};
          }
        }
      }
}

static inline int pad_to(int x, int m) { return (x + m - 1) / m * m; }
constexpr int QBN_MAX = 128;  // columns are padded for the widest variant

}  // namespace ecalj
