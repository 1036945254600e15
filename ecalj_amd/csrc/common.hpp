// common.hpp -- shared types and helpers for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ecalj_gpu.h"

namespace ecalj {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct cd { double re, im; };   // complex128, interleaved
struct cf { float re, im; };    // complex64, interleaved

// Two-level batch addressing: batch b -> lo = b % div, hi = b / div,
// offset = lo * s_lo + hi * s_hi (elements).
struct Batch2 {
  int64_t s_lo, s_hi;
  __device__ __host__ inline int64_t off(int b, int div) const {
    return (int64_t)(b % div) * s_lo + (int64_t)(b / div) * s_hi;
  }
};

#define ECALJ_HIP(call)                                        \
  do {                                                         \
    hipError_t e_ = (call);                                    \
    if (e_ != hipSuccess) return ECALJ_ERR_HIP;                \
  } while (0)

#define ECALJ_TRY(call)                                        \
  do {                                                         \
    ecalj_status s_ = (call);                                  \
    if (s_ != ECALJ_OK) return s_;                             \
  } while (0)

// GEMM epilogue variants.
enum EpiKind { EPI_STORE = 0, EPI_EPS = 1, EPI_GJ = 2 };

// Accumulator layouts of the 16x16 MFMA output tiles (checked on gfx950):
//   f64 16x16x4:            lane l, element v -> row (l/16) + 4v, column l%16
//   f32 16x16x4, bf16 x32:  lane l, element v -> row 4(l/16) + v, column l%16
// Operands of all three: A(i,k) in lane i + 16*(k/kpl), B(k,j) in lane j + 16*(k/kpl),
// kpl = k values per lane (1 for f64/f32 x4, 8 for bf16 x32).
__device__ inline int acc_row_f64(int lane, int v) { return (lane >> 4) + 4 * v; }
__device__ inline int acc_row_f32(int lane, int v) { return 4 * (lane >> 4) + v; }
__device__ inline int acc_col(int lane, int v) { (void)v; return lane & 15; }

}  // namespace ecalj
