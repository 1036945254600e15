// bf16split.hpp -- two FP32 values to bf16 hi + lo pairs in registers.
#pragma once
#include <hip/hip_runtime.h>

namespace ecalj {

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// The bf16 pairs (-1, 0) and (0, -1), kept in scalar registers the compiler
// cannot see into.  As a constant operand of v_dot2c_f32_bf16 the compiler
// encodes (-1, 0) as the inline constant -1.0, which the instruction reads as
// the FP32 bit pattern 0xbf800000, i.e. the pair (0, -1): the residual of the
// first value would then subtract the second value's hi part.
struct SplitMasks {
  bf16x2_t m0, m1;
};
__device__ __forceinline__ SplitMasks split_masks() {
  unsigned a, b;
  asm volatile("s_mov_b32 %0, 0xbf80" : "=s"(a));
  asm volatile("s_mov_b32 %0, 0xbf800000" : "=s"(b));
  return {__builtin_bit_cast(bf16x2_t, a), __builtin_bit_cast(bf16x2_t, b)};
}

struct HiLo2 {
  unsigned hi, lo;  // packed bf16 pairs
};

// hi = bf16(x) and lo = bf16(x - hi), round to nearest each.  The residual
// x - hi is exact in FP32 and is formed by one v_dot2c_f32_bf16 per value,
// hi . (-1, 0) + x0 and hi . (0, -1) + x1, instead of unpacking hi to FP32
// and subtracting (two instructions per value).
__device__ __forceinline__ HiLo2 split2_bf16(float x0, float x1, const SplitMasks& mk) {
  const bf16x2_t h = __builtin_convertvector((f32x2_t){x0, x1}, bf16x2_t);
  const float e0 = __builtin_amdgcn_fdot2_f32_bf16(h, mk.m0, x0, false);
  const float e1 = __builtin_amdgcn_fdot2_f32_bf16(h, mk.m1, x1, false);
  return {__builtin_bit_cast(unsigned, h),
          __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2_t){e0, e1}, bf16x2_t))};
}

// The same split with the residual by unpacking hi and subtracting (a shift,
// an AND and two subtractions).  Within one wave that issues MFMAs between
// these instructions it costs less: v_dot2c does not overlap a running
// 32x32x16 MFMA the way plain VALU instructions do (measured: the split of
// 19 value pairs between 72 MFMAs added 27% with v_dot2c, 11% this way).
// The subtraction is not contracted with the multiplication that forms x in
// the caller, so both forms split the same rounded x bit for bit.
__device__ __forceinline__ HiLo2 split2_bf16_alu(float x0, float x1) {
#pragma clang fp contract(off)
  const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2_t){x0, x1}, bf16x2_t));
  const float e0 = x0 - __uint_as_float(h << 16), e1 = x1 - __uint_as_float(h & 0xffff0000u);
  return {h, __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2_t){e0, e1}, bf16x2_t))};
}

}  // namespace ecalj
