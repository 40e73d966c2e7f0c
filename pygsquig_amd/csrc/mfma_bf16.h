// mfma_bf16.h -- bf16 conversion and the gfx950 bf16 MFMA used by the
// closure network's convolution kernels.
//
// v_mfma_f32_32x32x16_bf16: D(32x32, f32) += A(32x16, bf16) x B(16x32, bf16).
// Lane l of the 64-lane wave supplies A[l % 32][8 (l / 32) + e] and
// B[8 (l / 32) + e][l % 32] for e = 0..7 (one 16-byte register quad each) and
// holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] in accumulator r = 0..15.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned int u32x4 __attribute__((vector_size(16)));  // 8 bf16
typedef float f32x16 __attribute__((vector_size(64)));        // 32x32 MFMA tile

__device__ __forceinline__ f32x16 mfma_32x32x16(u32x4 a, u32x4 b, f32x16 c) {
  typedef __bf16 bf16x8_v __attribute__((ext_vector_type(8)));
  typedef float f32x16_v __attribute__((ext_vector_type(16)));
  f32x16_v r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_v, a),
                                                       __builtin_bit_cast(bf16x8_v, b),
                                                       __builtin_bit_cast(f32x16_v, c), 0, 0, 0);
  return __builtin_bit_cast(f32x16, r);
}

// Two floats -> packed bfloat16 pair (low half first), round to nearest
// even: one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t bf16x2_rne(float lo, float hi) {
  typedef __bf16 bf16x2_v __attribute__((ext_vector_type(2)));
  bf16x2_v v;
  v[0] = (__bf16)lo;
  v[1] = (__bf16)hi;
  return __builtin_bit_cast(uint32_t, v);
}
