// sph_mfma.hpp - BF16 matrix-core (MFMA) operand layouts for gfx950.
//
// v_mfma_f32_32x32x16_bf16: D(32x32, f32) = A(32x16, bf16) * B(16x32, bf16) + C
//   A: lane l supplies row (l % 32), k = 8 * (l / 32) + [0..7]  (8 bf16)
//   B: lane l supplies col (l % 32), k = 8 * (l / 32) + [0..7]  (8 bf16)
//   C/D: lane l holds col (l % 32) and 16 rows
//        row(i) = 8 * (i / 4) + 4 * (l / 32) + (i % 4),  i = 0..15
// v_mfma_f32_16x16x32_bf16: D(16x16) = A(16x32) * B(32x16) + C
//   A: lane l supplies row (l % 16), k = 8 * (l / 16) + [0..7]
//   B: lane l supplies col (l % 16), k = 8 * (l / 16) + [0..7]
//   C/D: lane l holds col (l % 16), rows 4 * (l / 16) + [0..3]
// The layouts are checked against a host GEMM by sph_mfma_selftest.
#pragma once

#include <hip/hip_runtime.h>

namespace sph {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int mfma32_row(int i, int lane) {
    return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
}

__device__ __forceinline__ f32x16 mfma32(const bf16x8& a, const bf16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace sph
