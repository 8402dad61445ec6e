// sph_mlp.hip - learned pair-force surrogate on BF16 matrix cores (gfx950).
//
// (This file currently holds the MFMA layout self-test; the surrogate
// kernels build on the layouts it checks.)
#include <cstring>

#include "sph_internal.hpp"
#include "sph_mfma.hpp"

namespace sph {
namespace {

// Row-major bf16 A (M x K) and B (K x N) in, row-major f32 C (M x N) out,
// through exactly the operand layouts documented in sph_mfma.hpp.
__global__ __launch_bounds__(64) void k_mfma_selftest32(const __bf16* __restrict__ a,
                                                       const __bf16* __restrict__ b,
                                                       float* __restrict__ c) {
    const int l = threadIdx.x;
    bf16x8 av, bv;
    const int row = l % 32, k0 = 8 * (l / 32);
    for (int e = 0; e < 8; ++e) {
        av[e] = a[row * 16 + k0 + e];
        bv[e] = b[(k0 + e) * 32 + row];  // column `row` of B
    }
    f32x16 acc;
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    acc = mfma32(av, bv, acc);
    for (int i = 0; i < 16; ++i) c[mfma32_row(i, l) * 32 + (l % 32)] = acc[i];
}

__global__ __launch_bounds__(64) void k_mfma_selftest16(const __bf16* __restrict__ a,
                                                       const __bf16* __restrict__ b,
                                                       float* __restrict__ c) {
    const int l = threadIdx.x;
    bf16x8 av, bv;
    const int row = l % 16, k0 = 8 * (l / 16);
    for (int e = 0; e < 8; ++e) {
        av[e] = a[row * 32 + k0 + e];
        bv[e] = b[(k0 + e) * 16 + row];
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma16(av, bv, acc);
    for (int i = 0; i < 4; ++i) c[(4 * (l / 16) + i) * 16 + (l % 16)] = acc[i];
}

}  // namespace
}  // namespace sph

extern "C" {

// One matrix-core GEMM through the documented operand layouts.  variant 0:
// A 32x16, B 16x32, C 32x32; variant 1: A 16x32, B 32x16, C 16x16.  A and B
// are row-major bf16 bit patterns, C row-major float.  Synchronous.
int sph_mfma_selftest(int variant, const uint16_t* a, const uint16_t* b, float* c) {
    if (!a || !b || !c || (variant != 0 && variant != 1)) return SPH_ERR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPH_ERR_NO_DEVICE;
    const size_t na = 512, nb = 512, nc = variant == 0 ? 1024 : 256;
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    hipError_t e = hipMalloc(&da, na * 2);
    if (e == hipSuccess) e = hipMalloc(&db, nb * 2);
    if (e == hipSuccess) e = hipMalloc(&dc, nc * 4);
    if (e == hipSuccess) e = hipMemcpy(da, a, na * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, nb * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (variant == 0)
            hipLaunchKernelGGL(sph::k_mfma_selftest32, dim3(1), dim3(64), 0, 0, (const __bf16*)da,
                               (const __bf16*)db, (float*)dc);
        else
            hipLaunchKernelGGL(sph::k_mfma_selftest16, dim3(1), dim3(64), 0, 0, (const __bf16*)da,
                               (const __bf16*)db, (float*)dc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(c, dc, nc * 4, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    return e == hipSuccess ? SPH_OK : SPH_ERR_DEVICE;
}

}  // extern "C"
