// sph_mlp.hpp - learned pair-force surrogate: parameter layout and the
// launches the device contexts use (kernels in sph_mlp.hip).
#pragma once

#include "sph_internal.hpp"

namespace sph {
namespace mlp {

constexpr int kHid = 64;       // hidden units of both pair layers
constexpr int kFeat = 16;      // edge features incl. the constant and padding
constexpr int kBiasFeat = 5;   // feature fixed at 1: column 5 of W1 is its bias
// W1 columns that carry a feature: 0..4 (position, distance, rho_i) and
// 8..11 (velocity, rho_j); 5 is the bias, the rest are padding.
__host__ __device__ constexpr bool is_feature_column(int k) { return k < kBiasFeat || (k >= 8 && k < 12); }
constexpr int kOut = 3;
constexpr int kNode = 3;
constexpr int kW3Cols = kHid + kNode;                 // 67
constexpr int kOffW1 = 0;                             // [kHid][kFeat]
constexpr int kOffW2 = kOffW1 + kHid * kFeat;         // [kHid out][kHid in]
constexpr int kOffB2 = kOffW2 + kHid * kHid;          // [kHid]
constexpr int kOffW3 = kOffB2 + kHid;                 // [kOut][kW3Cols]
constexpr int kOffB3 = kOffW3 + kOut * kW3Cols;       // [kOut]
constexpr int kNumParams = kOffB3 + kOut;
constexpr int kPairParams = kOffW3;                   // W1, W2, b2
constexpr int kNodeParams = kNumParams - kPairParams; // W3, b3
static_assert(kNumParams == SPH_MLP_NUM_PARAMS, "parameter layout");
constexpr uint32_t kRedChunks = 32;  // row chunks of the gradient reduction

}  // namespace mlp

// Packed BF16 operand fragments (16 bytes per lane each), rebuilt from the
// FP32 master weights by mlp_pack.
struct MlpPacked {
    void* w1f = nullptr;   // 2 x 64 fragments: W1 as A (and B) operand
    void* w2f = nullptr;   // 2 x 4 x 64: W2 rows, hidden units in the pair layer's k order
    void* w2tf = nullptr;  // 2 x 4 x 64: W2 columns, for the backward product with W2
};
size_t mlp_w1f_bytes();
size_t mlp_w2f_bytes();

// The particle array the surrogate reads (sorted order of its last sort).
struct MlpState {
    uint32_t n = 0;
    const float4* pos = nullptr;
    const float4* vel = nullptr;
    const float2* rhop = nullptr;
    const uint32_t* sid = nullptr;  // batch only
    const uint32_t* nbr = nullptr;  // neighbour lists, group-major (nbr_group_index), kNbrCap per slot
    const uint32_t* nnb = nullptr;  // neighbour counts
};

struct MlpTrainBuffers {
    float* h = nullptr;          // [n][kHid] neighbour sums
    float4* out = nullptr;       // [n] (network output, neighbour count)
    float* dh = nullptr;         // [n][kHid] loss gradient w.r.t. the sums
    float* part_pair = nullptr;  // [mlp_pair_grad_waves(n)][kPairParams]
    float* part_node = nullptr;  // [mlp_node_blocks(n)][kNodeParams]
    double* loss_part = nullptr; // [mlp_node_blocks(n)]
    double* chunks = nullptr;    // [kRedChunks][kNumParams + 1] reduction stage
    float* grad = nullptr;       // [kNumParams]
    double* loss = nullptr;      // [1]
};
uint32_t mlp_pair_grad_waves(uint32_t n);
uint32_t mlp_node_blocks(uint32_t n);

hipError_t mlp_pack(const float* params, const MlpPacked& pk, hipStream_t s);

// Surrogate on `st`: accel_out[i] = (a_i, 0) and vel_out[i] = vel[i] + a_i dt
// (dt of the slot's scene from dt_dev); either output may be null.
hipError_t mlp_infer(const LaunchCfg& c, const sph_mlp_norm& nm, const float* params,
                     const MlpPacked& pk, const MlpState& st, float4* vel_out, float4* accel_out,
                     const float* dt_dev, hipStream_t s);

// Loss and gradient against targets: the window sum of force increments per
// original index (target_orig), divided by (window steps x time_step of the
// slot's scene).  pid maps slots of `st` to original indices.
hipError_t mlp_gradient(const LaunchCfg& c, const sph_mlp_norm& nm, const float* params,
                        const MlpPacked& pk, const MlpState& st, const uint32_t* pid,
                        const float4* target_orig, uint32_t window, const MlpTrainBuffers& tb,
                        hipStream_t s);

// Adam on the FP32 master weights (bias corrections bc1 = 1/(1-b1^t),
// bc2 = 1/(1-b2^t) from the host), then the BF16 repack.
hipError_t mlp_adam(float* params, float* m, float* v, const float* grad, float lr, float b1,
                    float b2, float eps, float bc1, float bc2, const MlpPacked& pk, hipStream_t s);

// acc[pid[k]] = (first ? 0 : acc[pid[k]]) + dvel[k]
hipError_t mlp_accumulate_original(uint32_t n, const uint32_t* pid, const float4* dvel, float4* acc,
                                   bool first, hipStream_t s);

// out[0] = mean_k |rhop[k].x * inv_rho0 - 1| in FP64 (partials: 256 doubles).
// With counters: counters[first .. first + ncopy) (ncopy <= 256) are copied
// to cnt_out in the same launch.
hipError_t mlp_density_deviation(uint32_t n, const float2* rhop, float inv_rho0, double* partials,
                                 double* out, hipStream_t s, const uint32_t* counters = nullptr,
                                 uint32_t first = 0, uint32_t ncopy = 0, uint32_t* cnt_out = nullptr);

// dst[k] = src[k] * factor for nscenes floats (coarse time steps).
hipError_t mlp_scale_dt(uint32_t nscenes, const float* src, float factor, float* dst, hipStream_t s);

}  // namespace sph
