// sph_internal.hpp - declarations shared by the HIP translation units of the
// WCSPH step library.  Not part of the public C ABI (include/pibiti_sph.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pibiti_sph.h"

namespace sph {

constexpr uint32_t kEmptyCell = 0xFFFFFFFFu;

// Counters written by the step kernels (device memory, uint32 each).
enum Counter : int {
    kCntCapCells = 0,   // cells with more than max_par_in_cell particles
    kCntOutOfGrid = 1,  // particles whose own cell index was clamped
    kCntNum = 4
};

// Per-block FP64 partials of the diagnostics reduction.
enum Partial : int { kPMass = 0, kPKin = 1, kPPot = 2, kPVmax = 3, kPNum = 4 };

// ---- radix sort (sph_sort.hip) --------------------------------------------
struct RadixSortPlan {
    uint32_t n = 0;
    uint32_t num_tiles = 0;
    uint32_t* block_hist = nullptr;  // per-block digit histograms of every pass
    uint32_t* digit_off = nullptr;   // exclusive digit offsets per pass
    uint64_t* status = nullptr;      // look-back words, one per (tile, digit)
    uint64_t* ticket = nullptr;      // tile ticket counter (monotone)
    uint32_t* err = nullptr;         // set when a look-back wait hit its limit
};

// Bytes of device scratch the sort of n keys needs, and the plan carved out
// of such a buffer (clears the look-back state; not graph-capturable).
size_t radix_scratch_bytes(uint32_t n);
uint32_t radix_num_tiles(uint32_t n);
hipError_t radix_plan_init(RadixSortPlan& plan, uint32_t n, void* scratch, hipStream_t stream);

// Stable LSD sort of (keys, vals) by the low `bits` bits of keys.  keys0/vals0
// hold the input; keys1/vals1 are the ping-pong buffers.  On return
// *keys_out/*vals_out point at whichever pair holds the sorted result.
hipError_t radix_sort_pairs(const RadixSortPlan& plan, uint32_t bits, uint32_t* keys0,
                            uint32_t* vals0, uint32_t* keys1, uint32_t* vals1,
                            uint32_t** keys_out, uint32_t** vals_out, hipStream_t stream);

// ---- step kernels (sph_kernels.hip) ---------------------------------------
struct StepBuffers {
    float4* pos_cur;      // sorted-order state after the last reorder
    float4* vel_cur;
    uint32_t* pid_cur;    // original index of each slot
    float4* pos_alt;
    float4* vel_alt;
    uint32_t* pid_alt;
    float2* rhop;         // (density, pressure) per sorted slot
    float4* dvel;         // optional force increment per sorted slot
    uint32_t* key_pid;    // cell key in original-index order (sort input)
    uint32_t* val_pid;    // current slot in original-index order (sort input)
    uint32_t* key_tmp;
    uint32_t* val_tmp;
    uint32_t* cell_start;
    uint32_t* cell_end;
    uint32_t* counters;
    double* partials;     // kPNum * max partial blocks
    double* stats;        // kPNum finalised values + dt at [kPNum]
    float* dt_dev;        // time step the integrate/force kernels use
};

// integrate + boundary + hash on the current slots; writes the sort input in
// original-index order and clears cellStart.
hipError_t launch_integrate_hash(const sph_params& p, const StepBuffers& b, hipStream_t s);

// reorder into the alternate buffers + cellStart/cellEnd from the sorted keys.
hipError_t launch_reorder(const sph_params& p, const StepBuffers& b, const uint32_t* sorted_key,
                          const uint32_t* old_slot, hipStream_t s);

// density + EOS on the sorted slots of pos_cur.
hipError_t launch_density(const sph_params& p, const StepBuffers& b, hipStream_t s);

// force + velocity update: reads vel_cur, writes vel_out.
hipError_t launch_force(const sph_params& p, const StepBuffers& b, float4* vel_out, bool keep_dvel,
                        hipStream_t s);

// FP64 partials of (mass, kinetic, potential, max speed) over pos_cur/vel_cur,
// then the single-block finalisation into stats[0..4] and dt_dev.
// stats[kPNum] receives the time step, stats[kPNum + 1] the simulated clock
// (advanced by that step when advance_clock is set).
uint32_t stats_num_blocks(uint32_t n);
hipError_t launch_stats(const sph_params& p, const StepBuffers& b, bool advance_clock,
                        hipStream_t s);

// Largest particle count of one cell (atomicMax into *max_count).
hipError_t launch_cell_stats(const sph_params& p, const StepBuffers& b, uint32_t* max_count,
                             hipStream_t s);

// Host <-> device record transfer through the pid permutation.
hipError_t launch_scatter_to_original(uint32_t n, const uint32_t* pid, const void* src,
                                      void* dst, int record_floats, int src_stride_floats,
                                      hipStream_t s);
hipError_t launch_update_from_original(uint32_t n, const uint32_t* pid, const float4* staging,
                                       uint32_t first, uint32_t count, float4* dst, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* dst, uint32_t value, size_t count, hipStream_t s);
hipError_t launch_iota_u32(uint32_t* dst, uint32_t n, hipStream_t s);

}  // namespace sph
