// sph_internal.hpp - declarations shared by the HIP translation units of the
// WCSPH step library.  Not part of the public C ABI (include/pibiti_sph.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pibiti_sph.h"
#include "sph_derive.hpp"

namespace sph {

constexpr uint32_t kEmptyCell = 0xFFFFFFFFu;
// Neighbour-list capacity per particle (learned surrogate).
constexpr uint32_t kNbrCap = 128;  // a multiple of 4 (rows are written as uint4)

// Counters written by the step kernels (device memory, uint32 each).
enum Counter : int {
    kCntCapCells = 0,   // cells with more than max_par_in_cell particles
    kCntOutOfGrid = 1,  // particles whose own cell index was clamped
    kCntNbrOverflow = 2,  // particles with more than kNbrCap listed neighbours
    kCntPairs = 3,        // listed neighbour pairs (min(count, kNbrCap) summed)
    kCntSkinMoved = 4,    // particles farther than the skin from their position at the last sort
    kCntFullSort = 5,     // nonzero: this step's order comes from the full radix sort
    kCntNaN = 6,          // particles with a NaN coordinate (the walks then skip no cell)
    kCntNum = 7
};

// Per-scene FP64 diagnostics: stats[scene * kNumStats + k].
enum Stat : int {
    kPMass = 0,
    kPKin = 1,
    kPPot = 2,
    kPVmax = 3,
    kPNum = 4,     // number of reduced quantities
    kSDt = 4,      // time step picked for the next step
    kSClock = 5,   // simulated clock
    kNumStats = 8
};

// ---- radix sort (sph_sort.hip) --------------------------------------------
struct RadixSortPlan {
    uint32_t n = 0;
    uint32_t num_tiles = 0;
    uint32_t* block_hist = nullptr;  // per-block digit histograms of every pass
    uint32_t* digit_off = nullptr;   // exclusive digit offsets per pass
    uint64_t* status = nullptr;      // look-back words, one per (tile, digit)
    uint64_t* ticket = nullptr;      // tile ticket counter (monotone)
    uint32_t* err = nullptr;         // set when a look-back wait hit its limit
    uint32_t* done = nullptr;        // histogram blocks finished (the last one scans; zero between sorts)
};

// Bytes of device scratch the sort of n keys needs, and the plan carved out
// of such a buffer (clears the look-back state; not graph-capturable).
size_t radix_scratch_bytes(uint32_t n);
uint32_t radix_num_tiles(uint32_t n);
hipError_t radix_plan_init(RadixSortPlan& plan, uint32_t n, void* scratch, hipStream_t stream);

// Stable LSD sort of (keys, vals) by the low `bits` bits of keys.  keys0/vals0
// hold the input; keys1/vals1 are the ping-pong buffers.  On return
// *keys_out/*vals_out point at whichever pair holds the sorted result.
// With a gate, every kernel of the sort returns at once unless *gate != 0
// (a device word written by an earlier launch on the stream).
// Graph-capturable (no host-side state changes between launches).
hipError_t radix_sort_pairs(const RadixSortPlan& plan, uint32_t bits, uint32_t* keys0,
                            uint32_t* vals0, uint32_t* keys1, uint32_t* vals1,
                            uint32_t** keys_out, uint32_t** vals_out, hipStream_t stream,
                            const uint32_t* gate = nullptr);

// ---- incremental re-sort (sph_resort.hip) ----------------------------------
// Between steps most particles keep their cell.  The step's new (cell, index)
// order is built from the last one: the integrate kernel marks the slots
// whose cell changed (movers) and counts them into their new cells; a scan
// over the cells gives the new cell ranges; each slot's new position is its
// new cell's start plus the number of particles of that cell with a smaller
// original index (stayers counted from the old ranges and the mover bits,
// movers from the cell's list of incoming indices).  The cell tables are
// padded to a multiple of four cells.  The result equals the
// stable radix sort of (cell, index) exactly.  The full radix sort runs
// instead (decided on the device) when a cell receives more than
// in_cap movers or holds more than kMaxCellCount particles.  in_cap is
// kMaxIncoming, less where the grid is so large for its particles that the
// incoming lists (in_cap words per cell) would pass kInListBudget (0: every
// step with a mover takes the full sort).
constexpr uint32_t kMaxIncoming = 32;
constexpr uint64_t kInListBudgetMin = 64ull << 20;   // bytes
constexpr uint64_t kInListBudgetPerParticle = 128;   // bytes
constexpr uint32_t kDirectTiles = 2048;  // above this many scan tiles, k_tile_scan sums the tiles first
constexpr uint32_t kScanTile = 1024;  // cells per block of the cell scan
constexpr uint32_t kMaxCellCount = 4096;

struct ResortPlan {
    uint32_t n = 0, num_cells = 0, num_tiles = 0;
    uint32_t in_cap = 0;         // incoming-list entries per cell
    uint32_t* tile_start = nullptr;  // first slot of each tile (only above kDirectTiles tiles)
    uint32_t* valid = nullptr;   // device word: slot order, skey and the current cell ranges agree
    uint32_t* in_cnt[2] = {nullptr, nullptr};  // movers into each cell; steps alternate (parity)
    uint32_t* inlist = nullptr;  // original indices of the movers, in_cap per cell
    uint64_t* moved = nullptr;   // one bit per slot: the cell changed this step
    uint32_t* tile_total[2] = {nullptr, nullptr};  // particles per tile of kScanTile cells in the last order
    uint32_t* tile_delta[2] = {nullptr, nullptr};  // movers in less movers out per tile this step
    uint32_t* valid_step = nullptr;  // this step's copy of valid (the last kernel of a step sets valid)
    // Error word, shared with the radix sort's plan: kErrInvalid when a
    // step's order could not be built (results invalid; sticky until the
    // host reads and clears it).
    uint32_t* err = nullptr;
    // Repair flags of the current step: kRepairInputs from the integrate
    // kernel (a slot outside its cell's old range), kRepairCounts from the
    // cell scan (a tile's count disagrees two ways).  Either sends the step
    // through the full sort with the ranges taken from the sorted keys; the
    // place kernel clears the flags, leaves the order invalid for the next
    // step and counts the step in `repairs`.
    uint32_t* repair = nullptr;
    uint32_t* repairs = nullptr;  // steps repaired so far (cleared when the host reads it)
};
constexpr uint32_t kErrInvalid = 1u;
constexpr uint32_t kRepairInputs = 1u;
constexpr uint32_t kRepairCounts = 2u;
size_t resort_scratch_bytes(uint32_t n, uint32_t num_cells);
// Carves the plan out of `scratch`, clears it (not graph-capturable) and
// marks the order invalid.
hipError_t resort_plan_init(ResortPlan& plan, uint32_t n, uint32_t num_cells, void* scratch, hipStream_t s);
// Marks the order invalid (graph-capturable): the next step places every
// particle from scratch.
hipError_t resort_invalidate(const ResortPlan& plan, hipStream_t s);

// ---- step kernels (sph_kernels.hip) ---------------------------------------
struct StepBuffers {
    float4* pos_cur;      // sorted-order state after the last reorder
    float4* vel_cur;
    uint32_t* pid_cur;    // original index of each slot
    uint32_t* sid_cur;    // scene of each slot (batch only, else null)
    float4* pos_alt;
    float4* vel_alt;
    uint32_t* pid_alt;
    uint32_t* sid_alt;
    float2* rhop;         // (density, pressure) per sorted slot
    float4* dvel;         // optional force increment per sorted slot
    const uint32_t* skey; // cell key of each sorted slot (from the last sort)
    uint32_t* skey_alt;   // cell key of each slot of the next order (incremental re-sort)
    uint32_t* key_pid;    // cell key in original-index order (sort input)
    uint32_t* val_pid;    // current slot in original-index order (sort input)
    uint32_t* key_tmp;
    uint32_t* val_tmp;
    uint32_t* cell_start;
    uint32_t* cell_end;
    uint32_t* cell_start_alt;  // cell ranges of the next order (incremental re-sort)
    uint32_t* cell_end_alt;
    uint32_t* counters;
    uint32_t* work;       // persistent-kernel tickets: [ticket, done] per gather kernel
    double* partials;     // per (scene, chunk) FP64 partials
    double* stats;        // kNumStats doubles per scene
    float* dt_dev;        // time step of each scene
    uint32_t* masks;      // force pair masks of the density kernel (kMaskWords per slot), or null
    uint4* pairs;         // pair-entry lists (group-major), or null: see kPairSlotShift
    uint32_t* pair_cnt;   // entries | pairs << 16 per slot
    float4* rec;          // force records, 2 float4 per slot
    float* xyz;           // x, y, z per slot, packed (the pair-entry densities' candidates), or null
};

// Force pair masks: 27 cells x 16 bits in 14 words per slot (word w of slot
// i at masks[w n + i]), usable while every cell holds at most 16 candidates.
constexpr uint32_t kMaskWords = 14;
constexpr uint32_t kMaskCap = 16;
// Pair entries from the reference-order density walk to its force kernel:
// (first slot << 5) | (new cell << 4) | 4-bit mask of the force pairs among
// four consecutive candidates of one cell, in walk order; the new-cell flag
// marks a cell's first entry.  Group-major uint4 lists as in FAST mode.
constexpr uint32_t kPairSlotShift = 5;
constexpr uint32_t kPairNewCell = 1u << 4;
constexpr uint32_t kPairRing = 4;  // entries per lane in the density kernel's LDS ring

// What a launch covers: one scene (params by value) or a batch of scenes
// (device parameter table + per-slot scene ids + static slot ranges).
struct LaunchCfg {
    uint32_t n = 0;             // particles over all scenes
    uint32_t num_cells = 0;     // cells over all scenes
    uint32_t key_bits = 0;
    uint32_t nscenes = 1;
    uint32_t max_cap = 16;      // largest max_par_in_cell of any scene
    uint32_t max_scene_n = 0;   // largest scene
    uint32_t precision = 0;     // shared by every scene
    uint32_t tune[4] = {0, 0, 0, 0};
    uint32_t physics = 0;       // union of the scenes' sph_params::physics bits
    uint32_t colour = 0;        // some scene has a colour mode
    const sph_params* tab = nullptr;      // device table (batch) or null
    const uint32_t* scene_off = nullptr;  // device: first slot of each scene, nscenes + 1
    sph_params one{};                     // the scene (one-scene launches)
};

// Launch tuning words (LaunchCfg::tune, from sph_params::tune):
//   tune[0]  force kernel: 1 = fused candidate/pair loop, 2 = pair masks
//            written by the density kernel (max_par_in_cell <= 16; also the
//            default when every tune word is 0), 16/24/32/48 = neighbour-list
//            entries per lane filled by the force kernel's own candidate
//            tests (the default otherwise: 24)
//   tune[1]  gather grid: 0 = one thread per slot over all slots, else a
//            persistent grid of tune[1] blocks pulling 256-slot tiles
//   tune[3]  gather variant (tune[1] == 0): 0 = cell tables and candidate
//            positions read per lane, XCD-ordered tiles; 1 = candidate
//            positions staged in LDS, XCD-ordered; 2 = neighbour-row cell
//            tables staged in LDS, XCD-ordered; 3 = per lane, tiles in
//            launch order
uint32_t stats_chunks(const LaunchCfg& c);
size_t stats_partials_count(const LaunchCfg& c);

// The integration in place, keeping the slot order (the surrogate's steps
// between sorts).
hipError_t launch_integrate(const LaunchCfg& c, const StepBuffers& b, hipStream_t s);
// The step's integration for the incremental re-sort: the new cell key of
// each slot into nkey (slot order), the mover bits and per-cell mover counts
// into the plan (every slot is a mover while the plan's order is invalid).
hipError_t launch_integrate_moves(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, int parity,
                                  uint32_t* nkey, hipStream_t s);
// The rest of the re-sort after launch_integrate_moves (nkey = key_tmp),
// with the same parity (the step's mover counts; parity ^ 1 is cleared for
// the next step).  resort_order: new cell ranges into cell_start_alt /
// cell_end_alt.  resort_apply: the state in the new order into the *_alt
// buffers and skey_alt (through key_pid/val_pid, key_tmp/val_tmp and the
// radix plan on a full-sort step); marks the plan's order valid.
hipError_t resort_order(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, int parity, hipStream_t s);
hipError_t resort_apply(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, int parity,
                        const RadixSortPlan& sort, hipStream_t s);
// Particle colours and dye after the force update: b holds the step's state
// before the update (pos_cur, vel_cur, rhop, dvel, pid_cur).
hipError_t launch_colour(const LaunchCfg& c, const StepBuffers& b, float4* colour, float* dye, hipStream_t s);
hipError_t launch_density(const LaunchCfg& c, const StepBuffers& b, hipStream_t s);
// Density plus each slot's neighbour list, group-major: entries 4g .. 4g + 3
// of slot k are nbr[4 (g n + k) .. 4 (g n + k) + 3] (g < kNbrCap / 4), and
// nnb[k] = neighbour count (may exceed kNbrCap: counted in kCntNbrOverflow).
hipError_t launch_density_lists(const LaunchCfg& c, const StepBuffers& b, uint32_t* nbr, uint32_t* nnb,
                                hipStream_t s);
// Density, pressure and neighbour lists (as launch_density_lists) of one
// scene whose slots moved since the last sort (sph_skin.hip): walks every
// cell within (1 + skin) h of each slot's current position using the cell
// ranges of that sort, counts the pairs in kCntPairs and the slots farther
// than skin h from their position at the sort (bpos) in kCntSkinMoved; with
// none of those, the lists hold exactly the pairs now within h.
hipError_t launch_density_reach(const LaunchCfg& c, const StepBuffers& b, float skin, const float4* bpos,
                                uint32_t* nbr, uint32_t* nnb, hipStream_t s);
hipError_t launch_force(const LaunchCfg& c, const StepBuffers& b, float4* vel_out, bool keep_dvel,
                        hipStream_t s);
// The scene's force terms beyond the pair force (collider, accelerators,
// rotor, height map) after a force kernel left the pair increment in vel_out;
// finishes the velocity update (launch_force calls it when c.physics).
hipError_t launch_extra_terms(const LaunchCfg& c, const StepBuffers& b, float4* vel_out, bool keep_dvel,
                              hipStream_t s);
// SPH_PREC_FAST (sph_fast.hip): density + EOS writing each slot's force
// pair entries (group-major uint4 lists of fast_list_groups(max_par_in_cell)
// groups, counts in nnb) and force record (rec: 2 float4 per slot);
// launch_force then walks those entries with the fast arithmetic.
constexpr uint32_t kFastMaxCap = 64;
uint32_t fast_list_groups(uint32_t max_par_in_cell);
// Whether the step's gathers go through the pair entries (then each slot's
// count word holds its force pairs): FAST always, the other orders while
// the lists exist and the tune setting does not pick another path.
bool pair_entries_in_use(const LaunchCfg& c, const StepBuffers& b);

hipError_t launch_density_fast(const LaunchCfg& c, const StepBuffers& b, float4* rec, uint4* nbr, uint32_t* nnb,
                               hipStream_t s);
// FP64 per-scene reductions of (mass, kinetic, potential, max speed) over
// pos_cur/vel_cur, then per-scene finalisation: stats, the next time step
// (fixed, CFL, clipped at t_end) into dt_dev, and the clock when advance_clock.
hipError_t launch_stats(const LaunchCfg& c, const StepBuffers& b, bool advance_clock,
                        hipStream_t s);
// Largest particle count of one cell (atomicMax into *max_count).
hipError_t launch_cell_stats(const LaunchCfg& c, const StepBuffers& b, uint32_t* max_count,
                             hipStream_t s);

// Host <-> device record transfer through the pid permutation.
hipError_t launch_scatter_to_original(uint32_t n, const uint32_t* pid, const void* src,
                                      void* dst, int record_floats, int src_stride_floats,
                                      hipStream_t s);
hipError_t launch_update_from_original(uint32_t n, const uint32_t* pid, const float4* staging,
                                       uint32_t first, uint32_t count, float4* dst, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* dst, uint32_t value, size_t count, hipStream_t s);
// Every device allocation of the library: hipMalloc, or hipErrorOutOfMemory
// when the allocation-failure hook (sph_debug_fail_alloc) fires.
hipError_t device_malloc(void** ptr, size_t bytes);
hipError_t launch_iota_u32(uint32_t* dst, uint32_t n, hipStream_t s);
// sid[k] = scene of original index pid[k]; scene_off holds the first
// original index of each scene (scenes are contiguous in original order).
hipError_t launch_scene_ids(uint32_t n, uint32_t nscenes, const uint32_t* scene_off,
                            const uint32_t* pid, uint32_t* sid, hipStream_t s);
}  // namespace sph
