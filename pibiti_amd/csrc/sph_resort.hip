// sph_resort.hip - the step's new (cell, index) order built from the last
// one instead of sorting every key again.
//
// The reference re-sorts all particles each step (calcHash, the radix sort
// and reorderData, source/CUDA/System.cu:161-186).  The contract kept: after
// the step, slots hold the particles in ascending (cell key, original index)
// order, cellStart/cellEnd give each cell's slot range (empty cells
// kEmptyCell), and the state is reordered into the new slots.  Between two
// steps a particle moves a small fraction of a cell, so almost every slot
// keeps its cell; only the movers need placing:
//
//   k_integrate_moves  (sph_kernels.hip) the new key of each slot, mover
//                      bits, movers counted into their new cells and their
//                      original indices listed there (up to in_cap per
//                      cell), movers in less movers out per tile of
//                      kScanTile cells; the integrated state is not stored;
//   k_cell_scan        (here) per cell: stayers = old count - movers out
//                      (the old range and the mover bits), new count =
//                      stayers + movers in; exclusive scan over the tile's
//                      cells, the tile's start summed from the earlier
//                      tiles' counts of the last order and their deltas
//                      (by k_tile_scan first on grids of more than
//                      kDirectTiles tiles) ->
//                      new cell ranges; clears the other step's counts;
//                      flags the step for the full sort when a cell takes
//                      more than in_cap movers or holds more than
//                      kMaxCellCount particles, and when the step's inputs
//                      fail a consistency check (below);
//   k_resort_place     (sph_kernels.hip) slot -> new slot: cell start +
//                      stayers of the cell with a smaller index (a stayer:
//                      its rank among the old range's stayers; a mover: a
//                      binary search of the old range) + listed movers in
//                      with a smaller index; integrates the slot's state
//                      again (the same arithmetic) and writes it there with
//                      its index, scene and key.  On a full-sort step it
//                      gathers from the radix sort's output instead;
//   the radix sort     (sph_sort.hip) on the pairs the integrate kernel also
//                      writes, only on a full-sort step: launched every
//                      step before k_resort_place, its kernels return at
//                      once otherwise (so the launch sequence never changes
//                      and graphs replay it).
// While the order is invalid (first step, after the learned surrogate or
// the tuner rearranged the slots) every slot counts as a mover with no old
// ranges, which is a counting sort by cell with the same result.
//
// Consistency checks.  The order is only built upon while its inputs agree:
// the integrate kernel checks that every slot lies in its cell's old range,
// and the cell scan counts each tile two ways (the last order's tile total
// plus the step's deltas, against its own per-cell counts).  A failed check
// sends the step through the full sort, whose place kernel writes every
// occupied cell's range from the sorted keys (the failing tiles' cells are
// written empty by the scan), so the step's results are those of a clean
// sort; the order is then rebuilt from scratch on the next step.  An
// inconsistency that only the place kernel sees (or a sort look-back that
// hits its bound) cannot be repaired within the step: it sets kErrInvalid
// in the error word, which the host reports at its next synchronising call
// and which keeps later steps from building on the order.
#include "sph_device.hpp"

namespace sph {
namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr int kCellItems = 4;
static_assert(kScanTile == kBlock * kCellItems, "k_cell_scan: four cells per thread, one uint4 each");

// Inclusive scan of one value per thread over the block.
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* s_wave) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(v, off);
        v += lane >= off ? u : 0u;
    }
    if (lane == 63) s_wave[w] = v;
    __syncthreads();
    uint32_t add = 0;
    for (int k = 0; k < w; ++k) add += s_wave[k];
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(kBlock) void k_cell_scan(
    uint32_t num_cells, const uint32_t* __restrict__ valid, const uint32_t* __restrict__ old_start,
    const uint32_t* __restrict__ old_end, const uint32_t* __restrict__ in_cnt, uint32_t* __restrict__ in_next,
    const uint64_t* __restrict__ moved, uint32_t* __restrict__ new_start, uint32_t* __restrict__ new_end,
    uint32_t* __restrict__ counters, const uint32_t* __restrict__ tile_total, const uint32_t* __restrict__ tile_delta,
    uint32_t* __restrict__ total_next, uint32_t* __restrict__ delta_next, uint32_t* __restrict__ valid_step,
    uint32_t in_cap, const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ err,
    uint32_t* __restrict__ repair, uint32_t n) {
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_total, s_excl;
    __shared__ uint32_t s_bad;
    const uint32_t t = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    // an order left behind by a failed step is never built upon
    const bool ok = *valid != 0 && (*err & kErrInvalid) == 0;
    // the integrate kernel found a slot outside its cell's old range: the
    // counts below cannot be trusted, the step takes the full sort
    const bool rep_in = (*repair & kRepairInputs) != 0;
    if (tile == 0 && t == 0) *valid_step = ok ? 1u : 0u;  // for k_resort_place, which sets valid
    if (t == 0) s_bad = rep_in ? 1u : 0u;

    // four consecutive cells per thread (16-byte loads; the arrays are
    // padded to a multiple of four cells)
    const uint32_t c0 = tile * kScanTile + (uint32_t)kCellItems * t;
    uint32_t cnt[kCellItems] = {0, 0, 0, 0};
    bool big = false;
    if (c0 < num_cells) {
        const uint4 in4 = *reinterpret_cast<const uint4*>(in_cnt + c0);
        uint4 s4 = make_uint4(kEmptyCell, kEmptyCell, kEmptyCell, kEmptyCell), e4 = make_uint4(0, 0, 0, 0);
        if (ok) {
            s4 = *reinterpret_cast<const uint4*>(old_start + c0);
            e4 = *reinterpret_cast<const uint4*>(old_end + c0);
        }
        // the next step counts its movers into the other array
        *reinterpret_cast<uint4*>(in_next + c0) = make_uint4(0, 0, 0, 0);
        const uint32_t ins[kCellItems] = {in4.x, in4.y, in4.z, in4.w};
        const uint32_t ss[kCellItems] = {s4.x, s4.y, s4.z, s4.w};
        const uint32_t es[kCellItems] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
        for (int j = 0; j < kCellItems; ++j) {
            if (c0 + (uint32_t)j >= num_cells) break;
            uint32_t stay = 0, old = 0;
            if (ss[j] != kEmptyCell) {
                // an old range outside the slots is counted as empty (and
                // the tile's two-way count below then disagrees)
                if (es[j] > ss[j] && es[j] <= n) {
                    old = es[j] - ss[j];
                    stay = old - movers_between(moved, ss[j], es[j]);
                }
            }
            cnt[j] = stay + ins[j];
            big = big || ins[j] > in_cap || cnt[j] > kMaxCellCount || old > kMaxCellCount;
        }
    }
    big = big || rep_in;
    if (__any(big) && (t & 63u) == 0) atomicAdd(&counters[kCntFullSort], 1u);
    const uint32_t sum = cnt[0] + cnt[1] + cnt[2] + cnt[3];
    const uint32_t incl = block_incl_scan(sum, s_wave);
    if (t == kBlock - 1) s_total = incl;

    // the tile's first slot: the particles of every earlier tile, each the
    // tile's count of the previous order plus its movers in less its movers
    // out (no wait on other blocks)
    uint32_t pre = 0;
    if (tile_start) {
        if (t == 0) pre = tile_start[tile];  // summed by k_tile_scan (large grids)
    } else {
        for (uint32_t q = t; q < tile; q += kBlock) pre += (ok ? tile_total[q] : 0u) + tile_delta[q];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) pre += __shfl_xor(pre, off);
    __syncthreads();  // s_wave of the scan is read; s_total written
    if ((t & 63u) == 0) s_wave[t >> 6] = pre;
    __syncthreads();
    if (t == 0) {
        uint32_t e = 0;
        for (int q = 0; q < kBlock / 64; ++q) e += s_wave[q];
        s_excl = e;
        // the tile's own count two ways: on a disagreement the step takes the
        // full sort, and the tile's cells are written empty here (the place
        // kernel writes every occupied cell's range from the sorted keys)
        if ((ok ? tile_total[tile] : 0u) + tile_delta[tile] != s_total && !rep_in) {
            s_bad = 1u;
            atomicOr(repair, kRepairCounts);
            atomicAdd(&counters[kCntFullSort], 1u);
        }
        total_next[tile] = s_total;
        delta_next[tile] = 0;
    }
    __syncthreads();
    if (c0 < num_cells) {
        const bool bad = s_bad != 0;
        uint32_t s = s_excl + incl - sum;
        uint32_t st[kCellItems], en[kCellItems];
#pragma unroll
        for (int j = 0; j < kCellItems; ++j) {
            st[j] = cnt[j] && !bad ? s : kEmptyCell;
            s += cnt[j];
            en[j] = s;
        }
        if (c0 + kCellItems <= num_cells) {
            *reinterpret_cast<uint4*>(new_start + c0) = make_uint4(st[0], st[1], st[2], st[3]);
            *reinterpret_cast<uint4*>(new_end + c0) = make_uint4(en[0], en[1], en[2], en[3]);
        } else {
            for (uint32_t j = 0; c0 + j < num_cells; ++j) {
                new_start[c0 + j] = st[j];
                new_end[c0 + j] = en[j];
            }
        }
    }
}

// First slot of every tile (large grids): the exclusive prefix over tiles of
// each tile's count in the last order plus its movers in less its movers
// out, by one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_tile_scan(uint32_t num_tiles, const uint32_t* __restrict__ valid,
                                                    const uint32_t* __restrict__ err,
                                                    const uint32_t* __restrict__ tile_total,
                                                    const uint32_t* __restrict__ tile_delta,
                                                    uint32_t* __restrict__ tile_start) {
    __shared__ uint32_t s_wave[1024 / 64];
    __shared__ uint32_t s_carry;
    const bool ok = *valid != 0 && (*err & kErrInvalid) == 0;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < num_tiles; base += 1024) {
        const uint32_t q = base + threadIdx.x;
        const uint32_t v = q < num_tiles ? (ok ? tile_total[q] : 0u) + tile_delta[q] : 0u;
        const uint32_t incl = block_incl_scan(v, s_wave);
        const uint32_t carry = s_carry;
        if (q < num_tiles) tile_start[q] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = carry + incl;
        __syncthreads();
    }
}

inline uint32_t scan_tiles(uint32_t num_cells) {
    const uint32_t t = (uint32_t)(((uint64_t)num_cells + kScanTile - 1) / kScanTile);
    return t == 0 ? 1 : t;
}

inline size_t pad4(uint32_t cells) { return ((size_t)cells + 3) & ~(size_t)3; }

// Incoming-list entries per cell: kMaxIncoming, less when the lists of a
// grid that large for its particles would pass the budget.
inline uint32_t in_capacity(uint32_t n, uint32_t num_cells) {
    uint64_t budget = (uint64_t)n * kInListBudgetPerParticle;
    if (budget < kInListBudgetMin) budget = kInListBudgetMin;
    const uint64_t cap = budget / ((uint64_t)num_cells * 4);
    return cap < kMaxIncoming ? (uint32_t)cap : kMaxIncoming;
}

inline size_t tile_start_words(uint32_t num_cells) {
    return scan_tiles(num_cells) > kDirectTiles ? scan_tiles(num_cells) : 0;
}

}  // namespace

size_t resort_scratch_bytes(uint32_t n, uint32_t num_cells) {
    // mover counts x 2 (padded to four cells) | incoming lists | tile totals
    // x 2 | tile deltas x 2 | tile starts (large grids) | valid, err,
    // valid_step, repair, repairs (32 bytes) | mover bits
    return 2 * pad4(num_cells) * 4 + (size_t)num_cells * in_capacity(n, num_cells) * 4 +
           4 * (size_t)scan_tiles(num_cells) * 4 + tile_start_words(num_cells) * 4 + 32 + ((size_t)n / 64 + 2) * 8;
}

hipError_t resort_plan_init(ResortPlan& plan, uint32_t n, uint32_t num_cells, void* scratch, hipStream_t s) {
    plan.n = n;
    plan.num_cells = num_cells;
    plan.num_tiles = scan_tiles(num_cells);
    plan.in_cap = in_capacity(n, num_cells);
    char* p = static_cast<char*>(scratch);
    plan.in_cnt[0] = reinterpret_cast<uint32_t*>(p);
    plan.in_cnt[1] = plan.in_cnt[0] + pad4(num_cells);
    hipError_t e = hipMemsetAsync(p, 0, 2 * pad4(num_cells) * 4, s);
    p += 2 * pad4(num_cells) * 4;
    plan.inlist = reinterpret_cast<uint32_t*>(p);
    p += (size_t)num_cells * plan.in_cap * 4;
    char* q = p;
    for (int k = 0; k < 2; ++k) {
        plan.tile_total[k] = reinterpret_cast<uint32_t*>(p) + (size_t)k * plan.num_tiles;
        plan.tile_delta[k] = reinterpret_cast<uint32_t*>(p) + (size_t)(2 + k) * plan.num_tiles;
    }
    p += 4 * (size_t)plan.num_tiles * 4;
    plan.tile_start = tile_start_words(num_cells) ? reinterpret_cast<uint32_t*>(p) : nullptr;
    p += tile_start_words(num_cells) * 4;
    plan.valid = reinterpret_cast<uint32_t*>(p);
    plan.err = reinterpret_cast<uint32_t*>(p + 4);
    plan.valid_step = reinterpret_cast<uint32_t*>(p + 8);
    plan.repair = reinterpret_cast<uint32_t*>(p + 12);
    plan.repairs = reinterpret_cast<uint32_t*>(p + 16);
    p += 32;
    plan.moved = reinterpret_cast<uint64_t*>(p);
    p += ((size_t)n / 64 + 2) * 8;
    if (e == hipSuccess) e = hipMemsetAsync(q, 0, (size_t)(p - q), s);
    return e;
}

hipError_t resort_invalidate(const ResortPlan& plan, hipStream_t s) {
    return launch_fill_u32(plan.valid, 0u, 1, s);
}

hipError_t resort_order(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, int parity, hipStream_t s) {
    if (rp.n != c.n || rp.num_cells != c.num_cells) return hipErrorInvalidValue;
    if (rp.tile_start)
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, rp.num_tiles, rp.valid, rp.err, rp.tile_total[parity],
                           rp.tile_delta[parity], rp.tile_start);
    hipLaunchKernelGGL(k_cell_scan, dim3(rp.num_tiles), dim3(kBlock), 0, s, c.num_cells, rp.valid, b.cell_start,
                       b.cell_end, rp.in_cnt[parity], rp.in_cnt[parity ^ 1], rp.moved, b.cell_start_alt,
                       b.cell_end_alt, b.counters, rp.tile_total[parity], rp.tile_delta[parity],
                       rp.tile_total[parity ^ 1], rp.tile_delta[parity ^ 1], rp.valid_step, rp.in_cap, rp.tile_start,
                       rp.err, rp.repair, c.n);
    return hipGetLastError();
}

}  // namespace sph
