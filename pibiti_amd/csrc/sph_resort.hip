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
//   k_integrate_moves  (sph_kernels.hip) new key per slot, mover bits, movers
//                      counted into their new cells;
//   k_cell_scan        per cell: stayers = old count - movers out (from the
//                      old range and the mover bits), new count = stayers +
//                      movers in; exclusive scan over all cells (decoupled
//                      look-back) -> new cell ranges; flags the step for the
//                      full sort when a cell takes more than kMaxIncoming
//                      movers or holds more than kMaxCellCount particles;
//   k_place_movers     each mover's original index into its new cell's
//                      incoming list (the first entries of the cell's new
//                      range in a scratch array);
//   k_resort_scatter   slot -> new slot: cell start + stayers of the cell
//                      with a smaller index (a stayer: its rank among the old
//                      range's stayers; a mover: a binary search of the old
//                      range) + incoming entries with a smaller index; the
//                      state, index, scene and key scattered there.
//                      On a full-sort step it writes the radix sort's input
//                      instead;
//   radix sort + k_resort_gather  only on a full-sort step (gated on the
//                      device, so the launch sequence never changes and
//                      graphs replay it).
// While the order is invalid (first step, after the learned surrogate or
// the tuner rearranged the slots) every slot counts as a mover with no old
// ranges, which is a counting sort by cell with the same result.
#include "sph_device.hpp"

namespace sph {
namespace {

constexpr int kBlock = 256;
constexpr int kCellItems = 16;
constexpr uint32_t kCellTile = kBlock * kCellItems;  // cells per scan tile
constexpr uint32_t kSpinLimit = 1u << 22;
constexpr uint64_t kFlagAgg = 1ull << 30;
constexpr uint64_t kFlagInc = 2ull << 30;
constexpr uint64_t kValMask = (1ull << 30) - 1;

__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool is_mover(const uint64_t* __restrict__ moved, uint32_t k) {
    return ((moved[k >> 6] >> (k & 63u)) & 1ull) != 0;
}

// Movers among slots [a, b).
__device__ __forceinline__ uint32_t movers_between(const uint64_t* __restrict__ moved, uint32_t a, uint32_t b) {
    if (a >= b) return 0;
    const uint32_t wa = a >> 6, wb = b >> 6, la = a & 63u, lb = b & 63u;
    if (wa == wb) return (uint32_t)__popcll((moved[wa] >> la) & ((1ull << (lb - la)) - 1ull));
    uint32_t c = (uint32_t)__popcll(moved[wa] >> la);
    for (uint32_t w = wa + 1; w < wb; ++w) c += (uint32_t)__popcll(moved[w]);
    if (lb) c += (uint32_t)__popcll(moved[wb] & ((1ull << lb) - 1ull));
    return c;
}

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

// LDS index of tile item i, one word of padding per 16 so that the threads'
// runs of 16 consecutive items start in different banks.
__host__ __device__ constexpr uint32_t pad(uint32_t i) { return i + (i >> 4); }

__global__ __launch_bounds__(kBlock) void k_cell_scan(
    uint32_t num_cells, const uint32_t* __restrict__ valid, const uint32_t* __restrict__ old_start,
    const uint32_t* __restrict__ old_end, const uint32_t* __restrict__ in_cnt, const uint64_t* __restrict__ moved,
    uint32_t* __restrict__ new_start, uint32_t* __restrict__ new_end, uint32_t* __restrict__ counters,
    uint64_t* __restrict__ status, uint64_t* __restrict__ ticket, uint32_t num_tiles, uint32_t* __restrict__ err) {
    __shared__ uint32_t s_cnt[pad(kCellTile - 1) + 1];
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_total, s_excl;
    __shared__ uint64_t s_ticket;
    const uint32_t t = threadIdx.x;
    if (t == 0) s_ticket = atomicAdd((unsigned long long*)ticket, 1ull);
    __syncthreads();
    const uint32_t tile = (uint32_t)(s_ticket % num_tiles);
    const uint64_t epoch = (s_ticket / num_tiles) & 0xFFFFFFFFull;
    const bool ok = *valid != 0;
    const uint32_t base = tile * kCellTile;

    // counts, one cell per thread per row (coalesced)
    uint32_t cnt[kCellItems];
    bool big = false;
#pragma unroll
    for (int i = 0; i < kCellItems; ++i) {
        const uint32_t c = base + (uint32_t)i * kBlock + t;
        uint32_t v = 0;
        if (c < num_cells) {
            const uint32_t in = in_cnt[c];
            uint32_t stay = 0, old = 0;
            if (ok) {
                const uint32_t s = old_start[c];
                if (s != kEmptyCell) {
                    const uint32_t e = old_end[c];
                    old = e - s;
                    stay = old - movers_between(moved, s, e);
                }
            }
            v = stay + in;
            big = big || in > kMaxIncoming || v > kMaxCellCount || old > kMaxCellCount;
        }
        cnt[i] = v;
        s_cnt[pad((uint32_t)i * kBlock + t)] = v;
    }
    if (__any(big) && (t & 63u) == 0) atomicAdd(&counters[kCntFullSort], 1u);
    __syncthreads();

    // thread t scans items 16 t .. 16 t + 15 (cells in order)
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kCellItems; ++j) sum += s_cnt[pad(t * kCellItems + j)];
    const uint32_t incl = block_incl_scan(sum, s_wave);
    if (t == kBlock - 1) s_total = incl;
    uint32_t run = incl - sum;
#pragma unroll
    for (int j = 0; j < kCellItems; ++j) {
        const uint32_t q = pad(t * kCellItems + j);
        const uint32_t v = s_cnt[q];
        s_cnt[q] = run;
        run += v;
    }
    __syncthreads();

    // tile prefix: publish the aggregate, look back to an inclusive prefix
    if (t == 0) {
        const uint64_t ep = epoch << 32;
        const uint32_t total = s_total;
        uint32_t excl = 0;
        if (tile == 0) {
            st_status(&status[0], ep | kFlagInc | total);
        } else {
            st_status(&status[tile], ep | kFlagAgg | total);
            uint32_t spins = 0;
            int32_t j = (int32_t)tile - 1;
            while (j >= 0) {
                const uint64_t w = ld_status(&status[j]);
                if ((w >> 32) != epoch || (w & (kFlagAgg | kFlagInc)) == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(err, 1u);
                        break;
                    }
                    continue;
                }
                excl += (uint32_t)(w & kValMask);
                if (w & kFlagInc) break;
                --j;
            }
            st_status(&status[tile], ep | kFlagInc | ((excl + total) & kValMask));
        }
        s_excl = excl;
    }
    __syncthreads();
    const uint32_t excl = s_excl;
#pragma unroll
    for (int i = 0; i < kCellItems; ++i) {
        const uint32_t c = base + (uint32_t)i * kBlock + t;
        if (c < num_cells) {
            const uint32_t s = excl + s_cnt[pad((uint32_t)i * kBlock + t)];
            new_start[c] = cnt[i] ? s : kEmptyCell;
            new_end[c] = s + cnt[i];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_place_movers(uint32_t n, const uint32_t* __restrict__ nkey,
                                                        const uint64_t* __restrict__ moved,
                                                        uint32_t* __restrict__ in_cnt,
                                                        const uint32_t* __restrict__ new_start,
                                                        const uint32_t* __restrict__ pid,
                                                        uint32_t* __restrict__ inbuf) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !is_mover(moved, k)) return;
    const uint32_t c = nkey[k];
    // counting down leaves every count at zero for the next step
    const uint32_t t = atomicSub(&in_cnt[c], 1u) - 1u;
    const uint32_t q = new_start[c] + t;
    if (q < n) inbuf[q] = pid[k];
}

template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_resort_scatter(
    sph_params one, const sph_params* __restrict__ tab, uint32_t n, const uint32_t* __restrict__ nkey,
    const uint64_t* __restrict__ moved, const uint32_t* __restrict__ valid, const uint32_t* __restrict__ old_start,
    const uint32_t* __restrict__ old_end, const uint32_t* __restrict__ new_start,
    const uint32_t* __restrict__ new_end, const uint32_t* __restrict__ inbuf, const float4* __restrict__ pos_in,
    const float4* __restrict__ vel_in, const uint32_t* __restrict__ pid_in, const uint32_t* __restrict__ sid_in,
    float4* __restrict__ pos_out, float4* __restrict__ vel_out, uint32_t* __restrict__ pid_out,
    uint32_t* __restrict__ sid_out, uint32_t* __restrict__ skey_out, uint32_t* __restrict__ key_pid,
    uint32_t* __restrict__ val_pid, uint32_t* __restrict__ counters, uint32_t* __restrict__ err) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t c = nkey[k], id = pid_in[k];
    if (counters[kCntFullSort]) {
        // the radix sort's input: keys in original-index order
        key_pid[id] = c;
        val_pid[id] = k;
        return;
    }
    const bool mv = is_mover(moved, k);
    const uint32_t base = new_start[c], end = new_end[c];
    bool bad = base >= end || end > n;
    uint32_t os = kEmptyCell, oe = 0, stay = 0;
    if (*valid) {
        os = old_start[c];
        if (os != kEmptyCell) {
            oe = old_end[c];
            if (oe > n || oe <= os)
                bad = true;
            else
                stay = (oe - os) - movers_between(moved, os, oe);
        }
    }
    // a stayer lies in its cell's old range
    if (!mv && (os == kEmptyCell || k < os || k >= oe)) bad = true;
    if (bad || end - base < stay) {
        atomicOr(err, 1u);
        return;
    }
    const uint32_t nin = (end - base) - stay;
    uint32_t local = 0;
    if (!mv) {
        local = (k - os) - movers_between(moved, os, k);
    } else if (os != kEmptyCell) {
        uint32_t lo = os, hi = oe;
        while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (pid_in[mid] < id)
                lo = mid + 1;
            else
                hi = mid;
        }
        local = (lo - os) - movers_between(moved, os, lo);
    }
    uint32_t inr = 0;
    for (uint32_t q = 0; q < nin; ++q) inr += inbuf[base + q] < id ? 1u : 0u;
    const uint32_t r = base + local + inr;
    if (r >= end) {
        atomicOr(err, 1u);
        return;
    }
    pos_out[r] = pos_in[k];
    vel_out[r] = vel_in[k];
    pid_out[r] = id;
    skey_out[r] = c;
    uint32_t cap = one.max_par_in_cell;
    if constexpr (kBatch) {
        const uint32_t sc = sid_in[k];
        sid_out[r] = sc;
        cap = tab[sc].max_par_in_cell;
    }
    if (r + 1 == end && end - base > cap) atomicAdd(&counters[kCntCapCells], 1u);
}

// A full-sort step's reorder (the cell ranges come from k_cell_scan); marks
// the new order valid on every step.
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_resort_gather(
    sph_params one, const sph_params* __restrict__ tab, uint32_t n, const uint32_t* __restrict__ skey,
    const uint32_t* __restrict__ old_slot, const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,
    const uint32_t* __restrict__ pid_in, const uint32_t* __restrict__ sid_in, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ pid_out, uint32_t* __restrict__ sid_out,
    uint32_t* __restrict__ skey_out, uint32_t* __restrict__ counters, uint32_t* __restrict__ valid) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *valid = 1u;
    if (k >= n || counters[kCntFullSort] == 0) return;
    const uint32_t s = old_slot[k];
    pos_out[k] = pos_in[s];
    vel_out[k] = vel_in[s];
    pid_out[k] = pid_in[s];
    const uint32_t h = skey[k];
    skey_out[k] = h;
    uint32_t cap = one.max_par_in_cell;
    if constexpr (kBatch) {
        const uint32_t sc = sid_in[s];
        sid_out[k] = sc;
        cap = tab[sc].max_par_in_cell;
    }
    if ((k + 1 == n || skey[k + 1] != h) && k >= cap && skey[k - cap] == h)
        atomicAdd(&counters[kCntCapCells], 1u);
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

inline uint32_t scan_tiles(uint32_t num_cells) {
    const uint32_t t = (uint32_t)(((uint64_t)num_cells + kCellTile - 1) / kCellTile);
    return t == 0 ? 1 : t;
}

template <bool kBatch>
hipError_t apply_impl(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, const RadixSortPlan& sort,
                      hipStream_t s) {
    const dim3 grid(blocks(c.n)), block(kBlock);
    hipLaunchKernelGGL((k_resort_scatter<kBatch>), grid, block, 0, s, c.one, c.tab, c.n, b.key_tmp, rp.moved,
                       rp.valid, b.cell_start, b.cell_end, b.cell_start_alt, b.cell_end_alt, b.val_tmp,
                       b.pos_cur, b.vel_cur, b.pid_cur, b.sid_cur, b.pos_alt, b.vel_alt, b.pid_alt, b.sid_alt,
                       b.skey_alt, b.key_pid, b.val_pid, b.counters, rp.err);
    uint32_t *skey = nullptr, *oslot = nullptr;
    hipError_t e = radix_sort_pairs(sort, c.key_bits, b.key_pid, b.val_pid, b.key_tmp, b.val_tmp, &skey, &oslot, s,
                                    b.counters + kCntFullSort);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_resort_gather<kBatch>), grid, block, 0, s, c.one, c.tab, c.n, skey, oslot, b.pos_cur,
                       b.vel_cur, b.pid_cur, b.sid_cur, b.pos_alt, b.vel_alt, b.pid_alt, b.sid_alt, b.skey_alt,
                       b.counters, rp.valid);
    return hipGetLastError();
}

}  // namespace

size_t resort_scratch_bytes(uint32_t n, uint32_t num_cells) {
    // valid | err | ticket | status words | mover bits | mover counts
    return 16 + (size_t)scan_tiles(num_cells) * 8 + ((size_t)n / 64 + 2) * 8 + (size_t)num_cells * 4;
}

hipError_t resort_plan_init(ResortPlan& plan, uint32_t n, uint32_t num_cells, void* scratch, hipStream_t s) {
    plan.n = n;
    plan.num_cells = num_cells;
    plan.num_tiles = scan_tiles(num_cells);
    char* p = static_cast<char*>(scratch);
    plan.valid = reinterpret_cast<uint32_t*>(p);
    plan.err = reinterpret_cast<uint32_t*>(p + 4);
    plan.ticket = reinterpret_cast<uint64_t*>(p + 8);
    p += 16;
    plan.status = reinterpret_cast<uint64_t*>(p);
    p += (size_t)plan.num_tiles * 8;
    plan.moved = reinterpret_cast<uint64_t*>(p);
    p += ((size_t)n / 64 + 2) * 8;
    plan.in_cnt = reinterpret_cast<uint32_t*>(p);
    return hipMemsetAsync(scratch, 0, resort_scratch_bytes(n, num_cells), s);
}

hipError_t resort_invalidate(const ResortPlan& plan, hipStream_t s) {
    return launch_fill_u32(plan.valid, 0u, 1, s);
}

hipError_t resort_order(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, hipStream_t s) {
    if (rp.n != c.n || rp.num_cells != c.num_cells) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cell_scan, dim3(rp.num_tiles), dim3(kBlock), 0, s, c.num_cells, rp.valid, b.cell_start,
                       b.cell_end, rp.in_cnt, rp.moved, b.cell_start_alt, b.cell_end_alt, b.counters, rp.status,
                       rp.ticket, rp.num_tiles, rp.err);
    hipLaunchKernelGGL(k_place_movers, dim3(blocks(c.n)), dim3(kBlock), 0, s, c.n, b.key_tmp, rp.moved, rp.in_cnt,
                       b.cell_start_alt, b.pid_cur, b.val_tmp);
    return hipGetLastError();
}

hipError_t resort_apply(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, const RadixSortPlan& sort,
                        hipStream_t s) {
    return c.tab ? apply_impl<true>(c, b, rp, sort, s) : apply_impl<false>(c, b, rp, sort, s);
}

}  // namespace sph
