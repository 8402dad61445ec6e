// sph_sort.hip - stable LSD radix sort of (key, value) uint32 pairs for gfx950.
//
// Replaces RadixSum / RadixPrefixSum / RadixAddOffsetsAndShuffle
// (source/CUDA/radixsort_kernel.cu:70-432, driver :445-472).  Contract kept:
// a stable sort of 8-byte pairs by key, 8 key bits per pass, so that pairs
// (hash, index) given in index order come out in lexicographic
// (hash, index) order.  The pass count follows the key width the caller
// asks for (ceil(log2(numCells)) bits) instead of the reference's fixed 16
// or 32 bits.
//
// Structure (one read of the keys for all digit histograms, then one kernel
// per pass):
//   k_hist        per-block digit histograms of every pass in one read,
//   k_hist_scan   sum over blocks + exclusive scan -> global digit offsets,
//   k_pass        per tile: stable local ranking with 64-lane ballots (peer
//                 mask per digit, rank = popcount of lower peers), tile digit
//                 counts published and prefixed across tiles with a
//                 decoupled look-back, local reorder in LDS, coalesced runs
//                 written to each digit's global range.
// Tiles are numbered by an atomic ticket so a tile only ever waits for tiles
// that started before it; every wait is bounded by a spin limit that records
// an error instead of hanging.  Ticket counter and look-back status words
// carry the launch number, so nothing needs clearing between launches and
// replays of a captured graph stay correct.
#include "sph_internal.hpp"

namespace sph {
namespace {

constexpr int kSortBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kSortBlock / kWave;
constexpr int kItems = 8;
constexpr int kTile = kSortBlock * kItems;  // keys per tile
constexpr int kRadix = 256;
constexpr int kMaxPasses = 4;
constexpr int kHistBlocks = 64;
constexpr uint32_t kSpinLimit = 1u << 22;

constexpr uint64_t kFlagAgg = 1ull << 30;
constexpr uint64_t kFlagInc = 2ull << 30;
constexpr uint64_t kValMask = (1ull << 30) - 1;

static_assert(kSortBlock == kRadix, "one thread per digit in the tile bookkeeping");

__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kSortBlock) void k_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t bits, uint32_t* __restrict__ block_hist) {
    __shared__ uint32_t hist[kMaxPasses][kRadix];
    const int t = threadIdx.x;
    for (int p = 0; p < kMaxPasses; ++p) hist[p][t] = 0;
    __syncthreads();
    const uint32_t passes = (bits + 7) / 8;
    for (uint32_t i = blockIdx.x * kSortBlock + t; i < n; i += gridDim.x * kSortBlock) {
        const uint32_t k = keys[i];
        for (uint32_t p = 0; p < passes; ++p) atomicAdd(&hist[p][(k >> (8 * p)) & 0xFFu], 1u);
    }
    __syncthreads();
    for (uint32_t p = 0; p < passes; ++p)
        block_hist[((size_t)blockIdx.x * kMaxPasses + p) * kRadix + t] = hist[p][t];
}

// One block per pass: digit totals over the histogram blocks, exclusive scan.
__global__ __launch_bounds__(kRadix) void k_hist_scan(const uint32_t* __restrict__ block_hist,
                                                     uint32_t nblocks, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t s[kRadix];
    const int t = threadIdx.x;
    const int p = blockIdx.x;
    uint32_t part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // independent loads in flight: eight running sums
    for (uint32_t b = 0; b < nblocks; b += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (b + u < nblocks) part[u] += block_hist[((size_t)(b + u) * kMaxPasses + p) * kRadix + t];
    }
    uint32_t tot = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) tot += part[u];
    s[t] = tot;
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        const uint32_t add = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    offsets[p * kRadix + t] = s[t] - tot;
}

__global__ __launch_bounds__(kSortBlock) void k_pass(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
    uint32_t pass_bits, const uint32_t* __restrict__ digit_off, uint64_t* __restrict__ status,
    uint64_t* __restrict__ ticket, uint32_t num_tiles, uint32_t* __restrict__ err) {
    __shared__ uint32_t wcount[kWaves][kRadix];
    __shared__ uint32_t dig_off[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t skeys[kTile];
    __shared__ uint32_t svals[kTile];
    __shared__ uint64_t s_ticket;

    const int t = threadIdx.x;
    const int wave = t / kWave;
    const int lane = t % kWave;
    if (t == 0) s_ticket = atomicAdd((unsigned long long*)ticket, 1ull);
    for (int w = 0; w < kWaves; ++w) wcount[w][t] = 0;
    __syncthreads();
    const uint32_t tile = (uint32_t)(s_ticket % num_tiles);
    const uint64_t epoch = (s_ticket / num_tiles) & 0xFFFFFFFFull;
    const uint32_t mask = (1u << pass_bits) - 1u;

    const uint64_t tile_base = (uint64_t)tile * kTile;
    const uint64_t wbase = tile_base + (uint64_t)wave * (kWave * kItems);
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint32_t k[kItems], v[kItems], rank[kItems];

#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t e = wbase + (uint64_t)it * kWave + lane;
        const bool valid = e < n;
        const uint32_t key = valid ? keys_in[e] : 0u;
        const uint32_t val = valid ? vals_in[e] : 0u;
        const uint32_t d = (key >> shift) & mask;
        uint64_t peers = __ballot(valid);
        for (uint32_t bit = 0; bit < pass_bits; ++bit) {
            const bool set = (d >> bit) & 1u;
            const uint64_t bal = __ballot(set);
            peers &= set ? bal : ~bal;
        }
        const uint32_t r = (uint32_t)__popcll(peers & lt_mask);
        uint32_t pre = 0;
        if (valid) pre = wcount[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && r == 0) wcount[wave][d] = pre + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        k[it] = key;
        v[it] = val;
        rank[it] = valid ? pre + r : 0xFFFFFFFFu;
    }
    __syncthreads();

    // Per digit (thread t == digit): exclusive prefix over waves, tile total.
    uint32_t cnt = 0;
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = wcount[w][t];
        wcount[w][t] = cnt;
        cnt += c;
    }
    // Publish the tile's count, then look back over earlier tiles.
    uint64_t* my = status + (size_t)tile * kRadix + t;
    const uint64_t tag = epoch << 32;
    if (tile == 0) {
        st_status(my, tag | kFlagInc | cnt);
        gbase[t] = digit_off[t];
    } else {
        st_status(my, tag | kFlagAgg | cnt);
        uint32_t prefix = 0;
        int64_t kt = (int64_t)tile - 1;
        uint32_t spins = 0;
        while (kt >= 0) {
            const uint64_t s = ld_status(status + (size_t)kt * kRadix + t);
            if ((s >> 32) != epoch || (s & (kFlagAgg | kFlagInc)) == 0) {
                if (++spins > kSpinLimit) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += (uint32_t)(s & kValMask);
            if (s & kFlagInc) break;
            --kt;
        }
        st_status(my, tag | kFlagInc | (uint64_t)(prefix + cnt));
        gbase[t] = digit_off[t] + prefix;
    }
    dig_off[t] = cnt;
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        const uint32_t add = t >= off ? dig_off[t - off] : 0u;
        __syncthreads();
        dig_off[t] += add;
        __syncthreads();
    }
    const uint32_t incl = dig_off[t];
    __syncthreads();
    dig_off[t] = incl - cnt;
    __syncthreads();

#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        if (rank[it] == 0xFFFFFFFFu) continue;
        const uint32_t d = (k[it] >> shift) & mask;
        const uint32_t lp = dig_off[d] + wcount[wave][d] + rank[it];
        skeys[lp] = k[it];
        svals[lp] = v[it];
    }
    __syncthreads();

    const uint64_t rem = (uint64_t)n - tile_base;
    const uint32_t count = rem < (uint64_t)kTile ? (uint32_t)rem : (uint32_t)kTile;
    for (uint32_t i = t; i < count; i += kSortBlock) {
        const uint32_t key = skeys[i];
        const uint32_t d = (key >> shift) & mask;
        const uint32_t g = gbase[d] + (i - dig_off[d]);
        keys_out[g] = key;
        vals_out[g] = svals[i];
    }
}

}  // namespace

uint32_t radix_num_tiles(uint32_t n) {
    const uint32_t tiles = (n + kTile - 1) / kTile;
    return tiles == 0 ? 1 : tiles;
}

size_t radix_scratch_bytes(uint32_t n) {
    // block histograms | digit offsets | status words | ticket | error word
    return (size_t)kHistBlocks * kMaxPasses * kRadix * 4 + (size_t)kMaxPasses * kRadix * 4 +
           (size_t)radix_num_tiles(n) * kRadix * 8 + 64;
}

hipError_t radix_plan_init(RadixSortPlan& plan, uint32_t n, void* scratch, hipStream_t stream) {
    plan.n = n;
    plan.num_tiles = radix_num_tiles(n);
    char* b = static_cast<char*>(scratch);
    plan.block_hist = reinterpret_cast<uint32_t*>(b);
    b += (size_t)kHistBlocks * kMaxPasses * kRadix * 4;
    plan.digit_off = reinterpret_cast<uint32_t*>(b);
    b += (size_t)kMaxPasses * kRadix * 4;
    plan.status = reinterpret_cast<uint64_t*>(b);
    b += (size_t)plan.num_tiles * kRadix * 8;
    plan.ticket = reinterpret_cast<uint64_t*>(b);
    plan.err = reinterpret_cast<uint32_t*>(b + 8);
    // status words start at epoch 0 / no flag; the ticket and error word at 0
    return hipMemsetAsync(plan.status, 0, (size_t)plan.num_tiles * kRadix * 8 + 64, stream);
}

hipError_t radix_sort_pairs(const RadixSortPlan& plan, uint32_t bits, uint32_t* keys0,
                            uint32_t* vals0, uint32_t* keys1, uint32_t* vals1,
                            uint32_t** keys_out, uint32_t** vals_out, hipStream_t stream) {
    uint32_t* kin = keys0;
    uint32_t* vin = vals0;
    uint32_t* kout = keys1;
    uint32_t* vout = vals1;
    const uint32_t n = plan.n;
    const uint32_t passes = (bits + 7) / 8;
    if (passes > (uint32_t)kMaxPasses) return hipErrorInvalidValue;
    if (passes > 0) {
        hipLaunchKernelGGL(k_hist, dim3(kHistBlocks), dim3(kSortBlock), 0, stream, kin, n, bits,
                           plan.block_hist);
        hipLaunchKernelGGL(k_hist_scan, dim3(passes), dim3(kRadix), 0, stream, plan.block_hist,
                           (uint32_t)kHistBlocks, plan.digit_off);
    }
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = 8 * p;
        const uint32_t pb = (bits - shift) < 8 ? (bits - shift) : 8;
        hipLaunchKernelGGL(k_pass, dim3(plan.num_tiles), dim3(kSortBlock), 0, stream, kin, vin, kout,
                           vout, n, shift, pb, plan.digit_off + p * kRadix, plan.status, plan.ticket,
                           plan.num_tiles, plan.err);
        uint32_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    *keys_out = kin;
    *vals_out = vin;
    return hipGetLastError();
}

}  // namespace sph
