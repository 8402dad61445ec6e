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
// Per pass, three kernels:
//   upsweep   - per-tile digit histograms (wave-private LDS counters),
//   scan      - exclusive scan of the digit-major [digit][tile] table,
//   downsweep - stable local ranking with 64-lane ballots (peer mask per
//               digit, rank = popcount of lower peers), local reorder in LDS,
//               then coalesced runs written to the digit's global range.
// No spin waits: every kernel finishes independently of the others.
#include "sph_internal.hpp"

namespace sph {
namespace {

constexpr int kSortBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kSortBlock / kWave;
constexpr int kItems = 16;
constexpr int kTile = kSortBlock * kItems;  // keys per tile (block)
constexpr int kRadix = 256;
constexpr int kScanBlock = 1024;

static_assert(kSortBlock == kRadix, "one thread per digit in the downsweep bookkeeping");

__global__ __launch_bounds__(kSortBlock) void k_upsweep(const uint32_t* __restrict__ keys,
                                                       uint32_t n, uint32_t shift, uint32_t mask,
                                                       uint32_t* __restrict__ hist_out,
                                                       uint32_t num_tiles) {
    __shared__ uint32_t hist[kWaves][kRadix];
    const int t = threadIdx.x;
    const int wave = t / kWave;
    for (int w = 0; w < kWaves; ++w) hist[w][t] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll 4
    for (int it = 0; it < kItems; ++it) {
        const uint64_t e = base + (uint64_t)it * kSortBlock + t;
        if (e < n) atomicAdd(&hist[wave][(keys[e] >> shift) & mask], 1u);
    }
    __syncthreads();
    uint32_t s = 0;
    for (int w = 0; w < kWaves; ++w) s += hist[w][t];
    hist_out[(size_t)t * num_tiles + blockIdx.x] = s;
}

// Exclusive scan of `total` counters in place, one block.
__global__ __launch_bounds__(kScanBlock) void k_scan(uint32_t* __restrict__ data, uint32_t total) {
    __shared__ uint32_t sums[kScanBlock];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (total + kScanBlock - 1) / kScanBlock;
    const uint32_t b = t * chunk;
    const uint32_t e = min(total, b + chunk);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += data[i];
    sums[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums.
    for (uint32_t off = 1; off < kScanBlock; off <<= 1) {
        uint32_t v = t >= off ? sums[t - off] : 0u;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    uint32_t run = t == 0 ? 0u : sums[t - 1];
    for (uint32_t i = b; i < e; ++i) {
        uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kSortBlock) void k_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
    uint32_t mask, uint32_t pass_bits, const uint32_t* __restrict__ tile_off, uint32_t num_tiles) {
    __shared__ uint32_t wcount[kWaves][kRadix];
    __shared__ uint32_t dig_off[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t skeys[kTile];
    __shared__ uint32_t svals[kTile];

    const int t = threadIdx.x;
    const int wave = t / kWave;
    const int lane = t % kWave;
    for (int w = 0; w < kWaves; ++w) wcount[w][t] = 0;
    __syncthreads();

    const uint64_t tile_base = (uint64_t)blockIdx.x * kTile;
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
    uint32_t run = 0;
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = wcount[w][t];
        wcount[w][t] = run;
        run += c;
    }
    dig_off[t] = run;
    gbase[t] = tile_off[(size_t)t * num_tiles + blockIdx.x];
    __syncthreads();
    // Exclusive scan of the digit totals (Hillis-Steele, inclusive then shift).
    for (int off = 1; off < kRadix; off <<= 1) {
        const uint32_t add = t >= off ? dig_off[t - off] : 0u;
        __syncthreads();
        dig_off[t] += add;
        __syncthreads();
    }
    const uint32_t incl = dig_off[t];
    __syncthreads();
    dig_off[t] = incl - run;
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

size_t radix_scratch_bytes(uint32_t n) {
    const uint32_t tiles = (n + kTile - 1) / kTile;
    return (size_t)kRadix * (tiles == 0 ? 1 : tiles) * sizeof(uint32_t);
}

hipError_t radix_sort_pairs(const RadixSortPlan& plan, uint32_t bits, uint32_t* keys0,
                            uint32_t* vals0, uint32_t* keys1, uint32_t* vals1,
                            uint32_t** keys_out, uint32_t** vals_out, hipStream_t stream) {
    uint32_t* kin = keys0;
    uint32_t* vin = vals0;
    uint32_t* kout = keys1;
    uint32_t* vout = vals1;
    const uint32_t n = plan.n;
    const uint32_t tiles = plan.num_tiles;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        const uint32_t pb = (bits - shift) < 8 ? (bits - shift) : 8;
        const uint32_t mask = (1u << pb) - 1u;
        hipLaunchKernelGGL(k_upsweep, dim3(tiles), dim3(kSortBlock), 0, stream, kin, n, shift, mask,
                           plan.block_hist, tiles);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanBlock), 0, stream, plan.block_hist,
                           (uint32_t)(kRadix * tiles));
        hipLaunchKernelGGL(k_downsweep, dim3(tiles), dim3(kSortBlock), 0, stream, kin, vin, kout,
                           vout, n, shift, mask, pb, plan.block_hist, tiles);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
        uint32_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    *keys_out = kin;
    *vals_out = vin;
    return hipSuccess;
}

}  // namespace sph
