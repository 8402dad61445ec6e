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
//   k_hist        per-block digit histograms of every pass in one read
//                 (runs of equal digits in a wave counted by one atomic);
//                 the last block to finish sums them over the blocks and
//                 scans -> global digit offsets (scan_pass),
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
#include <algorithm>
#include <mutex>

#include "sph_internal.hpp"

namespace sph {
namespace {

constexpr int kSortBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kSortBlock / kWave;
constexpr int kItems = 16;
constexpr int kTile = kSortBlock * kItems;  // keys per tile
constexpr int kRadix = 256;
constexpr int kMaxPasses = 4;
constexpr int kHistBlocks = 64;
constexpr int kHistThreads = 1024;
constexpr int kHistKeys = 16;                       // keys per thread per round
constexpr int kHistTile = kHistThreads * kHistKeys;  // keys per block per round
constexpr int kScanThreads = 1024;
constexpr int kScanQ = kScanThreads / kRadix;        // threads per digit in scan_pass
static_assert(kHistBlocks % kScanQ == 0 && kHistBlocks / kScanQ <= 16, "scan_pass: one round of loads");
static_assert(kHistThreads == kScanThreads, "the last histogram block runs the scan");
constexpr uint32_t kSpinLimit = 1u << 22;

constexpr uint64_t kFlagAgg = 1ull << 30;
constexpr uint64_t kFlagInc = 2ull << 30;
constexpr uint64_t kValMask = (1ull << 30) - 1;

static_assert(kSortBlock == kRadix, "one thread per digit in the tile bookkeeping");

// Inclusive scan of one value per thread over the block: a shuffle scan in
// each wave, then the totals of the waves before it.
template <int kThreads>
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* s_wave) {
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t u = __shfl_up(v, off);
        v += lane >= off ? u : 0u;
    }
    if (lane == kWave - 1) s_wave[w] = v;
    __syncthreads();
    uint32_t add = 0;
    for (int k = 0; k < w; ++k) add += s_wave[k];
    __syncthreads();
    return v + add;
}

__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Digit totals of pass p over the histogram blocks (kScanQ threads per
// digit, every load of a thread in flight at once), then an exclusive scan
// into offsets; the whole block of kScanThreads takes part.
__device__ __forceinline__ void scan_pass(const uint32_t* __restrict__ block_hist, uint32_t nblocks,
                                          uint32_t* __restrict__ offsets, uint32_t p,
                                          uint32_t (*part)[kRadix], uint32_t* s_wave) {
    const int t = threadIdx.x % kRadix;
    const int q = threadIdx.x / kRadix;
    constexpr int kRows = kHistBlocks / kScanQ;
    uint32_t v[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
        const uint32_t b = (uint32_t)(q + u * kScanQ);
        v[u] = block_hist[((size_t)min(b, nblocks - 1) * kMaxPasses + p) * kRadix + t];
    }
    uint32_t sum = 0;
#pragma unroll
    for (int u = 0; u < kRows; ++u) sum += (uint32_t)(q + u * kScanQ) < nblocks ? v[u] : 0u;
    part[q][t] = sum;
    __syncthreads();
    uint32_t tot = 0;
    if (q == 0) {
#pragma unroll
        for (int r = 0; r < kScanQ; ++r) tot += part[r][t];
    }
    // threads past the first kRadix add nothing and come after every digit
    const uint32_t incl = block_incl_scan<kScanThreads>(tot, s_wave);
    if (q == 0) offsets[p * kRadix + t] = incl - tot;
}

// Digit histograms of every pass in one read of the keys.  Each thread
// loads kHistKeys keys before counting (the loads overlap), and within a
// wave a run of consecutive lanes with the same digit is counted by one
// LDS atomic of its first lane: the keys arrive nearly sorted (the previous
// step's order), so runs are long and the atomics rarely collide.
// kStride: 1 for a key array, 2 for (key, value) pairs (KeyValuePair layout).
template <int kStride>
__global__ __launch_bounds__(kHistThreads) void k_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                      uint32_t bits, uint32_t* __restrict__ block_hist,
                                                      uint32_t* __restrict__ offsets, uint32_t* __restrict__ done,
                                                      const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ uint32_t hist[kMaxPasses][kRadix];
    const int t = threadIdx.x;
    const int lane = t % kWave;
    for (int c = t; c < kMaxPasses * kRadix; c += kHistThreads) hist[c / kRadix][c % kRadix] = 0;
    __syncthreads();
    const uint32_t passes = (bits + 7) / 8;
    const uint64_t after_me = lane == kWave - 1 ? 0ull : (~0ull << (lane + 1));
    for (uint64_t base = (uint64_t)blockIdx.x * kHistTile; base < n; base += (uint64_t)gridDim.x * kHistTile) {
        uint32_t k[kHistKeys];
#pragma unroll
        for (int it = 0; it < kHistKeys; ++it) {
            const uint64_t e = base + (uint64_t)it * kHistThreads + t;
            k[it] = keys[(e < n ? e : n - 1) * kStride];
        }
#pragma unroll
        for (int it = 0; it < kHistKeys; ++it) {
            // the valid lanes of a wave are a prefix of it (indices grow with the lane)
            const bool valid = base + (uint64_t)it * kHistThreads + t < n;
            const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
            for (uint32_t p = 0; p < passes; ++p) {
                // the pass's digit: its bits of the low `bits` (k_pass masks the same way)
                const uint32_t pb = bits - 8 * p < 8 ? bits - 8 * p : 8;
                const uint32_t d = (k[it] >> (8 * p)) & ((1u << pb) - 1u);
                const uint32_t dprev = __shfl_up(d, 1);
                const bool head = valid && (lane == 0 || dprev != d);
                const uint64_t later = __ballot(head) & after_me;
                const uint32_t end = later ? (uint32_t)(__ffsll((unsigned long long)later) - 1) : nvalid;
                if (head) atomicAdd(&hist[p][d], end - (uint32_t)lane);
            }
        }
    }
    __syncthreads();
    for (int c = t; c < (int)passes * kRadix; c += kHistThreads)
        block_hist[((size_t)blockIdx.x * kMaxPasses + c / kRadix) * kRadix + c % kRadix] = hist[c / kRadix][c % kRadix];

    // the last block to finish sums and scans the digits of every pass
    __shared__ uint32_t part[kScanQ][kRadix];
    __shared__ uint32_t s_wave[kScanThreads / kWave];
    __shared__ bool s_last;
    __threadfence();
    __syncthreads();
    if (t == 0) s_last = atomicAdd(done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    for (uint32_t p = 0; p < passes; ++p) {
        scan_pass(block_hist, gridDim.x, offsets, p, part, s_wave);
        __syncthreads();
    }
    if (t == 0) *done = 0u;  // ready for the next sort
}

// kPairsIn / kPairsOut: the input / output is one array of (key, value)
// pairs (keys_in / keys_out; vals_in / vals_out unused) instead of a key
// array and a value array.
template <bool kPairsIn, bool kPairsOut>
__global__ __launch_bounds__(kSortBlock) void k_pass(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
    uint32_t pass_bits, const uint32_t* __restrict__ digit_off, uint64_t* __restrict__ status,
    uint64_t* __restrict__ ticket, uint32_t num_tiles, uint32_t* __restrict__ err,
    const uint32_t* __restrict__ gate) {
    // a gated-off launch takes no ticket, so the epochs stay in step
    if (gate && *gate == 0) return;
    __shared__ uint32_t wcount[kWaves][kRadix];
    __shared__ uint32_t dig_off[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t skeys[kTile];
    __shared__ uint32_t svals[kTile];
    __shared__ uint32_t s_wave[kWaves];
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
        uint32_t key = 0u, val = 0u;
        if (valid) {
            if constexpr (kPairsIn) {
                const uint2 kv = reinterpret_cast<const uint2*>(keys_in)[e];
                key = kv.x;
                val = kv.y;
            } else {
                key = keys_in[e];
                val = vals_in[e];
            }
        }
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
    const uint32_t incl = block_incl_scan<kSortBlock>(cnt, s_wave);
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
        if constexpr (kPairsOut) {
            reinterpret_cast<uint2*>(keys_out)[g] = make_uint2(key, svals[i]);
        } else {
            keys_out[g] = key;
            vals_out[g] = svals[i];
        }
    }
}

}  // namespace

uint32_t radix_num_tiles(uint32_t n) {
    const uint32_t tiles = (n + kTile - 1) / kTile;
    return tiles == 0 ? 1 : tiles;
}

namespace {

// Scratch layout: block histograms | digit offsets | ticket (8 bytes), error
// word, finished histogram blocks (64 bytes in all) | status words (one per
// tile and digit).
constexpr size_t kHistBytes = (size_t)kHistBlocks * kMaxPasses * kRadix * 4 + (size_t)kMaxPasses * kRadix * 4;

constexpr size_t scratch_bytes_for_tiles(uint64_t tiles) { return kHistBytes + 64 + (size_t)tiles * kRadix * 8; }

void carve(RadixSortPlan& plan, uint32_t n, void* scratch) {
    plan.n = n;
    plan.num_tiles = radix_num_tiles(n);
    char* b = static_cast<char*>(scratch);
    plan.block_hist = reinterpret_cast<uint32_t*>(b);
    b += (size_t)kHistBlocks * kMaxPasses * kRadix * 4;
    plan.digit_off = reinterpret_cast<uint32_t*>(b);
    b += (size_t)kMaxPasses * kRadix * 4;
    plan.ticket = reinterpret_cast<uint64_t*>(b);
    plan.err = reinterpret_cast<uint32_t*>(b + 8);
    plan.done = reinterpret_cast<uint32_t*>(b + 12);
    b += 64;
    plan.status = reinterpret_cast<uint64_t*>(b);
}

}  // namespace

size_t radix_scratch_bytes(uint32_t n) { return scratch_bytes_for_tiles(radix_num_tiles(n)); }

hipError_t radix_plan_init(RadixSortPlan& plan, uint32_t n, void* scratch, hipStream_t stream) {
    carve(plan, n, scratch);
    // the ticket, error word and histogram-block count at 0; status words at
    // epoch 0 / no flag (one contiguous range)
    return hipMemsetAsync(plan.ticket, 0, 64 + (size_t)plan.num_tiles * kRadix * 8, stream);
}

hipError_t radix_sort_pairs(const RadixSortPlan& plan, uint32_t bits, uint32_t* keys0,
                            uint32_t* vals0, uint32_t* keys1, uint32_t* vals1,
                            uint32_t** keys_out, uint32_t** vals_out, hipStream_t stream,
                            const uint32_t* gate) {
    uint32_t* kin = keys0;
    uint32_t* vin = vals0;
    uint32_t* kout = keys1;
    uint32_t* vout = vals1;
    const uint32_t n = plan.n;
    const uint32_t passes = (bits + 7) / 8;
    if (passes > (uint32_t)kMaxPasses) return hipErrorInvalidValue;
    if (passes > 0) {
        const uint32_t hb = std::min<uint32_t>(kHistBlocks, std::max<uint32_t>(1u, (n + kHistTile - 1) / kHistTile));
        hipLaunchKernelGGL(k_hist<1>, dim3(hb), dim3(kHistThreads), 0, stream, kin, n, bits, plan.block_hist,
                           plan.digit_off, plan.done, gate);
    }
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = 8 * p;
        const uint32_t pb = (bits - shift) < 8 ? (bits - shift) : 8;
        hipLaunchKernelGGL((k_pass<false, false>), dim3(plan.num_tiles), dim3(kSortBlock), 0, stream, kin, vin, kout,
                           vout, n, shift, pb, plan.digit_off + p * kRadix, plan.status, plan.ticket,
                           plan.num_tiles, plan.err, gate);
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

// Stable LSD sort of plan.n (key, value) pairs in place in `pairs` (the
// KeyValuePair layout: key, value interleaved) with `temp` (plan.n pairs) as
// the ping-pong buffer: the first pass reads the pairs and writes a key
// array and a value array into the other buffer, the middle passes sort
// those arrays, the last pass writes pairs again (into `pairs`, or into
// `temp` and a copy back after an odd number of passes).  The plan's
// look-back state is cleared first with fill kernels, so a plan carved from
// any scratch of radix_scratch_bytes(plan.n) bytes works whatever its last
// use was.
hipError_t radix_sort_kv(const RadixSortPlan& plan, uint32_t bits, uint32_t* pairs, uint32_t* temp,
                         hipStream_t stream) {
    const uint32_t n = plan.n;
    const uint32_t passes = (bits + 7) / 8;
    if (passes > (uint32_t)kMaxPasses) return hipErrorInvalidValue;
    if (passes == 0 || n == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(plan.ticket), 0u,
                                   (64 + (size_t)plan.num_tiles * kRadix * 8) / 4, stream);
    if (e != hipSuccess) return e;
    const uint32_t hb = std::min<uint32_t>(kHistBlocks, std::max<uint32_t>(1u, (n + kHistTile - 1) / kHistTile));
    hipLaunchKernelGGL(k_hist<2>, dim3(hb), dim3(kHistThreads), 0, stream, pairs, n, bits, plan.block_hist,
                       plan.digit_off, plan.done, nullptr);
    uint32_t* src = pairs;
    uint32_t* dst = temp;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = 8 * p;
        const uint32_t pb = (bits - shift) < 8 ? (bits - shift) : 8;
        const bool first = p == 0, last = p + 1 == passes;
        const dim3 grid(plan.num_tiles), block(kSortBlock);
        const uint32_t* off = plan.digit_off + p * kRadix;
        if (first && last)
            hipLaunchKernelGGL((k_pass<true, true>), grid, block, 0, stream, src, src + n, dst, dst + n, n, shift, pb,
                               off, plan.status, plan.ticket, plan.num_tiles, plan.err, nullptr);
        else if (first)
            hipLaunchKernelGGL((k_pass<true, false>), grid, block, 0, stream, src, src + n, dst, dst + n, n, shift,
                               pb, off, plan.status, plan.ticket, plan.num_tiles, plan.err, nullptr);
        else if (last)
            hipLaunchKernelGGL((k_pass<false, true>), grid, block, 0, stream, src, src + n, dst, dst + n, n, shift,
                               pb, off, plan.status, plan.ticket, plan.num_tiles, plan.err, nullptr);
        else
            hipLaunchKernelGGL((k_pass<false, false>), grid, block, 0, stream, src, src + n, dst, dst + n, n, shift,
                               pb, off, plan.status, plan.ticket, plan.num_tiles, plan.err, nullptr);
        std::swap(src, dst);
    }
    e = hipGetLastError();
    if (e == hipSuccess && src != pairs)
        e = hipMemcpyAsync(pairs, src, (size_t)n * 8, hipMemcpyDeviceToDevice, stream);
    return e;
}

}  // namespace sph

// --------------------------------------------------------------------------
// The reference's RadixSort seam (source/CUDA/radixsort.cuh:30-32) as C ABI
// entry points: sph_radix_sort allocates its scratch for the call,
// sph_radix_sort_noalloc uses a fixed device array instead.
namespace {

using sph::RadixSortPlan;

// Largest sort the fixed array serves: the reference's 1M-particle scenes.
constexpr uint32_t kNoAllocMaxPairs = SPH_SORT_NOALLOC_MAX;
constexpr size_t kNoAllocBytes = sph::scratch_bytes_for_tiles((kNoAllocMaxPairs + sph::kTile - 1) / sph::kTile);
constexpr int kMaxDevices = 64;

// The allocation-free sort's scratch: one copy per GPU, held until the
// process ends (about 0.8 MB).  Host code uses the per-device address from
// hipGetSymbolAddress (noalloc_ptr), never this name.
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex noalloc_mu;
void* noalloc_ptr[kMaxDevices] = {};

int status_of(hipError_t e) {
    if (e == hipSuccess) return SPH_OK;
    if (e == hipErrorOutOfMemory) return SPH_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SPH_ERR_NO_DEVICE;
    return SPH_ERR_DEVICE;
}

int check_sort_args(const uint32_t* pairs, const uint32_t* temp, uint32_t bits) {
    if (!pairs || !temp || bits > 32 || pairs == temp) return SPH_ERR_INVALID_ARGUMENT;
    return SPH_OK;
}

// Sort, wait for it, and report a look-back wait that hit its bound.
int sort_and_wait(RadixSortPlan& plan, uint32_t bits, uint32_t* pairs, uint32_t* temp, hipStream_t s) {
    hipError_t e = sph::radix_sort_kv(plan, bits, pairs, temp, s);
    uint32_t err = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&err, plan.err, sizeof(err), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    return err ? SPH_ERR_ORDER : SPH_OK;
}

}  // namespace

extern "C" {

int sph_radix_sort(uint32_t* pairs, uint32_t* temp, uint32_t n, uint32_t bits, void* stream) {
    int st = check_sort_args(pairs, temp, bits);
    if (st != SPH_OK) return st;
    if (n > SPH_SORT_MAX) return SPH_ERR_CAPACITY;
    if (n == 0 || bits == 0) return SPH_OK;
    void* buf = nullptr;
    hipError_t e = sph::device_malloc(&buf, sph::radix_scratch_bytes(n));
    if (e != hipSuccess) return status_of(e);
    RadixSortPlan plan;
    sph::carve(plan, n, buf);
    st = sort_and_wait(plan, bits, pairs, temp, static_cast<hipStream_t>(stream));
    (void)hipFree(buf);
    return st;
}

int sph_radix_sort_noalloc(uint32_t* pairs, uint32_t* temp, uint32_t n, uint32_t bits, void* stream) {
    int st = check_sort_args(pairs, temp, bits);
    if (st != SPH_OK) return st;
    if (n > kNoAllocMaxPairs) return SPH_ERR_CAPACITY;
    if (n == 0 || bits == 0) return SPH_OK;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SPH_ERR_NO_DEVICE;
    // one call per device at a time: the array is shared by every call on it
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return e != hipSuccess ? status_of(e) : SPH_ERR_DEVICE;
        noalloc_ptr[dev] = ptr;
    }
    RadixSortPlan plan;
    sph::carve(plan, n, noalloc_ptr[dev]);
    return sort_and_wait(plan, bits, pairs, temp, static_cast<hipStream_t>(stream));
}

}  // extern "C"
