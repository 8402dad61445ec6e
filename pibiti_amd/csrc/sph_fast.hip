// sph_fast.hip - the density gather of the FAST precision mode.
//
// The same step as the reference-order kernels (sph_kernels.hip), under the
// float contract of the task rather than bit-exactness: integer outputs stay
// exact (cell keys, order, cell ranges, and the neighbour sets with the
// per-cell cap and both predicates, r2 < h2 for the density and
// r2 <= force_r2_max for the force), every float is within 1e-5 of the FP32
// restatement of the reference and of the FP64 twin, condition-scaled.  That
// frees the arithmetic: the density is one running sum per particle with
// fused multiply-adds instead of per-cell partial sums added in cell order
// (source/CUDA/Kernel_Cell.cui:142-199); the candidate walk and predicates
// are the reference's.  The force is the reference order's pair-entry
// kernel (k_force_entries, sph_kernels.hip) with one running sum instead of
// per-cell partial sums.
//
// Pair lists.  The density walk tests candidates four at a time (four
// consecutive slots of one row's run); a group with at least one force pair
// leaves one list entry, (first slot << kPairSlotShift) | 4-bit mask of the
// pairs (no new-cell flag: FAST keeps no per-cell partial sums), so the walk
// spends one LDS write per group rather than per pair.  Entries are
// group-major (entries 4g .. 4g + 3 of slot i in the uint4 at g n + i) and
// there is room for every group the capped walk can produce (27 cells of at
// most ceil(max_par_in_cell / 4) groups), so the lists never overflow.  The
// count word of a slot holds the entries (low 16 bits) and the pairs (high
// 16 bits).
#include "sph_device.hpp"

namespace sph {
namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr uint32_t kRing = 4;  // ring entries per lane (a group of four adds at most one)
constexpr int kGroup = 4;      // candidates per iteration of the density walk

// Density + EOS, the force pair list and the force record of one slot per
// thread.
//
// The walk goes by rows: the neighbour cells (dx = -1, 0, 1) of one (dy, dz)
// have consecutive keys, so their slots form one run, and the candidates of
// a row are that run, four consecutive slots at a time (three 16-byte loads
// from the packed positions).  A run breaks only after a cell capped at
// max_par_in_cell (its slots past the cap are not candidates); the slots
// before the break are walked as a run of their own, and the own row's run
// is split at the particle itself.  The lanes of one cell then walk the same
// runs in step (in the own row, nearly) and load the same cache lines, and a
// row costs one loop instead of three.  Culling drops a side
// cell of a row whose box is farther than h (the middle cell of a row that
// is kept is never farther than the row).
//
// The force record of slot i is two float4: (x, y, z, rho) and
// (vx, vy, vz, p), everything the force reads of a particle in one aligned
// 32-byte block.
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_density_fast(sph_params one, const sph_params* __restrict__ tab,
                                                         const uint32_t* __restrict__ sid, uint32_t n,
                                                         const float4* __restrict__ pos,
                                                         const float* __restrict__ xyz,
                                                         const float4* __restrict__ vel,
                                                         const uint32_t* __restrict__ cell_start,
                                                         const uint32_t* __restrict__ cell_end,
                                                         float2* __restrict__ rhop, float4* __restrict__ rec,
                                                         uint4* __restrict__ nbr, uint32_t* __restrict__ nnb,
                                                         const uint32_t* __restrict__ counters) {
    __shared__ uint32_t ring[kRing * kBlock];
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, i, scene);
    const float4 xi = pos[i];
    CellWalk w;
    w.init(p, xi.x, xi.y, xi.z, walk_culled(counters));
    const float h2 = p.h2, r2f = p.force_r2_max;
    const float inv_h2 = 1.0f / h2;
    uint32_t* const slot0 = ring + threadIdx.x;
    uint32_t count = 0, flushed = 0, pairs = 0;
    float sum = 0.0f;  // of (1 - r2 / h2)^3; times h2^3 below
    // one list entry per four consecutive candidates with a pair; the
    // written slot is only kept when the count advances
    auto append = [&](uint32_t tag, uint32_t bits) {
        slot0[(count & (kRing - 1u)) * kBlock] = tag | bits;
        count += bits ? 1u : 0u;
        pairs += (uint32_t)__popc(bits);
        if (count - flushed == kRing) {
            nbr[(size_t)(flushed >> 2) * n + i] =
                make_uint4(slot0[0], slot0[kBlock], slot0[2 * kBlock], slot0[3 * kBlock]);
            flushed += kRing;
        }
    };
    // The candidates j0 .. j0 + 3; in the tail group of a run only the
    // first `left` of them (the others take r2 = +inf, which fails both
    // predicates).  The predicates use the reference's r2 exactly.  The
    // density weight is (1 - r2 / h2)^3 clamped to [0, 1], one fused
    // multiply-add with an output clamp: 0 for r2 >= h2 and for a NaN r2,
    // as the reference's (r2 < h2 ? (h2 - r2)^3 : 0) is.
    // The candidates' positions come from the packed array (packed_group:
    // three loads for four candidates); those loads are what bounds this
    // walk.
    auto group = [&](const float* q, uint32_t tag, uint32_t left, auto tail) {
        constexpr bool kTail = decltype(tail)::value;
        float cx[kGroup], cy[kGroup], cz[kGroup];
        packed_group(q, cx, cy, cz);
        uint32_t bits = 0;
#pragma unroll
        for (int k = 0; k < kGroup; ++k) {
            const float rx = xi.x - cx[k], ry = xi.y - cy[k], rz = xi.z - cz[k];
            float r2 = rx * rx + ry * ry + rz * rz;
            if constexpr (kTail)
                if (k > 0) r2 = left > (uint32_t)k ? r2 : __builtin_huge_valf();
            const float c = fminf(fmaxf(__builtin_fmaf(r2, -inv_h2, 1.0f), 0.0f), 1.0f);
            sum = __builtin_fmaf(c * c, c, sum);
            bits |= force_pair(r2, r2f) ? (1u << k) : 0u;
        }
        append(tag, bits);
    };
    // the candidates [s, e): full groups, then one tail group
    // (tag: the group's first slot, shifted into place for its entry)
    auto walk_run = [&](uint32_t s, uint32_t e) {
        const float* q = xyz + 3 * (size_t)s;
        uint32_t tag = s << kPairSlotShift, left = e - s;
        for (; left >= (uint32_t)kGroup; left -= kGroup, q += 3 * kGroup, tag += kGroup << kPairSlotShift)
            group(q, tag, kGroup, std::false_type{});
        if (left) group(q, tag, left, std::true_type{});
    };
    // a run of the particle's own row without the particle itself (it is in
    // the run unless its cell is capped before it)
    auto walk_own = [&](uint32_t s, uint32_t e) {
        if (i >= s && i < e) {
            walk_run(s, i);
            walk_run(i + 1, e);
        } else {
            walk_run(s, e);
        }
    };
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy) {
            const CellWalk::Row row = w.row(dy, dz);
            if (!row.ok) continue;
            const bool own = (dy | dz) == 0;
            uint32_t rs = 0, re = 0;  // the run so far (empty: rs == re)
#pragma unroll 1
            for (int dx = -1; dx <= 1; ++dx) {
                uint32_t s, e;
                if (!w.open_in(row, dx, cell_start, cell_end, s, e)) continue;
                if (s != re) {
                    // a break: a capped cell before (or the row's first open cell)
                    if (own)
                        walk_own(rs, re);
                    else
                        walk_run(rs, re);
                    rs = s;
                }
                re = e;
            }
            if (own)
                walk_own(rs, re);
            else
                walk_run(rs, re);
        }
    }
    if (count > flushed)
        nbr[(size_t)(flushed >> 2) * n + i] = make_uint4(slot0[0], slot0[kBlock], slot0[2 * kBlock], slot0[3 * kBlock]);
    nnb[i] = count | (pairs << 16);
    const float rho = (sum * (((p.poly6_kern * h2) * h2) * h2)) * p.particle_mass;
    const float pr = eos(p, rho);
    rhop[i] = make_float2(rho, pr);
    const float4 vi = vel[i];
    rec[2 * (size_t)i] = make_float4(xi.x, xi.y, xi.z, rho);
    rec[2 * (size_t)i + 1] = make_float4(vi.x, vi.y, vi.z, pr);
}

inline uint32_t blocks_for(uint32_t n) { return n == 0 ? 1u : (n + kBlock - 1) / kBlock; }

}  // namespace

uint32_t fast_list_groups(uint32_t max_par_in_cell) { return (27u * ((max_par_in_cell + 3u) / 4u) + 3u) / 4u; }

hipError_t launch_density_fast(const LaunchCfg& c, const StepBuffers& b, float4* rec, uint4* nbr, uint32_t* nnb,
                               hipStream_t s) {
    if (!b.xyz) return hipErrorInvalidValue;  // the place kernel packs the positions
    const dim3 grid(blocks_for(c.n)), block(kBlock);
    if (c.tab)
        hipLaunchKernelGGL(k_density_fast<true>, grid, block, 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur,
                           b.xyz, b.vel_cur, b.cell_start, b.cell_end, b.rhop, rec, nbr, nnb, b.counters);
    else
        hipLaunchKernelGGL(k_density_fast<false>, grid, block, 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur,
                           b.xyz, b.vel_cur, b.cell_start, b.cell_end, b.rhop, rec, nbr, nnb, b.counters);
    return hipGetLastError();
}


}  // namespace sph
