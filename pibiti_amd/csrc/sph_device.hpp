// sph_device.hpp - device helpers shared by the step kernels and the
// learned-surrogate kernels: scene parameter lookup, cell coordinates, the
// equation of state, compensated sums and the culled 27-cell neighbour walk.
#pragma once

#include "sph_internal.hpp"

namespace sph {
namespace dev {

// Parameters of the scene a slot belongs to: the by-value parameters of a
// one-scene launch, or the batch table entry of the slot's scene.
template <bool kBatch>
__device__ __forceinline__ const sph_params& scene_params(const sph_params& one,
                                                         const sph_params* __restrict__ tab,
                                                         const uint32_t* __restrict__ sid,
                                                         uint32_t slot, uint32_t& scene) {
    if constexpr (kBatch) {
        scene = sid[slot];
        return tab[scene];
    } else {
        scene = 0;
        return one;
    }
}

// Cell coordinate of one axis: floor((x - min)/cell), clamped into the grid.
__device__ __forceinline__ int cell_coord(float x, float wmin, float cs, uint32_t gsize, bool& out) {
    float f = floorf((x - wmin) / cs);
    float hi = (float)(gsize - 1);
    if (!(f >= 0.0f && f <= hi)) {
        out = true;
        f = (f != f) ? 0.0f : fminf(fmaxf(f, 0.0f), hi);
    }
    return (int)f;
}

// Incremental re-sort: one bit per slot, set when the slot's cell changed.
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

__device__ __forceinline__ uint32_t grid_hash(const sph_params& p, int gx, int gy, int gz) {
    return p.cell_base + (uint32_t)gz * p.grid_size_yx + (uint32_t)gy * p.grid_size[0] + (uint32_t)gx;
}

__device__ __forceinline__ float eos(const sph_params& p, float rho) {
    if (p.eos == SPH_EOS_TAIT) {
        float x = rho / p.rest_density;
        float x2 = x * x;
        float x4 = x2 * x2;
        float x7 = (x4 * x2) * x;
        float bcoef = (p.stiffness * p.rest_density) / 7.0f;
        return bcoef * (x7 - 1.0f);
    }
    return (rho - p.rest_density) * p.stiffness;
}

// IEEE square root (rounded to nearest).  For x >= 2^-96 the hardware
// square root (within 1 ulp) is moved to a neighbour when the residual
// x - neighbour * s (one fused multiply-add, exact in sign) says that
// neighbour is nearer; smaller arguments take the library routine.
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __int_as_float(__float_as_int(s) - 1);
    const float hi = __int_as_float(__float_as_int(s) + 1);
    float r = __builtin_fmaf(-lo, s, x) <= 0.0f ? lo : s;
    r = __builtin_fmaf(-hi, s, x) > 0.0f ? hi : r;
    return r;
}

__device__ __forceinline__ float sqrt_ieee(float x) {
    if (__builtin_expect(x < 0x1p-96f, 0)) return sqrtf(x);
    return sqrt_rn_normal(x);
}

// The force predicate of compForcePair (source/CUDA/Kernel_Cell.cui:214-218),
// max(minDist, sqrtf(r2)) < h, as one comparison with force_r2_max (the
// largest float r2 that passes; -1 when minDist >= h): r2 <= force_r2_max,
// and also a NaN r2 while minDist < h, because max(minDist, NaN) is minDist
// there (a particle at a NaN position is everyone's force pair in its 27
// cells, as in the reference).
__device__ __forceinline__ bool force_pair(float r2, float force_r2_max) {
    return !(r2 > force_r2_max) && force_r2_max >= 0.0f;
}

struct KahanF {
    float s, c;
    __device__ __forceinline__ void add(float v) {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// ------------------------------------------------------- neighbour-cell walk
// The 27 cells around a particle's cell in the reference's (dz, dy, dx)
// order (source/CUDA/Kernel_Cell.cui:188-192).  A cell whose box is farther
// than h from the particle is skipped: none of its particles at a finite
// position can pass r2 < h2 (or the force predicate, which is stricter), so
// its partial sum would be +0 and adding +0 leaves the total unchanged.  The
// test is conservative: each face distance is shrunk by a fuzz covering the
// rounding of cell assignment, and the bound carries a 1e-5 relative margin.
// A particle at a NaN position passes the force predicate from anywhere in
// its 27 cells (force_pair), so a step whose integration produced one
// (counter kCntNaN) walks all 27 cells (those inside the grid) unculled.
struct AxisWalk {
    int g;           // own cell coordinate
    int gmax;        // grid size - 1
    float lo2, hi2;  // squared distance to the own cell's lower/upper face (minus fuzz)

    __device__ __forceinline__ void init(float x, float wl, float wh, float cs, uint32_t gsize) {
        bool out = false;
        g = cell_coord(x, wl, cs, gsize, out);
        gmax = (int)gsize - 1;
        const float fuzz = 1e-6f * fmaxf(fabsf(wl), fabsf(wh)) + 1e-30f;
        const float clo = wl + (float)g * cs;
        const float chi = clo + cs;
        const float lo = fmaxf(x - clo - fuzz, 0.0f);
        const float hi = fmaxf(chi - x - fuzz, 0.0f);
        lo2 = lo * lo;
        hi2 = hi * hi;
    }
    // squared distance component of a neighbour offset d in {-1, 0, 1}
    __device__ __forceinline__ float d2(int d) const { return d < 0 ? lo2 : (d > 0 ? hi2 : 0.0f); }
    __device__ __forceinline__ bool ok(int d) const { return g + d >= 0 && g + d <= gmax; }
};

// Whether the step's walks may cull cells: no particle sits at a NaN
// position (the integration counts them in counters[kCntNaN]).
__device__ __forceinline__ bool walk_culled(const uint32_t* __restrict__ counters) {
    return counters[kCntNaN] == 0u;
}

struct CellWalk {
    AxisWalk ax, ay, az;
    float cull2;        // squared cull radius
    uint32_t base;      // cell key of (0, 0, 0) of this scene
    uint32_t gyx, gx;   // key strides
    uint32_t cap;       // max_par_in_cell

    // cull = false: no cell of the 27 is skipped for its distance.
    __device__ __forceinline__ void init(const sph_params& p, float x, float y, float z, bool cull) {
        ax.init(x, p.world_min[0], p.world_max[0], p.cell_size[0], p.grid_size[0]);
        ay.init(y, p.world_min[1], p.world_max[1], p.cell_size[1], p.grid_size[1]);
        az.init(z, p.world_min[2], p.world_max[2], p.cell_size[2], p.grid_size[2]);
        cull2 = cull ? p.h2 * 1.00001f : __builtin_huge_valf();
        base = p.cell_base;
        gyx = p.grid_size_yx;
        gx = p.grid_size[0];
        cap = p.max_par_in_cell;
    }
    // Slot range [s, e) of neighbour cell (dx, dy, dz), capped at max_par_in_cell;
    // false if the cell is outside the grid, culled or empty (s and e are then
    // meaningless).
    __device__ __forceinline__ bool open(int dx, int dy, int dz,
                                         const uint32_t* __restrict__ cell_start,
                                         const uint32_t* __restrict__ cell_end, uint32_t& s,
                                         uint32_t& e) const {
        if (!ax.ok(dx) || !ay.ok(dy) || !az.ok(dz)) return false;
        if (az.d2(dz) + ay.d2(dy) + ax.d2(dx) > cull2) return false;
        const uint32_t c = base + (uint32_t)(az.g + dz) * gyx + (uint32_t)(ay.g + dy) * gx +
                           (uint32_t)(ax.g + dx);
        // both loads go out together (an empty cell's end is stale and unused)
        s = cell_start[c];
        e = min(cell_end[c], s + cap);
        return s != kEmptyCell;
    }
    // The row of neighbour cells (dy, dz): whether it lies in the grid, its
    // squared distance so far and the key of its dx = 0 cell, computed once
    // for the row's three cells (the same values open() computes per cell).
    struct Row {
        bool ok;
        float d2;
        uint32_t key;
    };
    __device__ __forceinline__ Row row(int dy, int dz) const {
        Row r;
        r.ok = az.ok(dz) && ay.ok(dy);
        r.d2 = az.d2(dz) + ay.d2(dy);
        r.key = base + (uint32_t)(az.g + dz) * gyx + (uint32_t)(ay.g + dy) * gx + (uint32_t)ax.g;
        return r;
    }
    // open() for cell dx of a row.
    __device__ __forceinline__ bool open_in(const Row& r, int dx, const uint32_t* __restrict__ cell_start,
                                            const uint32_t* __restrict__ cell_end, uint32_t& s, uint32_t& e) const {
        if (!r.ok || !ax.ok(dx)) return false;
        if (r.d2 + ax.d2(dx) > cull2) return false;
        const uint32_t c = r.key + (uint32_t)dx;
        s = cell_start[c];
        e = min(cell_end[c], s + cap);
        return s != kEmptyCell;
    }
    // Cell key of the particle's own cell.
    __device__ __forceinline__ uint32_t own_key() const {
        return base + (uint32_t)az.g * gyx + (uint32_t)ay.g * gx + (uint32_t)ax.g;
    }
    // As open(), with the (start, end) pairs of the block's neighbour rows
    // staged in LDS: row q = 3 (dz + 1) + (dy + 1), entry xbase + dx, where
    // xbase = own key - (first staged key of the row).
    __device__ __forceinline__ bool open_staged(int dx, int dy, int dz, const uint2* tab, int cells,
                                                int xbase, uint32_t& s, uint32_t& e) const {
        if (!ax.ok(dx) || !ay.ok(dy) || !az.ok(dz)) return false;
        if (az.d2(dz) + ay.d2(dy) + ax.d2(dx) > cull2) return false;
        const uint2 r = tab[(3 * (dz + 1) + (dy + 1)) * cells + xbase + dx];
        s = r.x;
        if (s == kEmptyCell) return false;
        e = min(r.y, s + cap);
        return true;
    }
};

// ------------------------------------------------------- neighbour lists
// Group-major layout: entries 4g .. 4g + 3 of slot i's list are the uint4 at
// index g * n + i (n = slots of the launch), so the lanes of a wave, which
// hold consecutive slots, read and write consecutive 16-byte words.
__device__ __forceinline__ size_t nbr_group_index(uint32_t g, uint32_t n, uint32_t i) {
    return (size_t)g * n + i;
}

// Appends a slot's neighbours through a ring of kRingSlots entries per lane
// in LDS (slot s of lane t at ring[s * kRingLanes + t]): every candidate is
// written at the position of the next entry and only a hit advances the
// count, so lanes never branch per candidate.  After each group of at most
// four candidates a completed group of four entries (there is at most one)
// leaves as one 16-byte store; finish() stores the partly filled last group
// (its unused entries are never read).  Entries past kNbrCap are counted
// but not stored.  Every block that builds lists has kRingLanes threads.
constexpr uint32_t kRingSlots = 8;
constexpr uint32_t kRingLanes = 256;

struct ListRing {
    uint32_t* slot0;  // this lane's slot 0
    uint32_t count;
    uint32_t flushed;

    __device__ __forceinline__ void init(uint32_t* ring) {
        slot0 = ring + threadIdx.x;
        count = 0u;
        flushed = 0u;
    }
    __device__ __forceinline__ void put(uint32_t j, bool hit) {
        slot0[(count & (kRingSlots - 1u)) * kRingLanes] = j;
        count += hit ? 1u : 0u;
    }
    __device__ __forceinline__ uint4 group_at(uint32_t first) const {
        const uint32_t b = first & (kRingSlots - 1u);
        return make_uint4(slot0[b * kRingLanes], slot0[(b + 1u) * kRingLanes], slot0[(b + 2u) * kRingLanes],
                          slot0[(b + 3u) * kRingLanes]);
    }
    __device__ __forceinline__ void drain(uint4* __restrict__ nbr, uint32_t n, uint32_t i) {
        if (count - flushed >= 4u) {
            if (flushed < kNbrCap) nbr[nbr_group_index(flushed >> 2, n, i)] = group_at(flushed);
            flushed += 4u;
        }
    }
    __device__ __forceinline__ void finish(uint4* __restrict__ nbr, uint32_t n, uint32_t i) {
        if (count > flushed && flushed < kNbrCap) nbr[nbr_group_index(flushed >> 2, n, i)] = group_at(flushed);
    }
};

// XCD-aware tile order: the hardware hands consecutive workgroups to the
// eight XCDs in turn; renumbering gives each XCD a contiguous run of tiles,
// so tiles that share neighbour cells also share an L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t g) {
    const uint32_t x = b % kXcds, local = b / kXcds;
    const uint32_t per = g / kXcds, rem = g % kXcds;
    return x * per + min(x, rem) + local;
}

// The cell tables of a block's neighbour rows, staged in LDS.
struct CellView {
    const uint2* t;   // 9 rows of `cells` (start, end) pairs
    int cells;
    uint32_t kmin;    // smallest own-cell key of the block
};

// Four floats at a 4-byte aligned address: a 16-byte global load needs only
// dword alignment.  Four candidates of the packed position array (three
// floats a slot) are three of them, where the float4 array takes four; the
// vector memory unit's cost goes by the wave's load instructions, not by
// their width.
struct alignas(4) F4 {
    float x, y, z, w;
};

// The positions of four consecutive slots from the packed array, q = the
// first one's x (the array is padded past its end, so the loads need no
// clamp).
__device__ __forceinline__ void packed_group(const float* __restrict__ xq, float (&cx)[4], float (&cy)[4],
                                             float (&cz)[4]) {
    const F4* q = reinterpret_cast<const F4*>(xq);
    const F4 a = q[0];  // x0 y0 z0 x1
    const F4 b = q[1];  // y1 z1 x2 y2
    const F4 d = q[2];  // z2 x3 y3 z3
    cx[0] = a.x, cy[0] = a.y, cz[0] = a.z;
    cx[1] = a.w, cy[1] = b.x, cz[1] = b.y;
    cx[2] = b.z, cy[2] = b.w, cz[2] = d.x;
    cx[3] = d.y, cy[3] = d.z, cz[3] = d.w;
}

}  // namespace dev
}  // namespace sph
