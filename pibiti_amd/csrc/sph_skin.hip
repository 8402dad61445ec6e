// sph_skin.hip - density and neighbour lists of the learned surrogate's
// steps between sorts.
//
// A surrogate step needs, per particle, its density and the list of
// neighbours within h.  Sorting every step costs a sixth of the step, so
// the surrogate sorts once per run of steps and the steps in between keep
// the slot order and the cell ranges of that sort: a slot stays listed in
// the cell that held it at the sort.  While no particle has moved more than
// d = skin h since the sort, a particle j now within h of particle i was
// within h + d of i's current position at the sort, so walking every cell
// within h + d of i finds it.  The kernel counts particles that have moved
// farther, and the caller sorts again before a step that has any.
//
// The walk covers the cells that the (h + d) sphere reaches: two to four
// per axis, z, y, x ascending, slots ascending within a cell (the order of
// the 27-cell walk), culled by the distance from the particle to the cell's
// box; cells hold at most max_par_in_cell slots, as in the plain walk.
#include "sph_device.hpp"

namespace sph {
namespace {

using namespace dev;

constexpr int kBlock = 256;
static_assert(kBlock == (int)kRingLanes, "list rings are sized for the block");

// Cells of one axis that [x - r, x + r] reaches, clamped into the grid, and
// the squared distance from x to a cell (zero inside), shrunk by a fuzz that
// covers the rounding of cell assignment so that culling stays conservative.
struct ReachAxis {
    int lo, hi;
    float x, wl, cs, fuzz;

    __device__ __forceinline__ void init(float x_, float r, float wl_, float wh, float cs_, uint32_t gsize) {
        bool out = false;
        lo = cell_coord(x_ - r, wl_, cs_, gsize, out);
        hi = cell_coord(x_ + r, wl_, cs_, gsize, out);
        x = x_;
        wl = wl_;
        cs = cs_;
        fuzz = 1e-6f * fmaxf(fabsf(wl_), fabsf(wh)) + 1e-30f;
    }
    __device__ __forceinline__ float d2(int g) const {
        const float clo = wl + (float)g * cs;
        const float chi = clo + cs;
        const float d = fmaxf(fmaxf(clo - x, x - chi) - fuzz, 0.0f);
        return d * d;
    }
};

__global__ __launch_bounds__(kBlock) void k_density_reach(sph_params p, uint32_t n, float reach, float move2,
                                                         const float4* __restrict__ pos,
                                                         const float4* __restrict__ bpos,
                                                         const uint32_t* __restrict__ cell_start,
                                                         const uint32_t* __restrict__ cell_end,
                                                         float2* __restrict__ rhop, uint4* __restrict__ nbr,
                                                         uint32_t* __restrict__ nnb,
                                                         uint32_t* __restrict__ counters) {
    __shared__ uint32_t ring[kRingSlots * kRingLanes];
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    uint32_t listed = 0;
    if (i < n) {
        const float4 xi = pos[i];
        const float4 x0 = bpos[i];
        const float mx = xi.x - x0.x, my = xi.y - x0.y, mz = xi.z - x0.z;
        if (mx * mx + my * my + mz * mz > move2) atomicAdd(&counters[kCntSkinMoved], 1u);
        ReachAxis ax, ay, az;
        ax.init(xi.x, reach, p.world_min[0], p.world_max[0], p.cell_size[0], p.grid_size[0]);
        ay.init(xi.y, reach, p.world_min[1], p.world_max[1], p.cell_size[1], p.grid_size[1]);
        az.init(xi.z, reach, p.world_min[2], p.world_max[2], p.cell_size[2], p.grid_size[2]);
        const float cull2 = (reach * reach) * 1.00001f;
        const float h2 = p.h2;
        const uint32_t cap = p.max_par_in_cell;
        ListRing lw;
        lw.init(ring);
        float sum = 0.0f;
        for (int gz = az.lo; gz <= az.hi; ++gz) {
            const float dz2 = az.d2(gz);
            if (dz2 > cull2) continue;
            for (int gy = ay.lo; gy <= ay.hi; ++gy) {
                const float dzy2 = dz2 + ay.d2(gy);
                if (dzy2 > cull2) continue;
                const uint32_t rowkey =
                    p.cell_base + (uint32_t)gz * p.grid_size_yx + (uint32_t)gy * p.grid_size[0];
                for (int gx = ax.lo; gx <= ax.hi; ++gx) {
                    if (dzy2 + ax.d2(gx) > cull2) continue;
                    const uint32_t c = rowkey + (uint32_t)gx;
                    const uint32_t s = cell_start[c];
                    if (s == kEmptyCell) continue;
                    const uint32_t e = min(cell_end[c], s + cap);
                    float cell = 0.0f;
                    // candidates four at a time (indices past the cell repeat its last slot)
                    for (uint32_t j0 = s; j0 < e; j0 += 4) {
                        float4 xq[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) xq[k] = pos[min(j0 + k, e - 1)];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const uint32_t j = j0 + k;
                            const float rx = xi.x - xq[k].x, ry = xi.y - xq[k].y, rz = xi.z - xq[k].z;
                            const float r2 = rx * rx + ry * ry + rz * rz;
                            const bool hit = (r2 < h2) && (j != i) && (j < e);
                            const float c3 = h2 - r2;
                            const float wgt = (c3 * c3) * c3;
                            cell = cell + (hit ? wgt : 0.0f);
                            lw.put(j, hit);
                        }
                        lw.drain(nbr, n, i);
                    }
                    sum = sum + cell;
                }
            }
        }
        lw.finish(nbr, n, i);
        nnb[i] = lw.count;
        if (lw.count > kNbrCap) atomicAdd(&counters[kCntNbrOverflow], 1u);
        const float rho = (sum * p.poly6_kern) * p.particle_mass;
        rhop[i] = make_float2(rho, eos(p, rho));
        listed = min(lw.count, kNbrCap);
    }
    for (int off = 32; off > 0; off >>= 1) listed += __shfl_xor(listed, off);
    if ((threadIdx.x & 63) == 0 && listed) atomicAdd(&counters[kCntPairs], listed);
}

}  // namespace

hipError_t launch_density_reach(const LaunchCfg& c, const StepBuffers& b, float skin, const float4* bpos,
                                uint32_t* nbr, uint32_t* nnb, hipStream_t s) {
    if (c.tab || !(skin > 0.0f) || c.n == 0) return hipErrorInvalidValue;
    const float d = skin * c.one.h;
    const uint32_t blocks = (c.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_density_reach, dim3(blocks), dim3(kBlock), 0, s, c.one, c.n, c.one.h + d, d * d,
                       b.pos_cur, bpos, b.cell_start, b.cell_end, b.rhop, reinterpret_cast<uint4*>(nbr), nnb,
                       b.counters);
    return hipGetLastError();
}

}  // namespace sph
