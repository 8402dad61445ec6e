// sph_mlp.hip - learned pair-force surrogate on BF16 matrix cores (gfx950).
//
// Work decomposition.  The density kernel lists every particle's
// neighbours (r2 < h2, j != i, walk order) in group-major lists.  One
// wavefront serves 32 target particles (consecutive sorted slots): tile t
// holds each target's t-th neighbour as one of 32 pair columns, lane halves
// g = 0/1 holding the two halves of the pair's 16 input features, and both
// pair layers run on v_mfma_f32_32x32x16_bf16.  The columns of a tile are
// distinct targets, so the neighbour sum accumulates in registers: no
// atomics, and a fixed summation order (the list order).  A wave runs as
// many tiles as its largest neighbour count; shorter lists are padded with
// masked pairs.
//
// Layouts (operand conventions in sph_mfma.hpp).  The forward pass runs
// transposed, weights as the A operand and pairs as columns (lane = pair):
//   C1^T (units x pairs) = W1 (units x features) * E^T
//   C2^T (units x pairs) = W2 * relu(C1^T) + b2
// Lane l of C1^T holds pair l%32 and units 32t + row(i, l).  Read as the B
// operand of the next product, those 16 values form two k-blocks, so the
// hidden units enter layer 2 in the order
//   unit(kb, g, e) = 32(kb/2) + 16(kb%2) + 8(e/4) + 4g + e%4   (g = l/32)
// and W2 is packed with its input units in that order (any permutation of
// the contraction index is exact as long as both operands share it).
// Swapping the A and B operands of a product yields its transpose with lane
// = unit and the pairs along the accumulator rows: the backward pass uses
// this to form every weight gradient (contraction over pairs) directly from
// registers, with no LDS transposes.  The transposed copy of E comes from
// one product with the identity.
#include <cstring>

#include "sph_device.hpp"
#include "sph_mfma.hpp"
#include "sph_mlp.hpp"

namespace sph {
namespace {

using namespace dev;
using namespace mlp;

constexpr int kWaves = 4;  // waves per block
constexpr int kBlk = 64 * kWaves;
constexpr int kGroup = 32;     // target particles per wave
constexpr int kDhStride = 68;  // LDS row stride (floats) of a staged dL/dH row
constexpr int kNodeBlk = 256;
constexpr uint32_t kMaxGradBlocks = 512;

__device__ __forceinline__ int hid_perm(int kb, int g, int e) {
    return 32 * (kb >> 1) + 16 * (kb & 1) + 8 * (e >> 2) + 4 * g + (e & 3);
}

__device__ __forceinline__ float relu(float x) { return x > 0.0f ? x : 0.0f; }

__device__ __forceinline__ __bf16 bf(float x) { return (__bf16)x; }

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
#pragma unroll
    for (int e = 0; e < 16; ++e) z[e] = 0.0f;
    return z;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct MlpArgs {
    sph_params one;
    const sph_params* tab;
    MlpState st;
    sph_mlp_norm nm;
    const float* params;
    const bf16x8* w1f;
    const bf16x8* w2f;
    const bf16x8* w2tf;
    uint32_t ngroups;
    float4* vel_out;      // inference
    float4* accel_out;    // inference
    const float* dt_dev;  // inference
    float* h;             // training forward
    float4* out;          // training forward
    const float* dh;      // backward
    float* part_pair;     // backward
};

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short short2v __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Two floats as a packed BF16 pair (round to nearest even).
__device__ __forceinline__ uint32_t cvt2(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// What one lane gathers for one pair: the neighbour's position (g = 0
// lanes) or velocity (g = 1 lanes), and its density.
struct PairRaw {
    float4 a;
    float b;
};

// A lane's stream over its target's neighbours.  Entry t of target i is
// word 4 (floor(t / 4) n + i) + t % 4 of the group-major lists; entries at
// and past the target's count stand for the target itself, so every load
// is in bounds and no lane branches (those pairs are masked out of the
// features).
struct PairStream {
    const uint32_t* words;  // this lane's target's words of group 0
    const float4* src;      // positions (g = 0) or velocities (g = 1)
    const float2* rhop;
    size_t gstride;         // words from one group to the next: 4 n
    uint32_t self;
    int cnt;

    __device__ __forceinline__ uint32_t index(int t) const {
        const size_t off = t < cnt ? (size_t)(t >> 2) * gstride + (size_t)(t & 3) : 0;
        const uint32_t w = words[off];
        return t < cnt ? w : self;
    }
    __device__ __forceinline__ PairRaw load(uint32_t j) const {
        PairRaw r;
        r.a = src[j];
        r.b = rhop[j].x;
        return r;
    }
};

// Per-lane constants of a target for the edge features: T is the target's
// position (g = 0) or velocity (g = 1), s the scale with the sign that
// turns (neighbour - target) into the feature, w2 the fixed third word.
struct PairConst {
    float4 T;
    float s;
    uint32_t w2;
};

// Edge features of one pair as this lane's half of the k dimension, the
// same code on both halves:
//   g = 0 (k 0..7):  (x_i - x_j)/h (3), |x_i - x_j|/h, s(rho_i/rho0 - 1), 1, 0, 0
//   g = 1 (k 8..15): (v_j - v_i)/v (3), s(rho_j/rho0 - 1), 0 (4)
// (x_j - x_i) (-1/h) equals (x_i - x_j) (1/h) bit for bit.
__device__ __forceinline__ bf16x8 edge_features(int g, bool valid, const PairRaw& r, const PairConst& k,
                                                const sph_mlp_norm& nm) {
    const float dx = r.a.x - k.T.x, dy = r.a.y - k.T.y, dz = r.a.z - k.T.z;
    const float r2 = (dx * dx + dy * dy) + dz * dz;
    const float dist = sqrt_rn_normal(r2) * nm.inv_h;  // r2 is 0 or normal for a real pair
    const float rj = (r.b * nm.inv_rho0 - 1.0f) * nm.rho_scale;
    u32x4 w;
    w[0] = valid ? cvt2(dx * k.s, dy * k.s) : 0u;
    w[1] = valid ? cvt2(dz * k.s, g ? rj : dist) : 0u;
    w[2] = valid ? k.w2 : 0u;
    w[3] = 0u;
    return __builtin_bit_cast(bf16x8, w);
}

// Layer 1, transposed: the hidden units of pair (lane % 32) as the four
// k-blocks of layer 2's B operand.
typedef short short8 __attribute__((ext_vector_type(8)));

// relu on packed BF16: a negative BF16 has its sign bit set, so as a 16-bit
// integer it is negative, and max(x, 0) over int16 lanes gives
// bf16(relu(v)) == relu(bf16(v)) two values per instruction.
__device__ __forceinline__ bf16x8 relu_bf16(const bf16x8& v) {
    const short8 s = __builtin_elementwise_max(__builtin_bit_cast(short8, v), (short8)0);
    return __builtin_bit_cast(bf16x8, s);
}

__device__ __forceinline__ bf16x8 keep_if(const bf16x8& v, bool keep) {
    const short8 s = keep ? __builtin_bit_cast(short8, v) : (short8)0;
    return __builtin_bit_cast(bf16x8, s);
}

// Two floats to a relu'd BF16 pair: one packed conversion, one packed
// 16-bit max.
__device__ __forceinline__ uint32_t cvt_relu2(float a, float b) {
    const short2v s = __builtin_bit_cast(short2v, cvt2(a, b));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(s, (short2v)0));
}

// 16 accumulator values as two k-blocks of a B operand (BF16, relu applied).
__device__ __forceinline__ void pack_relu(const f32x16& c, bf16x8& lo, bf16x8& hi) {
    const u32x4 a = {cvt_relu2(c[0], c[1]), cvt_relu2(c[2], c[3]), cvt_relu2(c[4], c[5]), cvt_relu2(c[6], c[7])};
    const u32x4 b = {cvt_relu2(c[8], c[9]), cvt_relu2(c[10], c[11]), cvt_relu2(c[12], c[13]),
                     cvt_relu2(c[14], c[15])};
    lo = __builtin_bit_cast(bf16x8, a);
    hi = __builtin_bit_cast(bf16x8, b);
}

__device__ __forceinline__ void layer1_t(const bf16x8 (&w1f)[2], const bf16x8& E, bf16x8 (&h1t)[4]) {
#pragma unroll
    for (int t1 = 0; t1 < 2; ++t1) {
        const f32x16 c = mfma32(w1f[t1], E, zero16());
        pack_relu(c, h1t[2 * t1], h1t[2 * t1 + 1]);
    }
}

// Identity fragments that move a B operand in the permuted hidden order
// (rows of the transposed accumulator) back into the accumulator layout:
// I_par[e] = 1 where hidden unit 16 par + 8 (e / 4) + 4 g + e % 4 is the
// lane's row p within its 32-unit tile.
__device__ __forceinline__ bf16x8 unit_identity(int par, int g, int p) {
    bf16x8 f;
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] = bf((16 * par + 8 * (e >> 2) + 4 * g + (e & 3)) == p ? 1.0f : 0.0f);
    return f;
}

// Per-wave target group: slots [32 grp, 32 grp + 32), one target per lane
// pair (lane % 32), its neighbours processed as one tile row per neighbour.
struct Group {
    uint32_t iT;   // target slot of this lane
    uint32_t ii;   // iT, or the group's first slot for a lane past the end
    int cT;        // its listed neighbours
    int m;         // tiles (largest count of the group)
    float4 xT, vT;
    float rT;
};

__device__ __forceinline__ Group open_group(uint32_t grp, int p, const MlpState& st) {
    Group G;
    G.iT = grp * kGroup + p;
    const bool act = G.iT < st.n;
    G.ii = act ? G.iT : grp * kGroup;
    G.cT = act ? (int)min(st.nnb[G.ii], kNbrCap) : 0;
    G.m = wave_max(G.cT);
    G.xT = st.pos[G.ii];
    G.vT = st.vel[G.ii];
    G.rT = st.rhop[G.ii].x;
    return G;
}

__device__ __forceinline__ PairStream pair_stream(const Group& G, int g, const MlpState& st) {
    PairStream ps;
    ps.words = st.nbr + 4 * (size_t)G.ii;
    ps.src = g ? st.vel : st.pos;
    ps.rhop = st.rhop;
    ps.gstride = 4 * (size_t)st.n;
    ps.self = G.ii;
    ps.cnt = G.cT;
    return ps;
}

__device__ __forceinline__ PairConst pair_const(const Group& G, int g, const sph_mlp_norm& nm) {
    PairConst k;
    k.T = g ? G.vT : G.xT;
    k.s = g ? nm.inv_v : -nm.inv_h;
    k.w2 = g ? 0u : cvt2((G.rT * nm.inv_rho0 - 1.0f) * nm.rho_scale, 1.0f);
    return k;
}

// Inference (kTrain = false: velocity update and/or acceleration) and the
// training forward pass (kTrain = true: neighbour sums H and outputs).
template <bool kBatch, bool kTrain>
__global__ __launch_bounds__(kBlk) void k_mlp_fwd(const MlpArgs a) {
    __shared__ float s_w3[kNodeParams];
    __shared__ float s_b2[kHid];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int g = lane >> 5, p = lane & 31;
    for (int k = threadIdx.x; k < kNodeParams; k += kBlk) s_w3[k] = a.params[kOffW3 + k];
    for (int k = threadIdx.x; k < kHid; k += kBlk) s_b2[k] = a.params[kOffB2 + k];
    __syncthreads();
    const sph_mlp_norm nm = a.nm;
    const MlpState st = a.st;

    bf16x8 w1f[2], w2f[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t) w1f[t] = a.w1f[t * 64 + lane];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) w2f[mt][kb] = a.w2f[(mt * 4 + kb) * 64 + lane];
    // b2 enters layer 2 as one more k-block: A holds b2 split into two BF16
    // parts (k slots 0 and 1 of lane half 0), B is 1 in those slots for a
    // valid pair, so an invalid pair gets pre2 = 0 and adds nothing to H.
    bf16x8 b2a[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int e = 0; e < 8; ++e) b2a[mt][e] = bf(0.0f);
        if (g == 0) {
            const float b = s_b2[32 * mt + p];
            const __bf16 hi = bf(b);
            b2a[mt][0] = hi;
            b2a[mt][1] = bf(b - (float)hi);
        }
    }
    bf16x8 one_valid, zero8;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        one_valid[e] = bf((g == 0 && e < 2) ? 1.0f : 0.0f);
        zero8[e] = bf(0.0f);
    }

    const bf16x8 ident[2] = {unit_identity(0, g, p), unit_identity(1, g, p)};

    const uint32_t nwaves = gridDim.x * kWaves;
    for (uint32_t grp = blockIdx.x * kWaves + wv; grp < a.ngroups; grp += nwaves) {
        // keeps the LDS-resident W3 in LDS (hoisting it out of the loop would
        // take 96 registers per lane)
        asm volatile("" ::: "memory");
        const Group G = open_group(grp, p, st);
        f32x16 hacc[2] = {zero16(), zero16()};
        // neighbour indices two tiles ahead, the neighbour's data one tile ahead
        const PairStream ps = pair_stream(G, g, st);
        const PairConst kc = pair_const(G, g, nm);
        PairRaw rn = ps.load(ps.index(0));
        uint32_t jn = ps.index(1);
        for (int t = 0; t < G.m; ++t) {
            const bool valid = t < G.cT;
            const PairRaw rc = rn;
            rn = ps.load(jn);
            jn = ps.index(t + 2);
            const bf16x8 E = edge_features(g, valid, rc, kc, nm);
            bf16x8 h1t[4];
            layer1_t(w1f, E, h1t);
            const bf16x8 bias_b = valid ? one_valid : zero8;
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                f32x16 c = mfma32(b2a[mt], bias_b, zero16());  // pre2 = b2 + W2 h1
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) c = mfma32(w2f[mt][kb], h1t[kb], c);
                // H += bf16(relu(pre2)) on the matrix cores (FP32 accumulation)
                bf16x8 m0, m1;
                pack_relu(c, m0, m1);
                hacc[mt] = mfma32(ident[0], m0, hacc[mt]);
                hacc[mt] = mfma32(ident[1], m1, hacc[mt]);
            }
        }

        // node layer: out = W3 [agg H ; q] + b3
        float o[kOut];
#pragma unroll
        for (int oo = 0; oo < kOut; ++oo) {
            float acc = 0.0f;
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    acc += s_w3[oo * kW3Cols + 32 * mt + mfma32_row(r, lane)] * hacc[mt][r];
            o[oo] = acc;
        }
#pragma unroll
        for (int oo = 0; oo < kOut; ++oo) o[oo] += __shfl_xor(o[oo], 32);
        if (G.iT >= st.n) continue;
        const uint32_t iT = G.iT;
        if constexpr (kTrain) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    *reinterpret_cast<float4*>(a.h + (size_t)iT * kHid + 32 * mt + 8 * q + 4 * g) =
                        make_float4(hacc[mt][4 * q], hacc[mt][4 * q + 1], hacc[mt][4 * q + 2],
                                    hacc[mt][4 * q + 3]);
        }
        if (g != 0) continue;
        const uint32_t nnb = st.nnb[iT];
        const float2 rp = st.rhop[iT];
        const float q0 = (rp.x * nm.inv_rho0 - 1.0f) * nm.rho_scale;
        const float q1 = rp.y * nm.inv_p;
        const float q2 = (float)nnb * nm.inv_nnb;
        float y[kOut];
#pragma unroll
        for (int oo = 0; oo < kOut; ++oo) {
            const float* w3 = s_w3 + oo * kW3Cols;
            y[oo] = o[oo] * nm.agg_scale + w3[kHid] * q0 + w3[kHid + 1] * q1 + w3[kHid + 2] * q2 +
                    s_w3[kOut * kW3Cols + oo];
        }
        if constexpr (kTrain) {
            a.out[iT] = make_float4(y[0], y[1], y[2], (float)nnb);
        } else {
            const float ax = y[0] * nm.a_scale, ay = y[1] * nm.a_scale, az = y[2] * nm.a_scale;
            if (a.accel_out) a.accel_out[iT] = make_float4(ax, ay, az, 0.0f);
            if (a.vel_out) {
                uint32_t sc = 0;
                if constexpr (kBatch) sc = st.sid[iT];
                const float dt = a.dt_dev[sc];
                const float4 v = G.vT;
                a.vel_out[iT] = make_float4(v.x + ax * dt, v.y + ay * dt, v.z + az * dt, v.w);
            }
        }
    }
}

// Backward pass of the pair network: recomputes the forward per tile and
// accumulates dL/dW1, dL/dW2, dL/db2 over every pair the wave serves; each
// wave writes its sums as one row of part_pair.
template <bool kBatch>
__global__ __launch_bounds__(kBlk) void k_mlp_bwd(const MlpArgs a) {
    __shared__ bf16x8 s_w2tf[8 * 64];
    __shared__ float s_b2[kHid];
    __shared__ float s_dh[kWaves * kGroup * kDhStride];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int g = lane >> 5, p = lane & 31;
    for (int k = threadIdx.x; k < 8 * 64; k += kBlk) s_w2tf[k] = a.w2tf[k];
    for (int k = threadIdx.x; k < kHid; k += kBlk) s_b2[k] = a.params[kOffB2 + k];
    __syncthreads();
    float* D = s_dh + wv * kGroup * kDhStride;
    const sph_mlp_norm nm = a.nm;
    const MlpState st = a.st;

    bf16x8 w1f[2], w2f[2][4], ifrag;
#pragma unroll
    for (int t = 0; t < 2; ++t) w1f[t] = a.w1f[t * 64 + lane];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) w2f[mt][kb] = a.w2f[(mt * 4 + kb) * 64 + lane];
#pragma unroll
    for (int e = 0; e < 8; ++e) ifrag[e] = bf((8 * g + e) == p ? 1.0f : 0.0f);
    f32x16 dw2[2][2], dw1[2];
    float db2[2] = {0.0f, 0.0f};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        dw1[nt] = zero16();
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) dw2[mt][nt] = zero16();
    }
    const float b2n[2] = {s_b2[p], s_b2[32 + p]};

    const uint32_t nwaves = gridDim.x * kWaves;
    for (uint32_t grp = blockIdx.x * kWaves + wv; grp < a.ngroups; grp += nwaves) {
        asm volatile("" ::: "memory");
        const Group G = open_group(grp, p, st);
        // stage dL/dH of the group's 32 targets (rows of kDhStride floats)
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int f = k * 64 + lane;  // float4 index within the 32 x 64 block
            const int tt = f >> 4, u4 = (f & 15) * 4;
            const uint32_t it = grp * kGroup + tt;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (it < st.n) v = *reinterpret_cast<const float4*>(a.dh + (size_t)it * kHid + u4);
            *reinterpret_cast<float4*>(D + tt * kDhStride + u4) = v;
        }
        wave_lds_sync();
        if (G.m == 0) continue;
        // dL/dH of this lane's target in the transposed layout (rows = units)
        float dht[2][16];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 v = *reinterpret_cast<const float4*>(D + p * kDhStride + 32 * mt + 8 * q + 4 * g);
                dht[mt][4 * q + 0] = v.x;
                dht[mt][4 * q + 1] = v.y;
                dht[mt][4 * q + 2] = v.z;
                dht[mt][4 * q + 3] = v.w;
            }

        // neighbour indices two tiles ahead, the neighbour's data one tile ahead
        const PairStream ps = pair_stream(G, g, st);
        const PairConst kc = pair_const(G, g, nm);
        PairRaw rn = ps.load(ps.index(0));
        uint32_t jn = ps.index(1);
        for (int t = 0; t < G.m; ++t) {
            const bool valid = t < G.cT;
            const PairRaw rc = rn;
            rn = ps.load(jn);
            jn = ps.index(t + 2);
            const bf16x8 E = edge_features(g, valid, rc, kc, nm);
            const uint32_t vm = (uint32_t)__ballot(valid);  // bit r: pair row r valid

            bf16x8 h1t[4];
            layer1_t(w1f, E, h1t);
            // layer 1 with lane = hidden unit, rows = pairs
            bf16x8 h1n[2][2];
            uint32_t m1 = 0;  // bit 16 t1 + r: pre-activation > 0
#pragma unroll
            for (int t1 = 0; t1 < 2; ++t1) {
                const f32x16 c = mfma32(E, w1f[t1], zero16());
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    h1n[t1][r >> 3][r & 7] = bf(relu(c[r]));
                    m1 |= (c[r] > 0.0f ? 1u : 0u) << (16 * t1 + r);
                }
            }
            // dL/dpre2 with lane = pair (A operand over output units)
            bf16x8 dp2t[4];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                f32x16 c = zero16();
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) c = mfma32(w2f[mt][kb], h1t[kb], c);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 b = *reinterpret_cast<const float4*>(s_b2 + 32 * mt + 8 * q + 4 * g);
                    const float bb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const int r = 4 * q + r4;
                        const float v = (valid && c[r] + bb[r4] > 0.0f) ? dht[mt][r] : 0.0f;
                        dp2t[2 * mt + (r >> 3)][r & 7] = bf(v);
                    }
                }
            }
            // dL/dpre2 with lane = output unit, rows = pairs
            bf16x8 dp2n[2][2];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                f32x16 c = zero16();
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) c = mfma32(h1t[kb], w2f[nt][kb], c);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int row = mfma32_row(r, lane);
                    const bool rv = ((vm >> row) & 1u) != 0u;
                    const float v = (rv && c[r] + b2n[nt] > 0.0f) ? D[row * kDhStride + 32 * nt + p] : 0.0f;
                    db2[nt] += v;
                    dp2n[nt][r >> 3][r & 7] = bf(v);
                }
            }
            // dW2 (out x in) += dpre2^T h1, contraction over pairs
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                    for (int k2 = 0; k2 < 2; ++k2) dw2[mt][nt] = mfma32(dp2n[mt][k2], h1n[nt][k2], dw2[mt][nt]);
            // dh1 = dpre2 W2 with lane = hidden unit; dpre1 = dh1 masked
            bf16x8 dp1n[2][2];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                f32x16 c = zero16();
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) c = mfma32(dp2t[kb], s_w2tf[(nt * 4 + kb) * 64 + lane], c);
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    dp1n[nt][r >> 3][r & 7] = bf(((m1 >> (16 * nt + r)) & 1u) ? c[r] : 0.0f);
            }
            // features with lane = feature, rows = pairs (E times the identity)
            bf16x8 enb[2];
            {
                const f32x16 c = mfma32(E, ifrag, zero16());
#pragma unroll
                for (int r = 0; r < 16; ++r) enb[r >> 3][r & 7] = bf(c[r]);
            }
            // dW1 (units x features) += dpre1^T E
#pragma unroll
            for (int t1 = 0; t1 < 2; ++t1)
#pragma unroll
                for (int k2 = 0; k2 < 2; ++k2) dw1[t1] = mfma32(dp1n[t1][k2], enb[k2], dw1[t1]);
        }
    }

    float* P = a.part_pair + (size_t)(blockIdx.x * kWaves + wv) * kPairParams;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                P[kOffW2 + (32 * mt + mfma32_row(r, lane)) * kHid + 32 * nt + p] = dw2[mt][nt][r];
    if (p < kFeat) {
#pragma unroll
        for (int t1 = 0; t1 < 2; ++t1)
#pragma unroll
            for (int r = 0; r < 16; ++r) P[kOffW1 + (32 * t1 + mfma32_row(r, lane)) * kFeat + p] = dw1[t1][r];
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const float s = db2[nt] + __shfl_xor(db2[nt], 32);
        if (g == 0) P[kOffB2 + 32 * nt + p] = s;
    }
}

// Node layer backward, one thread per slot: loss, dL/dout, dL/dH and the
// block's partial sums of dL/dW3, dL/db3 (fixed order).
template <bool kBatch>
__global__ __launch_bounds__(kNodeBlk) void k_mlp_node_grad(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float2* __restrict__ rhop, const float* __restrict__ h, const float4* __restrict__ out,
    const uint32_t* __restrict__ pid, const float4* __restrict__ tgt, float inv_window,
    sph_mlp_norm nm, const float* __restrict__ params, float* __restrict__ dh,
    float* __restrict__ part_node, double* __restrict__ loss_part) {
    __shared__ float s_w3[kNodeParams];
    __shared__ float s_d[kNodeBlk][kOut];
    __shared__ float s_q[kNodeBlk][kNode];
    __shared__ double s_l[kNodeBlk];
    const int t = threadIdx.x;
    const uint32_t i = blockIdx.x * kNodeBlk + t;
    for (int k = t; k < kNodeParams; k += kNodeBlk) s_w3[k] = params[kOffW3 + k];
    __syncthreads();
    float d[kOut] = {0.0f, 0.0f, 0.0f}, q[kNode] = {0.0f, 0.0f, 0.0f};
    double l = 0.0;
    if (i < n) {
        uint32_t scene;
        const sph_params& sp = scene_params<kBatch>(one, tab, sid, i, scene);
        const float4 o = out[i];
        const float4 y4 = tgt[pid[i]];
        const float ys = inv_window / (sp.time_step * nm.a_scale);
        const float diff[kOut] = {o.x - y4.x * ys, o.y - y4.y * ys, o.z - y4.z * ys};
        const float two_n = 2.0f / (float)n;
#pragma unroll
        for (int k = 0; k < kOut; ++k) {
            l += (double)diff[k] * (double)diff[k];
            d[k] = diff[k] * two_n;
        }
        const float2 rp = rhop[i];
        q[0] = (rp.x * nm.inv_rho0 - 1.0f) * nm.rho_scale;
        q[1] = rp.y * nm.inv_p;
        q[2] = o.w * nm.inv_nnb;
#pragma unroll
        for (int u4 = 0; u4 < kHid; u4 += 4) {
            float v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int u = u4 + c;
                v[c] = (d[0] * s_w3[u] + d[1] * s_w3[kW3Cols + u] + d[2] * s_w3[2 * kW3Cols + u]) *
                       nm.agg_scale;
            }
            *reinterpret_cast<float4*>(dh + (size_t)i * kHid + u4) = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
#pragma unroll
    for (int k = 0; k < kOut; ++k) s_d[t][k] = d[k];
#pragma unroll
    for (int k = 0; k < kNode; ++k) s_q[t][k] = q[k];
    s_l[t] = l;
    __syncthreads();
    for (int st = kNodeBlk / 2; st > 0; st >>= 1) {
        if (t < st) s_l[t] += s_l[t + st];
        __syncthreads();
    }
    if (t == 0) loss_part[blockIdx.x] = s_l[0];
    const uint32_t first = blockIdx.x * kNodeBlk;
    const int cnt = (int)min((uint32_t)kNodeBlk, n - first);
    for (int k = t; k < kNodeParams; k += kNodeBlk) {
        float s = 0.0f;
        if (k < kOut * kW3Cols) {
            const int o = k / kW3Cols, c = k % kW3Cols;
            if (c < kHid) {
                for (int r = 0; r < cnt; ++r) s += s_d[r][o] * (h[(size_t)(first + r) * kHid + c] * nm.agg_scale);
            } else {
                for (int r = 0; r < cnt; ++r) s += s_d[r][o] * s_q[r][c - kHid];
            }
        } else {
            const int o = k - kOut * kW3Cols;
            for (int r = 0; r < cnt; ++r) s += s_d[r][o];
        }
        part_node[(size_t)blockIdx.x * kNodeParams + k] = s;
    }
}

// Gradient reduction in two deterministic stages: block (x, y) sums rows
// y, y + kRedChunks, ... of parameter columns 256 x .. 256 x + 255, then one
// thread per parameter adds the chunk sums in order.
__global__ __launch_bounds__(256) void k_mlp_reduce_chunks(uint32_t rows_pair, uint32_t rows_node,
                                                           const float* __restrict__ part_pair,
                                                           const float* __restrict__ part_node,
                                                           const double* __restrict__ loss_part,
                                                           double* __restrict__ chunks) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = blockIdx.y;
    double s = 0.0;
    if (k < kPairParams) {
        for (uint32_t r = y; r < rows_pair; r += kRedChunks) s += (double)part_pair[(size_t)r * kPairParams + k];
    } else if (k < kNumParams) {
        const int kk = k - kPairParams;
        for (uint32_t r = y; r < rows_node; r += kRedChunks) s += (double)part_node[(size_t)r * kNodeParams + kk];
    } else if (k == kNumParams) {
        for (uint32_t r = y; r < rows_node; r += kRedChunks) s += loss_part[r];
    }
    if (k <= kNumParams) chunks[(size_t)y * (kNumParams + 1) + k] = s;
}

__global__ __launch_bounds__(256) void k_mlp_reduce_final(uint32_t n, const double* __restrict__ chunks,
                                                          float* __restrict__ grad, double* __restrict__ loss) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k > kNumParams) return;
    double s = 0.0;
    for (uint32_t y = 0; y < kRedChunks; ++y) s += chunks[(size_t)y * (kNumParams + 1) + k];
    if (k < kNumParams)
        grad[k] = (float)s;
    else
        loss[0] = s / (double)n;
}

__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ prm, float* __restrict__ m,
                                                  float* __restrict__ v, const float* __restrict__ g,
                                                  float lr, float b1, float b2, float eps, float bc1,
                                                  float bc2) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= kNumParams) return;
    const float gg = g[k];
    const float mm = b1 * m[k] + (1.0f - b1) * gg;
    const float vv = b2 * v[k] + (1.0f - b2) * (gg * gg);
    m[k] = mm;
    v[k] = vv;
    prm[k] = prm[k] - lr * (mm * bc1) / (sqrtf(vv * bc2) + eps);
}

constexpr int kPackFrags = 2 + 8 + 8;

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ prm, bf16x8* __restrict__ w1f,
                                                  bf16x8* __restrict__ w2f, bf16x8* __restrict__ w2tf) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= kPackFrags * 64) return;
    const int frag = k >> 6, lane = k & 63, g = lane >> 5, p = lane & 31;
    bf16x8 f;
    if (frag < 2) {
#pragma unroll
        for (int e = 0; e < 8; ++e) f[e] = bf(prm[kOffW1 + (32 * frag + p) * kFeat + 8 * g + e]);
        w1f[frag * 64 + lane] = f;
    } else if (frag < 10) {
        const int idx = frag - 2, mt = idx >> 2, kb = idx & 3;
#pragma unroll
        for (int e = 0; e < 8; ++e) f[e] = bf(prm[kOffW2 + (32 * mt + p) * kHid + hid_perm(kb, g, e)]);
        w2f[idx * 64 + lane] = f;
    } else {
        const int idx = frag - 10, nt = idx >> 2, kb = idx & 3;
#pragma unroll
        for (int e = 0; e < 8; ++e) f[e] = bf(prm[kOffW2 + hid_perm(kb, g, e) * kHid + 32 * nt + p]);
        w2tf[idx * 64 + lane] = f;
    }
}

__global__ __launch_bounds__(256) void k_accumulate_original(uint32_t n, const uint32_t* __restrict__ pid,
                                                             const float4* __restrict__ dvel,
                                                             float4* __restrict__ acc, int first) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t o = pid[k];
    const float4 d = dvel[k];
    if (first) {
        acc[o] = make_float4(d.x, d.y, d.z, 0.0f);
    } else {
        const float4 s = acc[o];
        acc[o] = make_float4(s.x + d.x, s.y + d.y, s.z + d.z, 0.0f);
    }
}

constexpr int kDevBlocks = 256;

__global__ __launch_bounds__(256) void k_density_dev_partials(uint32_t n, const float2* __restrict__ rhop,
                                                              float inv_rho0, double* __restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256)
        s += fabs((double)rhop[k].x * (double)inv_rho0 - 1.0);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// With counters: counters[first .. first + ncopy) are copied to cnt_out.
__global__ __launch_bounds__(256) void k_density_dev_final(uint32_t n, const double* __restrict__ part,
                                                           double* __restrict__ out,
                                                           const uint32_t* __restrict__ counters,
                                                           uint32_t first, uint32_t ncopy,
                                                           uint32_t* __restrict__ cnt_out) {
    __shared__ double sh[256];
    sh[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] / (double)n;
    if (counters && threadIdx.x < ncopy) cnt_out[threadIdx.x] = counters[first + threadIdx.x];
}

__global__ void k_scale_dt(uint32_t n, const float* __restrict__ src, float f, float* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k] * f;
}

// Row-major bf16 A (M x K) and B (K x N) in, row-major f32 C (M x N) out,
// through exactly the operand layouts documented in sph_mfma.hpp.
__global__ __launch_bounds__(64) void k_mfma_selftest32(const __bf16* __restrict__ a,
                                                       const __bf16* __restrict__ b,
                                                       float* __restrict__ c) {
    const int l = threadIdx.x;
    bf16x8 av, bv;
    const int row = l % 32, k0 = 8 * (l / 32);
    for (int e = 0; e < 8; ++e) {
        av[e] = a[row * 16 + k0 + e];
        bv[e] = b[(k0 + e) * 32 + row];  // column `row` of B
    }
    f32x16 acc;
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    acc = mfma32(av, bv, acc);
    for (int i = 0; i < 16; ++i) c[mfma32_row(i, l) * 32 + (l % 32)] = acc[i];
}

__global__ __launch_bounds__(64) void k_mfma_selftest16(const __bf16* __restrict__ a,
                                                       const __bf16* __restrict__ b,
                                                       float* __restrict__ c) {
    const int l = threadIdx.x;
    bf16x8 av, bv;
    const int row = l % 16, k0 = 8 * (l / 16);
    for (int e = 0; e < 8; ++e) {
        av[e] = a[row * 32 + k0 + e];
        bv[e] = b[(k0 + e) * 16 + row];
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma16(av, bv, acc);
    for (int i = 0; i < 4; ++i) c[(4 * (l / 16) + i) * 16 + (l % 16)] = acc[i];
}

inline uint32_t groups_of(uint32_t n) { return (n + kGroup - 1) / kGroup; }

MlpArgs make_args(const LaunchCfg& c, const sph_mlp_norm& nm, const float* params,
                  const MlpPacked& pk, const MlpState& st) {
    MlpArgs a{};
    a.one = c.one;
    a.tab = c.tab;
    a.st = st;
    a.nm = nm;
    a.params = params;
    a.w1f = static_cast<const bf16x8*>(pk.w1f);
    a.w2f = static_cast<const bf16x8*>(pk.w2f);
    a.w2tf = static_cast<const bf16x8*>(pk.w2tf);
    a.ngroups = groups_of(st.n);
    return a;
}

template <bool kTrain>
hipError_t launch_fwd(bool batch, const MlpArgs& a, hipStream_t s) {
    const dim3 grid((a.ngroups + kWaves - 1) / kWaves), block(kBlk);
    if (batch)
        hipLaunchKernelGGL((k_mlp_fwd<true, kTrain>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((k_mlp_fwd<false, kTrain>), grid, block, 0, s, a);
    return hipGetLastError();
}

}  // namespace

size_t mlp_w1f_bytes() { return 2 * 64 * sizeof(bf16x8); }
size_t mlp_w2f_bytes() { return 8 * 64 * sizeof(bf16x8); }

uint32_t mlp_pair_grad_waves(uint32_t n) {
    const uint32_t blocks = std::min<uint32_t>((groups_of(n) + kWaves - 1) / kWaves, kMaxGradBlocks);
    return blocks * kWaves;
}

uint32_t mlp_node_blocks(uint32_t n) { return (n + kNodeBlk - 1) / kNodeBlk; }

hipError_t mlp_pack(const float* params, const MlpPacked& pk, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((kPackFrags * 64 + 255) / 256), dim3(256), 0, s, params,
                       static_cast<bf16x8*>(pk.w1f), static_cast<bf16x8*>(pk.w2f),
                       static_cast<bf16x8*>(pk.w2tf));
    return hipGetLastError();
}

hipError_t mlp_infer(const LaunchCfg& c, const sph_mlp_norm& nm, const float* params,
                     const MlpPacked& pk, const MlpState& st, float4* vel_out, float4* accel_out,
                     const float* dt_dev, hipStream_t s) {
    if (st.n == 0) return hipSuccess;
    MlpArgs a = make_args(c, nm, params, pk, st);
    a.vel_out = vel_out;
    a.accel_out = accel_out;
    a.dt_dev = dt_dev;
    return launch_fwd<false>(c.tab != nullptr, a, s);
}

hipError_t mlp_gradient(const LaunchCfg& c, const sph_mlp_norm& nm, const float* params,
                        const MlpPacked& pk, const MlpState& st, const uint32_t* pid,
                        const float4* target_orig, uint32_t window, const MlpTrainBuffers& tb,
                        hipStream_t s) {
    if (st.n == 0 || window == 0) return hipErrorInvalidValue;
    const bool batch = c.tab != nullptr;
    MlpArgs a = make_args(c, nm, params, pk, st);
    a.h = tb.h;
    a.out = tb.out;
    hipError_t e = launch_fwd<true>(batch, a, s);
    if (e != hipSuccess) return e;
    const uint32_t nb = mlp_node_blocks(st.n);
    const float inv_window = 1.0f / (float)window;
    if (batch)
        hipLaunchKernelGGL((k_mlp_node_grad<true>), dim3(nb), dim3(kNodeBlk), 0, s, c.one, c.tab, st.sid,
                           st.n, st.rhop, tb.h, tb.out, pid, target_orig, inv_window, nm, params, tb.dh,
                           tb.part_node, tb.loss_part);
    else
        hipLaunchKernelGGL((k_mlp_node_grad<false>), dim3(nb), dim3(kNodeBlk), 0, s, c.one, c.tab, st.sid,
                           st.n, st.rhop, tb.h, tb.out, pid, target_orig, inv_window, nm, params, tb.dh,
                           tb.part_node, tb.loss_part);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t waves = mlp_pair_grad_waves(st.n);
    a.dh = tb.dh;
    a.part_pair = tb.part_pair;
    if (batch)
        hipLaunchKernelGGL((k_mlp_bwd<true>), dim3(waves / kWaves), dim3(kBlk), 0, s, a);
    else
        hipLaunchKernelGGL((k_mlp_bwd<false>), dim3(waves / kWaves), dim3(kBlk), 0, s, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_reduce_chunks, dim3((kNumParams + 1 + 255) / 256, kRedChunks), dim3(256), 0, s,
                       waves, nb, tb.part_pair, tb.part_node, tb.loss_part, tb.chunks);
    hipLaunchKernelGGL(k_mlp_reduce_final, dim3((kNumParams + 1 + 255) / 256), dim3(256), 0, s, st.n, tb.chunks,
                       tb.grad, tb.loss);
    return hipGetLastError();
}

hipError_t mlp_adam(float* params, float* m, float* v, const float* grad, float lr, float b1,
                    float b2, float eps, float bc1, float bc2, const MlpPacked& pk, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, s, params, m, v, grad,
                       lr, b1, b2, eps, bc1, bc2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return mlp_pack(params, pk, s);
}

hipError_t mlp_accumulate_original(uint32_t n, const uint32_t* pid, const float4* dvel, float4* acc,
                                   bool first, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_accumulate_original, dim3((n + 255) / 256), dim3(256), 0, s, n, pid, dvel, acc,
                       first ? 1 : 0);
    return hipGetLastError();
}

hipError_t mlp_density_deviation(uint32_t n, const float2* rhop, float inv_rho0, double* partials,
                                 double* out, hipStream_t s, const uint32_t* counters, uint32_t first,
                                 uint32_t ncopy, uint32_t* cnt_out) {
    hipLaunchKernelGGL(k_density_dev_partials, dim3(kDevBlocks), dim3(256), 0, s, n, rhop, inv_rho0,
                       partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_density_dev_final, dim3(1), dim3(256), 0, s, n, partials, out, counters, first, ncopy,
                       cnt_out);
    return hipGetLastError();
}

hipError_t mlp_scale_dt(uint32_t nscenes, const float* src, float factor, float* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_dt, dim3((nscenes + 255) / 256), dim3(256), 0, s, nscenes, src, factor, dst);
    return hipGetLastError();
}

}  // namespace sph

extern "C" {

// One matrix-core GEMM through the documented operand layouts.  variant 0:
// A 32x16, B 16x32, C 32x32; variant 1: A 16x32, B 32x16, C 16x16.  A and B
// are row-major bf16 bit patterns, C row-major float.  Synchronous.
int sph_mfma_selftest(int variant, const uint16_t* a, const uint16_t* b, float* c) {
    if (!a || !b || !c || (variant != 0 && variant != 1)) return SPH_ERR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPH_ERR_NO_DEVICE;
    const size_t na = 512, nb = 512, nc = variant == 0 ? 1024 : 256;
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    hipError_t e = sph::device_malloc(&da, na * 2);
    if (e == hipSuccess) e = sph::device_malloc(&db, nb * 2);
    if (e == hipSuccess) e = sph::device_malloc(&dc, nc * 4);
    if (e == hipSuccess) e = hipMemcpy(da, a, na * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, nb * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (variant == 0)
            hipLaunchKernelGGL(sph::k_mfma_selftest32, dim3(1), dim3(64), 0, 0, (const __bf16*)da,
                               (const __bf16*)db, (float*)dc);
        else
            hipLaunchKernelGGL(sph::k_mfma_selftest16, dim3(1), dim3(64), 0, 0, (const __bf16*)da,
                               (const __bf16*)db, (float*)dc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(c, dc, nc * 4, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    return e == hipSuccess ? SPH_OK : SPH_ERR_DEVICE;
}

}  // extern "C"
