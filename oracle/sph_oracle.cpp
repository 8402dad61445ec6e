// sph_oracle.cpp - CPU oracle of the WCSPH step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the parity tests, the
// smoke check and the CPU-baseline leg of bench.py compare the HIP path with.
// The product library never links, loads or calls it.
//
// It restates, in plain C++, the arithmetic of the reference's CUDA step
// (the reference has no CPU path of its own):
//   integrate + boundary ....... source/CUDA/System.cu:41-159, :165-205
//   height map, rotor, collider,
//   accelerators ............... source/CUDA/System.cu:253-399, source/CUDA/Kernel_Cell.cui:78-115
//   cell hash ................... source/CUDA/Kernel_Cell.cui:5-34
//   stable key/value sort ....... source/CUDA/radixsort_kernel.cu:445-472 (LSD, stable)
//   cellStart + reorder ......... source/CUDA/Kernel_Cell.cui:40-67, source/CUDA/System.cu:689-713
//   density + EOS ............... source/CUDA/Kernel_Cell.cui:142-199
//   force + velocity update ..... source/CUDA/Kernel_Cell.cui:210-261, source/CUDA/System.cu:228-250,402
//   step order .................. source/SPH/SPH_Update.cpp:39-80
//   vector helpers .............. source/external/cutil_math.h:392-407 (dot left to right)
//
// Modes:
//   float, ORACLE_PREC_REFERENCE: same operation order as the reference
//     kernels (per-cell partial sums, (c*c)*c for pow(c,3)), no FMA
//     contraction (built with -ffp-contract=off).
//   float, ORACLE_PREC_KAHAN: one compensated accumulator per sum over the
//     same candidate order.
//   double: the FP64 twin.  Takes the same float inputs and float-valued
//     constants, computes every step in double, and also returns per-particle
//     error scales (the sums re-evaluated with absolute values) that the
//     tests use to normalise FP32-vs-FP64 differences.
//
// Deviations from the reference that are shared with the HIP path:
//   - neighbour cells outside the grid are skipped (the reference reads
//     aliased or out-of-range cells whose particles cannot pass r < h);
//   - a particle's own cell index is clamped into the grid and counted.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../include/pibiti_sph.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr float kBndEps = 0.00001f;  // EPS of boundary(), source/CUDA/System.cu:51

struct Counters {
    uint32_t cap_cells;
    uint32_t out_of_grid;
    uint32_t max_cell_count;
    uint32_t reserved;
};

// ---------------------------------------------------------------- integrate
template <typename T>
inline void bnd_wall(T diff, T nx, T ny, T nz, T stiff, T damp, T dt, T v[3]) {
    // addB()/addC(): accBnd = stiff*diff - damp*dot(norm, vel); vel += accBnd*norm*dt
    if (diff > (T)kBndEps) {
        T dot = nx * v[0] + ny * v[1] + nz * v[2];
        T acc = stiff * diff - damp * dot;
        v[0] = v[0] + (acc * nx) * dt;
        v[1] = v[1] + (acc * ny) * dt;
        v[2] = v[2] + (acc * nz) * dt;
    }
}

// sin, cos and atan as the nearest T of the double result (the CUDA float
// functions the reference calls are not reproduced; both sides of the parity
// tests use correctly rounded values instead).
template <typename T>
inline T sin_t(T a) { return (T)std::sin((double)a); }
template <typename T>
inline T cos_t(T a) { return (T)std::cos((double)a); }
template <typename T>
inline T atan_t(T a) { return (T)std::atan((double)a); }

// boundary() (source/CUDA/System.cu:41-159) for one particle: the soft walls
// of the boundary type and z effect in the reference's order, the wrap/cycle
// position shifts and the pump.  addB walls damp with bndDamp, addC walls
// (z walls of the box, cylinders, sphere) with bndDampC.
template <typename T>
inline void boundary(const sph_params& p, T x[3], T v[3]) {
    const T b = (T)p.dist_bnd_soft, stiff = (T)p.bnd_stiff, damp = (T)p.bnd_damp,
            damp2 = (T)p.bnd_damp_c, dt = (T)p.time_step;
    T wmin[3] = {(T)p.world_min[0], (T)p.world_min[1], (T)p.world_min[2]};
    const T wmax[3] = {(T)p.world_max[0], (T)p.world_max[1], (T)p.world_max[2]};
    const T o = 0, one = 1;
    const uint32_t type = p.bnd_type, eff = p.bnd_eff_z;
    const bool cyl_y = type == SPH_BND_CYL_Y, cyl_z = type == SPH_BND_CYL_Z;
    const bool cyl_yz = type == SPH_BND_CYL_YZ, sphere = type == SPH_BND_SPHERE;
    const bool wave = eff == SPH_BND_EFF_WAVE, none = eff == SPH_BND_EFF_NONE,
               cycle = eff == SPH_BND_EFF_CYCLE;
    if (wave) {  // :55-61
        const T sl = -(T)p.rot2_angle;
        bnd_wall<T>((b - (x[1] - wmin[1])) - (x[2] - wmin[2]) * sl, o, one - sl, sl, stiff, damp, dt, v);
        wmin[2] = wmin[2] + (T)p.rot_twist * (one + sin_t<T>((T)p.rot_angle));
    }
    if (!sphere) {  // :64-77
        if (!cyl_y) {
            if (none || wave) bnd_wall<T>((b - x[2]) + wmin[2], o, o, one, stiff, damp2, dt, v);
            if (!cycle) bnd_wall<T>((b + x[2]) - wmax[2], o, o, -one, stiff, damp2, dt, v);
        }
        if (!cyl_y && !cyl_z) {
            bnd_wall<T>((b - x[0]) + wmin[0], one, o, o, stiff, damp, dt, v);
            bnd_wall<T>((b + x[0]) - wmax[0], -one, o, o, stiff, damp, dt, v);
        }
        if (!cyl_z) {
            bnd_wall<T>((b - x[1]) + wmin[1], o, one, o, stiff, damp, dt, v);
            bnd_wall<T>((b + x[1]) - wmax[1], o, -one, o, stiff, damp, dt, v);
        }
    } else {  // :78-81
        const T len = std::sqrt((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]);
        const T diff = (b + len) + wmin[1];
        if (diff > (T)kBndEps) bnd_wall<T>(diff, -x[0] / len, -x[1] / len, -x[2] / len, stiff, damp2, dt, v);
    }
    if (cyl_y || cyl_yz) {  // :84-86
        const T len = std::sqrt(x[0] * x[0] + x[2] * x[2]);
        const T diff = (b + len) - wmax[0];
        if (diff > (T)kBndEps) bnd_wall<T>(diff, -x[0] / len, o, -x[2] / len, stiff, damp2, dt, v);
    }
    if (cyl_z || cyl_yz) {  // :89-91
        const T len = std::sqrt(x[0] * x[0] + x[1] * x[1]);
        const T diff = (b + len) + wmin[1];
        if (diff > (T)kBndEps) bnd_wall<T>(diff, -x[0] / len, -x[1] / len, o, stiff, damp2, dt, v);
    }
    if (!wave && !none) {  // :94-97
        const T dr = (T)1 * (T)p.particle_r;
        const T ev = (T)p.exit_vel;
        if (cycle && v[2] > ev && x[2] > (wmax[2] - b) - dr)
            x[2] = x[2] - (((wmax[2] - wmin[2]) - (T)2 * b) - dr);
        else if (v[2] < -ev && x[2] < (wmin[2] + b) + dr)
            x[2] = x[2] + (((wmax[2] - wmin[2]) - (T)2 * b) - dr);
    }
    if (type == SPH_BND_PUMP_Y) {  // :101-158
        const T rad = wmax[0], ang = (T)p.pump_ang_out, hc = (T)p.h_close;
        const T rin = rad * (T)p.pump_rad_in;
        T sv[6];
        for (int k = 0; k < 6; ++k) sv[k] = (T)p.shape_s[k];
        const T len = std::sqrt(x[0] * x[0] + x[1] * x[1]);
        T diff = (b + len) - rad;
        if (diff > (T)kBndEps) {
            const T a = atan_t<T>(x[0] / x[1]);
            bool wall;
            if (ang < (T)0.5f)
                wall = a < -ang || a > ang || x[1] < o;
            else
                wall = x[1] < o || (len < rad * sv[4] && a < ang);
            if (wall) bnd_wall<T>(diff, -x[0] / len, -x[1] / len, o, stiff, damp, dt, v);
        }
        T xs;
        if (ang < (T)0.5f) {
            xs = sin_t<T>(ang * sv[2]) * rad;
            const T zs = (cos_t<T>(ang * sv[2]) * rad) * sv[3];
            if (x[1] > zs) {
                bnd_wall<T>((b - x[0]) - xs, one, o, o, stiff, damp, dt, v);
                bnd_wall<T>((b + x[0]) - xs, -one, o, o, stiff, damp, dt, v);
            }
        } else {
            xs = (T)0.09f * sv[3];
            if (len >= rad * sv[5]) bnd_wall<T>((b - x[0]) + xs, one, o, o, stiff, damp, dt, v);
        }
        if (x[2] > hc - b * sv[0]) {
            diff = (b + len) - rin;
            if (diff > (T)kBndEps) bnd_wall<T>(diff, -x[0] / len, -x[1] / len, o, stiff, damp, dt, v);
        }
        if (x[2] < hc - b * sv[1]) bnd_wall<T>((b + x[2]) - hc, o, o, -one, stiff, damp, dt, v);
        diff = (x[1] - wmax[1]) + (T)p.exit_dist;
        if (diff > (T)kBndEps && v[1] > (T)p.exit_vel) {
            T rr, aa;
            const T zz = std::fabs(hc - wmin[2]);
            if (ang < (T)0.5f) {
                const T xx = xs * (T)2;
                rr = ((x[0] + xx / (T)2) / xx) * (T)0.7f;
                aa = ((x[2] - zz / (T)2) / zz) * (T)1.6f;
            } else {
                rr = ((wmax[0] - x[0]) / xs) * (T)0.45f;
                aa = ((x[2] - zz / (T)2) / zz) * (T)1.8f;
            }
            rr = rr * rin;
            aa = aa * ((T)2.0f * (T)3.141592654f);
            const T nx = cos_t<T>(aa) * rr, ny = sin_t<T>(aa) * rr;
            const T nz = (T)((double)(wmax[2] - b) - (double)std::fabs(v[1] - (T)p.exit_vel) * 0.01);
            x[0] = nx;
            x[1] = ny;
            x[2] = nz;
            const T vy = v[1];
            v[1] = v[2];
            v[2] = -vy;
        }
    }
}

// integrateD + boundary() for one particle (source/CUDA/System.cu:41-159,165-205).
template <typename T>
inline void integrate_one(const sph_params& p, T x[3], T v[3]) {
    const T dt = (T)p.time_step;
    const T wmin[3] = {(T)p.world_min[0], (T)p.world_min[1], (T)p.world_min[2]};
    const T wmax[3] = {(T)p.world_max[0], (T)p.world_max[1], (T)p.world_max[2]};
    boundary<T>(p, x, v);
    for (int c = 0; c < 3; ++c) v[c] = v[c] + (T)p.gravity[c] * dt;
    for (int c = 0; c < 3; ++c) v[c] = v[c] * (T)p.global_damping;
    for (int c = 0; c < 3; ++c) x[c] = x[c] + v[c] * dt;
    const T bh = (T)p.dist_bnd_hard;
    for (int c = 0; c < 3; ++c) {
        T hi = wmax[c] - bh, lo = wmin[c] + bh;
        if (x[c] > hi) x[c] = hi;
        if (x[c] < lo) x[c] = lo;
    }
}

// ------------------------------------------------------------------ hashing
// calcGridPos (source/CUDA/Kernel_Cell.cui:5-13) with the index clamped.
template <typename T>
inline void grid_pos(const sph_params& p, const T x[3], int g[3], uint32_t* oob) {
    bool out = false;
    for (int c = 0; c < 3; ++c) {
        T q = (x[c] - (T)p.world_min[c]) / (T)p.cell_size[c];
        T f = std::floor(q);
        T hi = (T)(p.grid_size[c] - 1);
        if (!(f >= (T)0 && f <= hi)) {
            out = true;
            f = std::isnan(f) ? (T)0 : std::min(std::max(f, (T)0), hi);
        }
        g[c] = (int)f;
    }
    if (out && oob) ++*oob;
}

// calcGridHash (source/CUDA/Kernel_Cell.cui:15-19).
inline uint32_t grid_hash(const sph_params& p, int gx, int gy, int gz) {
    return (uint32_t)gz * p.grid_size_yx + (uint32_t)gy * p.grid_size[0] + (uint32_t)gx;
}

// -------------------------------------------------------------- accumulators
template <typename T>
struct Plain {
    T s = 0;
    void add(T v) { s = s + v; }
};
template <typename T>
struct Kahan {
    T s = 0, c = 0;
    void add(T v) {
        T y = v - c;
        T t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

template <typename T>
struct Grid {
    std::vector<uint32_t> cell_start, cell_end;  // cellEnd valid for non-empty cells
};

// ---------------------------------------------------------------- density
// computeDensityD/compDensCell (source/CUDA/Kernel_Cell.cui:142-199) for the
// particle in sorted slot i.  Returns the raw kernel sum (before Poly6*m).
// scale (if given) receives sum_j [c^3 + 3 c^2 (h2 + r2)]: each term plus
// its first-order sensitivity to a relative perturbation of the operands of
// c = h2 - r2, the error scale of the sum in floating point.
template <typename T, bool kKahan>
T density_sum(const sph_params& p, uint32_t i, const std::vector<T>& spos,
              const std::vector<uint32_t>& cs, const std::vector<uint32_t>& ce, T h2,
              T* scale = nullptr) {
    T sc = 0;
    const T xi[3] = {spos[4 * i], spos[4 * i + 1], spos[4 * i + 2]};
    int g[3];
    grid_pos<T>(p, xi, g, nullptr);
    Plain<T> total;
    Kahan<T> ktotal;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int cx = g[0] + dx, cy = g[1] + dy, cz = g[2] + dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)p.grid_size[0] ||
                    cy >= (int)p.grid_size[1] || cz >= (int)p.grid_size[2])
                    continue;
                uint32_t c = grid_hash(p, cx, cy, cz);
                uint32_t s = cs[c];
                if (s == kEmpty) continue;
                uint32_t e = std::min(ce[c], s + p.max_par_in_cell);
                Plain<T> cell;
                for (uint32_t j = s; j < e; ++j) {
                    if (j == i) continue;
                    T rx = xi[0] - spos[4 * j], ry = xi[1] - spos[4 * j + 1],
                      rz = xi[2] - spos[4 * j + 2];
                    T r2 = rx * rx + ry * ry + rz * rz;
                    if (r2 < h2) {
                        T c3 = h2 - r2;
                        T w = (c3 * c3) * c3;
                        if (kKahan)
                            ktotal.add(w);
                        else
                            cell.add(w);
                        if (scale) sc += w + (T)3 * (c3 * c3) * (h2 + r2);
                    }
                }
                if (!kKahan) total.add(cell.s);
            }
    if (scale) *scale = sc;
    return kKahan ? ktotal.s : total.s;
}

// |p| evaluated with absolute values (the EOS terms before they cancel) and
// |dp/drho|: together the error scale of a pressure computed from a density.
template <typename T>
inline T eos_abs(const sph_params& p, T rho) {
    if (p.eos == SPH_EOS_TAIT) {
        T x = std::fabs(rho) / (T)p.rest_density;
        T x2 = x * x, x4 = x2 * x2, x7 = (x4 * x2) * x;
        return ((T)p.stiffness * (T)p.rest_density / (T)7) * (x7 + (T)1);
    }
    return (std::fabs(rho) + (T)p.rest_density) * (T)p.stiffness;
}
template <typename T>
inline T eos_slope(const sph_params& p, T rho) {
    if (p.eos == SPH_EOS_TAIT) {
        T x = std::fabs(rho) / (T)p.rest_density;
        T x2 = x * x;
        return (T)p.stiffness * (x2 * x2 * x2);
    }
    return (T)p.stiffness;
}

template <typename T>
inline T eos(const sph_params& p, T rho) {
    if (p.eos == SPH_EOS_TAIT) {
        // p = (k rho0 / 7) * ((rho/rho0)^7 - 1), powers as explicit products.
        T x = rho / (T)p.rest_density;
        T x2 = x * x;
        T x4 = x2 * x2;
        T x7 = (x4 * x2) * x;
        T bcoef = ((T)p.stiffness * (T)p.rest_density) / (T)7;
        return bcoef * (x7 - (T)1);
    }
    return (rho - (T)p.rest_density) * (T)p.stiffness;
}

// ------------------------------------------------------------------ force
// collideSpheresR (source/CUDA/Kernel_Cell.cui:101-115) added to out: the
// spring and damping force of a rotor or height-map sphere at offset rel
// (relative velocity u = -v); a zero vector outside its reach.
template <typename T>
inline void collide_r(const sph_params& p, const T rel[3], const T u[3], T rad, T out[3]) {
    const T dist = std::sqrt((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2]);
    T f[3] = {0, 0, 0};
    if (dist < rad) {
        const T inv = (T)1 / dist;  // float3 / float (cutil_math.h:358-362)
        const T spring = (T)p.coll_spring * (dist - rad);
        for (int k = 0; k < 3; ++k) f[k] = spring * (rel[k] * inv) + (T)p.coll_damping * u[k];
    }
    for (int k = 0; k < 3; ++k) out[k] = out[k] + f[k];
}

constexpr float kPiRef = 3.141592654f;  // PI, source/CUDA/Params.cuh:18

// What computeForceD adds to addVel after the pair sum
// (source/CUDA/System.cu:253-399), in its order: the height map; the rotor
// or, with no rotor, the sphere collider's DEM force (collideSpheres,
// source/CUDA/Kernel_Cell.cui:78-96; float3/float is a multiplication by the
// reciprocal, source/external/cutil_math.h:358-362); each accelerator whose
// shape holds the particle.
template <typename T>
inline void extra_terms(const sph_params& p, const T xi[3], const T vi[3], T out[3]) {
    const T dt = (T)p.time_step;
    const T pi = (T)kPiRef;
    T u[3];
    for (int k = 0; k < 3; ++k) u[k] = vi[k] * (T)-1;
    if (p.hmap_type > 0) {  // :254-306
        const bool xz = p.hmap_type == SPH_HMAP_XZ;
        const int ax = xz ? 0 : 1;  // the in-plane axis besides z
        const int lv = xz ? 1 : 0;  // the axis the sheet's height is along
        const T spc = (T)p.rot_spc, rr = (T)p.particle_r + (T)p.rot_r;
        // worldSizeD (Scene.cpp:99-103)
        const T bd = (T)p.dist_bnd_soft - (T)p.particle_r;
        T wsd[3];
        for (int c = 0; c < 3; ++c) wsd[c] = ((T)p.world_max[c] - bd) - ((T)p.world_min[c] + bd);
        const int ia = (int)((xi[ax] - (T)p.world_min[ax]) / spc);
        const int iz = (int)((xi[2] - (T)p.world_min[2]) / spc);
        const T af = (T)ia * spc + (T)p.world_min[ax];
        const T zf = (T)iz * spc + (T)p.world_min[2];
        for (int j = -2; j <= 2; ++j)
            for (int i = -2; i <= 2; ++i) {
                const T ah = af + (T)i * spc, zh = zf + (T)j * spc;
                for (int k = 0; k <= 1; ++k) {
                    // xh / worldSizeD.x * PI2 with PI2 = 2.f*PI unparenthesised
                    const T an = ((ah / wsd[ax]) * (T)2) * pi;
                    const T zn = ((zh / wsd[2]) * (T)2) * pi;
                    const T ss = sin_t<T>((T)p.shape_s[0] * an + ((T)p.shape_s[1] * pi) / (T)180) *
                                 sin_t<T>((T)p.shape_s[2] * zn + ((T)p.shape_s[3] * pi) / (T)180);
                    if ((T)p.shape_s[4] <= (T)-1 || ss > (T)p.shape_s[4]) {
                        const T lf = ((ss * (T)p.h_close) + (T)p.rot2_angle) + (T)p.world_min[lv];
                        const T lh = lf - (T)k * spc;
                        T rel[3];
                        rel[ax] = ah - xi[ax];
                        rel[lv] = lh - xi[lv];
                        rel[2] = zh - xi[2];
                        collide_r<T>(p, rel, u, rr, out);
                    }
                }
            }
    }
    if (p.rot_type > 0) {  // :309-372
        const int sx = p.rot_size[0], sy = p.rot_size[1], sz = p.rot_size[2], cb = p.rot_blades;
        const T sp = (T)p.rot_spc, ca = ((T)2 * pi) / (T)cb;
        const T x2 = (T)sx * (T)0.5f, y2 = (T)sy * (T)0.5f;
        const T rr = (T)p.particle_r + (T)p.rot_r;
        const T cp[3] = {(T)p.coll_pos[0], (T)p.coll_pos[1], (T)p.coll_pos[2]};
        if (p.rot_type == SPH_ROTOR_PUMP_Z || p.rot_type == SPH_ROTOR_PUMP_Y) {
            const int ax = p.rot_type == SPH_ROTOR_PUMP_Z ? 2 : 1;  // blade axis
            const int bx = p.rot_type == SPH_ROTOR_PUMP_Z ? 1 : 2;  // the other turning axis
            if (xi[ax] > cp[ax] - rr && xi[ax] < (cp[ax] + rr) + sp * (T)sy)
                for (int z = 1; z <= sz; ++z)
                    for (int x = 0; x <= sx; ++x)
                        for (int c = 0; c < cb; ++c) {
                            const T a = (-(T)p.rot_angle + (T)c * ca) + (T)z * (T)p.rot_twist;
                            const T cs = cos_t<T>(a) * sp, sn = -sin_t<T>(a) * sp;
                            for (int y = 0; y <= sy; ++y) {
                                T pc[3];
                                pc[0] = ((T)x - x2) * cs - (T)z * sn;
                                pc[bx] = ((T)x - x2) * sn + (T)z * cs;
                                pc[ax] = sp * (T)y;
                                T rel[3];
                                for (int k = 0; k < 3; ++k) rel[k] = (cp[k] + pc[k]) - xi[k];
                                collide_r<T>(p, rel, u, rr, out);
                            }
                        }
        } else {
            const T half = (sp * (T)sz) / (T)2;
            if (xi[2] > (cp[2] - rr) - half && xi[2] < (cp[2] + rr) + half) {
                T rot[3] = {cp[0], cp[1], cp[2]};
                T aa = (T)p.rot_angle, tw = (T)p.rot_twist;
                if (p.rot2_dist > 0.0f) {
                    if (xi[0] > (T)0) {
                        rot[0] = rot[0] + (T)p.rot2_dist * (T)0.5f;
                    } else {
                        rot[0] = rot[0] - (T)p.rot2_dist * (T)0.5f;
                        aa = (T)p.rot2_angle;
                        tw = tw * (T)p.rot2_twist;
                    }
                }
                for (int c = 0; c < cb; ++c) {
                    const T a = aa + (T)c * ca;
                    for (int h = 0; h <= sy; ++h) {
                        T dh = 0;
                        if (h == sy - 1)
                            dh = (T)0.5f;
                        else if (h == sy)
                            dh = (T)1.2f;
                        for (int x = 0; x <= sz; ++x) {
                            T d = dh;
                            if (x == sz && d == (T)0) d = (T)0.4f;
                            const T k = cos_t<T>(((T)x - x2) * (T)0.2f + pi * (T)0.6f);
                            const T at = tw * (d * (T)-0.05f - ((T)h - y2) * k);
                            const T ac = a + at, cs = cos_t<T>(ac) * sp, sn = -sin_t<T>(ac) * sp;
                            const T pc[3] = {(T)x * cs, (T)x * sn, (sp * ((-d + (T)h) - y2)) * k};
                            const T r = (T)p.rot_r * std::fabs((T)1 - d);
                            T rel[3];
                            for (int q = 0; q < 3; ++q) rel[q] = (rot[q] + pc[q]) - xi[q];
                            collide_r<T>(p, rel, u, (T)p.particle_r + r, out);
                        }
                    }
                }
            }
        }
    } else if (p.collider) {
        T rel[3], nrm[3], f[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) rel[k] = (T)p.coll_pos[k] - xi[k];
        const T dist = std::sqrt((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2]);
        const T rad = (T)p.particle_r + (T)p.coll_r;
        if (dist < rad) {
            const T inv = (T)1 / dist;
            for (int k = 0; k < 3; ++k) nrm[k] = rel[k] * inv;
            const T d = (u[0] * nrm[0] + u[1] * nrm[1]) + u[2] * nrm[2];
            const T spring = (T)p.coll_spring * (dist - rad);
            for (int k = 0; k < 3; ++k) {
                const T tan = u[k] - d * nrm[k];
                f[k] = (spring * nrm[k] + (T)p.coll_damping * u[k]) + (T)p.coll_shear * tan;
            }
        }
        for (int k = 0; k < 3; ++k) out[k] = out[k] + f[k];
    }
    for (int a = 0; a < SPH_NUM_ACCEL; ++a) {  // :379-399
        const sph_accel& ac = p.acc[a];
        if (ac.type == SPH_ACC_OFF) continue;
        T rel[3], size[3], acc[3];
        for (int k = 0; k < 3; ++k) {
            rel[k] = xi[k] - (T)ac.pos[k];
            size[k] = (T)ac.size[k];
            acc[k] = (T)ac.acc[k];
        }
        if (ac.type == SPH_ACC_BOX) {
            if (std::fabs(rel[0]) < size[0] && std::fabs(rel[1]) < size[1] && std::fabs(rel[2]) < size[2])
                for (int k = 0; k < 3; ++k) out[k] = out[k] + acc[k] * dt;
        } else if (ac.type == SPH_ACC_CYL_Y) {
            const T qx = rel[0] / size[0], qz = rel[2] / size[2];
            const T r = std::sqrt(qx * qx + qz * qz);
            if (std::fabs(rel[1]) < size[1] && r < (T)1)
                for (int k = 0; k < 3; ++k) out[k] = out[k] + acc[k] * dt;
        } else {
            const T r = std::sqrt(rel[0] * rel[0] + rel[2] * rel[2]);
            if (std::fabs(rel[1]) < size[1] && r < size[0]) {
                const T g = (T)1 - r / size[2];
                for (int k = 0; k < 3; ++k) out[k] = out[k] + (acc[k] * g) * dt;
            }
        }
    }
}

// computeForceD/compForceCell/compForcePair (source/CUDA/Kernel_Cell.cui:210-261,
// source/CUDA/System.cu:228-250) for sorted slot i.  Writes the velocity
// increment addVel = F*m*dt.  For double also returns the absolute-value
// evaluation of the same sum (error scale).
template <typename T, bool kKahan>
void force_sum(const sph_params& p, uint32_t i, const std::vector<T>& spos,
               const std::vector<T>& svel, const std::vector<T>& dens, const std::vector<T>& pres,
               const std::vector<T>& pscale, const std::vector<uint32_t>& cs,
               const std::vector<uint32_t>& ce, T out[3], T* scale) {
    const T xi[3] = {spos[4 * i], spos[4 * i + 1], spos[4 * i + 2]};
    const T vi[3] = {svel[4 * i], svel[4 * i + 1], svel[4 * i + 2]};
    const T di = dens[i], pi = pres[i];
    const bool dbl = sizeof(T) == 8;
    // error scale of p_i + p_j: the pressures' own error scales
    auto pabs = [&](uint32_t k) -> T { return pscale[k]; };
    const T pai = dbl ? pabs(i) : (T)0;
    int g[3];
    grid_pos<T>(p, xi, g, nullptr);
    const T h = (T)p.h, spiky = (T)p.spiky_kern, lap = (T)p.lap_kern, visc = (T)p.viscosity,
            min_dist = (T)p.min_dist, min_dens = (T)p.min_dens;
    Plain<T> tot[3];
    Kahan<T> ktot[3];
    T sc = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int cx = g[0] + dx, cy = g[1] + dy, cz = g[2] + dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)p.grid_size[0] ||
                    cy >= (int)p.grid_size[1] || cz >= (int)p.grid_size[2])
                    continue;
                uint32_t c = grid_hash(p, cx, cy, cz);
                uint32_t s = cs[c];
                if (s == kEmpty) continue;
                uint32_t e = std::min(ce[c], s + p.max_par_in_cell);
                Plain<T> cell[3];
                for (uint32_t j = s; j < e; ++j) {
                    if (j == i) continue;
                    T rel[3] = {xi[0] - spos[4 * j], xi[1] - spos[4 * j + 1],
                                xi[2] - spos[4 * j + 2]};
                    T rv[3] = {svel[4 * j] - vi[0], svel[4 * j + 1] - vi[1],
                               svel[4 * j + 2] - vi[2]};
                    T len = std::sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
                    T r = std::fmax(min_dist, len);
                    if (r < h) {
                        T d12 = std::fmin(min_dens, (T)1 / (di * dens[j]));
                        T cc = h - r;
                        T pterm = ((cc * spiky) * (pi + pres[j])) / r;
                        T vterm = lap * visc;
                        T cd = cc * d12;
                        for (int k = 0; k < 3; ++k) {
                            T f = (pterm * rel[k] + vterm * rv[k]) * cd;
                            if (kKahan)
                                ktot[k].add(f);
                            else
                                cell[k].add(f);
                        }
                        if (dbl) {
                            // |term| with absolute values, plus its first-order
                            // sensitivity to c = h - r (which cancels near r = h):
                            // term = (c A' + B) c d12 -> d/dc = (2 A + B) d12,
                            // and c carries an error proportional to h + r.
                            T pa = ((cc * spiky) * (pai + pabs(j))) / r;
                            T vn = std::sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
                            T A = pa * len, B = std::fabs(vterm) * vn;
                            sc += (A + B) * std::fabs(cd) + ((T)2 * A + B) * std::fabs(d12) * (h + r);
                        }
                    }
                }
                if (!kKahan)
                    for (int k = 0; k < 3; ++k) tot[k].add(cell[k].s);
            }
    const T mdt = (T)p.particle_mass * (T)p.time_step;
    for (int k = 0; k < 3; ++k) out[k] = (kKahan ? ktot[k].s : tot[k].s) * mdt;
    extra_terms<T>(p, xi, vi, out);
    if (scale) *scale = sc * std::fabs(mdt);
}

// ------------------------------------------------------------- whole step
template <typename T>
struct StepOut {
    // all in original index order
    T* pos4;
    T* vel4;
    T* vel_int4;    // velocity after integrate, before the force update
    T* dens;
    T* pres;
    T* dvel4;
    T* dvel_scale;  // double only
    T* dens_scale;  // double only
    T* pres_scale;  // double only
    uint32_t* sorted_hash;
    uint32_t* perm;
    uint32_t* cell_start;
    uint32_t* cell_end;
    Counters* cnt;
    T* colour4;     // particle colours (clr_type != NONE)
    T* dye;         // dye intensity, read and written
};

// ----------------------------------------------------------------- colours
// computeForceD's colouring and dye (source/CUDA/System.cu:405-546) for
// sorted slot k; dv is the step's whole velocity increment.  Returns false
// for CLR_None (no colour, no dye update).
template <typename T>
inline T dot4(const T* a, const T* b) {
    return ((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) + a[3] * b[3];  // cutil_math.h:540-543
}
template <typename T>
inline T phi_clamp(T v, T lo, T hi) {  // phiClamp (System.cu:221-223)
    return (std::fmin(v, hi) - std::fmin(v, lo)) / (hi - lo);
}

template <typename T>
bool colour_one(const sph_params& p, uint32_t k, uint32_t si, const std::vector<T>& spos,
                const std::vector<T>& svel, const std::vector<T>& dens, const T* dv,
                const std::vector<uint32_t>& cs, const std::vector<uint32_t>& ce, T* dye, T col[3]) {
    if (p.clr_type == SPH_CLR_NONE) return false;
    const T dt = (T)p.time_step, br = (T)p.brightness, ct = (T)p.contrast, rho0 = (T)p.rest_density;
    const T* x = &spos[4 * (size_t)k];
    const T* v = &svel[4 * (size_t)k];
    auto len3 = [](T a, T b, T c) { return std::sqrt((a * a + b * b) + c * c); };
    col[0] = (T)0.2f;
    col[1] = (T)0.5f;
    col[2] = (T)1;
    T intens = 0;
    const uint32_t t = p.clr_type;
    if (t == SPH_CLR_VEL_ACC || t == SPH_CLR_DENS_ACC) {
        const T f = ((T)0.02f * len3(dv[0], dv[1], dv[2])) / dt;
        const T kf = ct * f;
        T kb;
        if (t == SPH_CLR_VEL_ACC)
            kb = br + ct * ((T)2.5f * len3(v[0], v[1], v[2]));
        else
            kb = br + ct * (((T)4 * (dens[k] - rho0)) / rho0);
        const T add[3] = {(T)0.7f, t == SPH_CLR_VEL_ACC ? (T)0.35f : (T)0.7f, (T)0};
        for (int c = 0; c < 3; ++c) col[c] = col[c] * kb + add[c] * kf;
    } else if (t == SPH_CLR_VEL) {
        intens = br + ct * ((T)2.5f * len3(v[0], v[1], v[2]));
        // foam potential (:424-486): cells in x, y, z order
        int g[3];
        grid_pos<T>(p, x, g, nullptr);
        T vdiff = 0;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const int cx = g[0] + dx, cy = g[1] + dy, cz = g[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)p.grid_size[0] || cy >= (int)p.grid_size[1] ||
                        cz >= (int)p.grid_size[2])
                        continue;
                    const uint32_t c = grid_hash(p, cx, cy, cz);
                    if (cs[c] == kEmpty) continue;
                    const uint32_t e = std::min(ce[c], cs[c] + p.max_par_in_cell);
                    for (uint32_t j = cs[c]; j < e; ++j) {
                        if (j == k) continue;
                        const T* x2 = &spos[4 * (size_t)j];
                        const T* v2 = &svel[4 * (size_t)j];
                        T sv[4], sx[4], ddv[4], ddx[4];
                        for (int q = 0; q < 4; ++q) {
                            sv[q] = v[q] + v2[q];
                            sx[q] = x[q] + x2[q];
                            ddv[q] = v[q] - v2[q];
                            ddx[q] = x[q] - x2[q];
                        }
                        const T iv = (T)1 / std::sqrt(dot4(sv, sv)), ix = (T)1 / std::sqrt(dot4(sx, sx));
                        for (int q = 0; q < 4; ++q) {
                            sv[q] = sv[q] * iv;  // normalize (cutil_math.h:552-556)
                            sx[q] = sx[q] * ix;
                        }
                        const T r = std::sqrt(dot4(ddx, ddx));
                        const T w = r <= (T)3 ? ((T)1 - r) / (T)3 : (T)0;
                        vdiff = vdiff + (std::sqrt(dot4(ddv, ddv)) * ((T)1 - dot4(sv, sx))) * w;
                    }
                }
        const T trapped = phi_clamp<T>(vdiff, (T)5, (T)20);
        const T ke = (((T)p.particle_mass * (T)1000) * dot4(v, v)) * (T)0.5;
        const T nd = phi_clamp<T>(ke, (T)0.1f, (T)1) * ((T)3 * trapped + (T)3 * (T)0);
        col[0] = col[1] = col[2] = nd;
    } else if (t == SPH_CLR_VEL_RGB) {
        for (int c = 0; c < 3; ++c) col[c] = (T)0.5f + v[c] * (ct * (T)2.5f);
    } else {  // ACCEL, DENS
        const T kb = t == SPH_CLR_ACCEL ? br + ct * (((T)0.02f * len3(dv[0], dv[1], dv[2])) / dt)
                                        : br + ct * (((T)4 * (dens[k] - rho0)) / rho0);
        intens = kb;
        for (int c = 0; c < 3; ++c) col[c] = col[c] * kb;
    }
    if (p.hue == 1) {
        col[0] = intens;
        col[1] = 0;
    }
    if (p.dye_clear > 0) {
        dye[si] = 0;
    } else if (p.dye_type > 0) {
        T dc = dye[si];
        const T rel[3] = {x[0] - (T)p.dye_pos[0], x[1] - (T)p.dye_pos[1], x[2] - (T)p.dye_pos[2]};
        if (p.dye_type == SPH_DYE_BOX) {
            if (std::fabs(rel[0]) < (T)p.dye_size[0] && std::fabs(rel[1]) < (T)p.dye_size[1] &&
                std::fabs(rel[2]) < (T)p.dye_size[2])
                dc = 1;
        } else if (len3(rel[0], rel[1], rel[2]) < (T)p.dye_size[1]) {
            dc = 1;
        }
        dc = dc - dt * (T)p.dye_fade;
        if (dc < 0) dc = 0;
        if (dc >= 0) dye[si] = dc;
        if (p.hue == 0) {
            col[0] = col[0] + (T)0.9f * dc;
            col[1] = col[1] + (T)0.9f * dc;
            col[2] = col[2] + (T)1 * dc;
        } else {
            col[1] = dc;
        }
    }
    return true;
}

template <typename T>
void step_impl(const sph_params& p, int prec, uint32_t n, const T* pos_in, const T* vel_in,
               bool skip_integrate, const StepOut<T>& o) {
    const uint32_t ncell = p.num_cells;
    std::vector<T> pos(pos_in, pos_in + 4 * (size_t)n), vel(vel_in, vel_in + 4 * (size_t)n);
    Counters cnt{};

    // 1. integrate in original order (SPH_Update.cpp:39-47).
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (skip_integrate) continue;
        T x[3] = {pos[4 * i], pos[4 * i + 1], pos[4 * i + 2]};
        T v[3] = {vel[4 * i], vel[4 * i + 1], vel[4 * i + 2]};
        integrate_one<T>(p, x, v);
        for (int c = 0; c < 3; ++c) {
            pos[4 * i + c] = x[c];
            vel[4 * i + c] = v[c];
        }
    }
    if (o.vel_int4) std::copy(vel.begin(), vel.end(), o.vel_int4);
    // 2. hash in original order (SPH_Update.cpp:56).
    std::vector<uint32_t> hash(n);
    uint32_t oob = 0;
    for (uint32_t i = 0; i < n; ++i) {
        T x[3] = {pos[4 * i], pos[4 * i + 1], pos[4 * i + 2]};
        int g[3];
        grid_pos<T>(p, x, g, &oob);
        hash[i] = grid_hash(p, g[0], g[1], g[2]);
    }
    cnt.out_of_grid = oob;
    // 3. stable sort of (hash, index) by hash (SPH_Update.cpp:57-58).
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return hash[a] < hash[b]; });
    // 4. cellStart/cellEnd + reorder (Kernel_Cell.cui:40-67).
    std::vector<uint32_t> cs(ncell, kEmpty), ce(ncell, kEmpty);
    std::vector<T> spos(4 * (size_t)n), svel(4 * (size_t)n);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t hk = hash[perm[k]];
        if (k == 0 || hk != hash[perm[k - 1]]) cs[hk] = k;
        if (k + 1 == n || hk != hash[perm[k + 1]]) ce[hk] = k + 1;
        for (int c = 0; c < 4; ++c) {
            spos[4 * (size_t)k + c] = pos[4 * (size_t)perm[k] + c];
            svel[4 * (size_t)k + c] = vel[4 * (size_t)perm[k] + c];
        }
    }
    for (uint32_t c = 0; c < ncell; ++c)
        if (cs[c] != kEmpty) {
            uint32_t cnt_c = ce[c] - cs[c];
            if (cnt_c > p.max_par_in_cell) ++cnt.cap_cells;
            cnt.max_cell_count = std::max(cnt.max_cell_count, cnt_c);
        }
    // 5. density + EOS in sorted order (Kernel_Cell.cui:178-199).
    const bool dbl = sizeof(T) == 8;
    const T h2 = dbl ? (T)p.h * (T)p.h : (T)p.h2;
    std::vector<T> dens(n), pres(n), rsc(n, (T)0), psc(n, (T)0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        T sc = 0;
        T* scp = dbl ? &sc : nullptr;
        T s = prec == SPH_PREC_KAHAN ? density_sum<T, true>(p, (uint32_t)i, spos, cs, ce, h2, scp)
                                     : density_sum<T, false>(p, (uint32_t)i, spos, cs, ce, h2, scp);
        T rho = (s * (T)p.poly6_kern) * (T)p.particle_mass;
        dens[i] = rho;
        pres[i] = eos<T>(p, rho);
        if (dbl) {
            // error scales: density sum, and the EOS terms plus dp/drho times that
            rsc[i] = sc * (T)p.poly6_kern * (T)p.particle_mass;
            psc[i] = eos_abs<T>(p, rho) + eos_slope<T>(p, rho) * rsc[i];
        }
    }
    // 6. force + velocity update, scattered to original order (System.cu:228-250,402).
    std::vector<T> dv(4 * (size_t)n, (T)0), dsc(n, (T)0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        T f[3];
        T* scp = dbl ? &dsc[i] : nullptr;
        if (prec == SPH_PREC_KAHAN)
            force_sum<T, true>(p, (uint32_t)i, spos, svel, dens, pres, psc, cs, ce, f, scp);
        else
            force_sum<T, false>(p, (uint32_t)i, spos, svel, dens, pres, psc, cs, ce, f, scp);
        for (int c = 0; c < 3; ++c) dv[4 * i + c] = f[c];
    }
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t si = perm[k];
        for (int c = 0; c < 3; ++c) vel[4 * (size_t)si + c] = svel[4 * (size_t)k + c] + dv[4 * (size_t)k + c];
        vel[4 * (size_t)si + 3] = svel[4 * (size_t)k + 3] + (T)0;
        if (o.dens) o.dens[si] = dens[k];
        if (o.pres) o.pres[si] = pres[k];
        if (o.dvel4)
            for (int c = 0; c < 4; ++c) o.dvel4[4 * (size_t)si + c] = c < 3 ? dv[4 * (size_t)k + c] : (T)0;
        if (o.dvel_scale) o.dvel_scale[si] = dsc[k];
        if (o.dens_scale) o.dens_scale[si] = rsc[k];
        if (o.pres_scale) o.pres_scale[si] = psc[k];
        if (o.sorted_hash) o.sorted_hash[k] = hash[si];
        if (o.perm) o.perm[k] = si;
    }
    if (o.dye)
        for (uint32_t k = 0; k < n; ++k) {
            T col[3];
            const uint32_t si = perm[k];
            if (colour_one<T>(p, k, si, spos, svel, dens, &dv[4 * (size_t)k], cs, ce, o.dye, col) && o.colour4) {
                for (int c = 0; c < 3; ++c) o.colour4[4 * (size_t)si + c] = col[c];
                o.colour4[4 * (size_t)si + 3] = 1;
            }
        }
    if (o.pos4) std::copy(pos.begin(), pos.end(), o.pos4);
    if (o.vel4) std::copy(vel.begin(), vel.end(), o.vel4);
    if (o.cell_start) std::copy(cs.begin(), cs.end(), o.cell_start);
    if (o.cell_end) std::copy(ce.begin(), ce.end(), o.cell_end);
    if (o.cnt) *o.cnt = cnt;
}

// App::UpdateEmitter runs before every cSPH::Update (source/App/Render.cpp:10-11)
// and advances the rotor angle, which is also the wave phase:
// rAngle += rVel * timeStep (source/App/Update.cpp:15-16).
inline sph_params advanced(const sph_params& p) {
    sph_params q = p;
    if (q.dye_clear > 0) q.dye_clear = q.dye_clear - 1;  // source/App/Update.cpp:23
    if (q.rot_vel != 0.0f) {
        q.rot_angle = q.rot_angle + q.rot_vel * q.time_step;
        if (q.rot2_dist > 0.0f) {  // the second propeller (source/App/Update.cpp:17-20)
            q.rot2_angle = q.rot2_angle + (q.rot_vel * q.rot2_vel) * q.time_step;
            q.rot2_twist = q.rot2_vel < 0.0f ? -1.0f : 1.0f;
        }
    }
    return q;
}

}  // namespace

extern "C" {

// Number of OpenMP threads the oracle uses (1 when built without OpenMP).
int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// Every step entry point below first advances rot_angle by
// rot_vel * time_step (and the second propeller's angle; no-op for a still
// scene), as each library step does; oracle_run_f32 carries the angles over
// its steps.  The stage
// functions further down use the parameters as given.
//
// One FP32 step.  pos4/vel4 are updated in place (original order); every
// other output pointer may be NULL.  prec is SPH_PREC_REFERENCE or
// SPH_PREC_KAHAN.  counters receives {cap_cells, out_of_grid, max_cell_count, 0}.
int oracle_step_f32(const sph_params* p, int prec, uint32_t n, float* pos4, float* vel4,
                    float* dens, float* pres, float* dvel4, uint32_t* sorted_hash, uint32_t* perm,
                    uint32_t* cell_start, uint32_t* cell_end, uint32_t* counters) {
    if (!p || !pos4 || !vel4 || n != p->num_particles) return 1;
    StepOut<float> o{pos4, vel4, nullptr, dens, pres, dvel4, nullptr, nullptr, nullptr, sorted_hash,
                     perm, cell_start, cell_end, reinterpret_cast<Counters*>(counters)};
    step_impl<float>(advanced(*p), prec, n, pos4, vel4, false, o);
    return 0;
}

// The same step, also returning the velocity after integrate (before the
// force update) in vel_int4 (original order, float4).
int oracle_step_f32_ex(const sph_params* p, int prec, uint32_t n, float* pos4, float* vel4,
                       float* vel_int4, float* dens, float* pres, float* dvel4,
                       uint32_t* sorted_hash, uint32_t* perm, uint32_t* cell_start,
                       uint32_t* cell_end, uint32_t* counters) {
    if (!p || !pos4 || !vel4 || n != p->num_particles) return 1;
    StepOut<float> o{pos4, vel4, vel_int4, dens, pres, dvel4, nullptr, nullptr, nullptr,
                     sorted_hash, perm, cell_start, cell_end, reinterpret_cast<Counters*>(counters)};
    step_impl<float>(advanced(*p), prec, n, pos4, vel4, false, o);
    return 0;
}

// Several FP32 steps in a row (outputs of the last step only).
int oracle_run_f32(const sph_params* p, int prec, uint32_t n, uint32_t nsteps, float* pos4,
                   float* vel4, float* dens, float* pres) {
    if (!p || !pos4 || !vel4 || n != p->num_particles) return 1;
    sph_params q = *p;
    for (uint32_t s = 0; s < nsteps; ++s) {
        StepOut<float> o{pos4, vel4, nullptr, dens, pres, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr, nullptr, nullptr, nullptr};
        q = advanced(q);
        step_impl<float>(q, prec, n, pos4, vel4, false, o);
    }
    return 0;
}

// nsteps FP32 steps with the colour output: dye (float per particle,
// original order) is read and updated; colour4 receives the last step's
// colours (float4, original order; untouched for SPH_CLR_NONE).
int oracle_run_f32_colour(const sph_params* p, int prec, uint32_t n, uint32_t nsteps, float* pos4,
                          float* vel4, float* dye, float* colour4) {
    if (!p || !pos4 || !vel4 || !dye || n != p->num_particles) return 1;
    sph_params q = *p;
    for (uint32_t s = 0; s < nsteps; ++s) {
        StepOut<float> o{pos4, vel4, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr, nullptr, nullptr, nullptr, colour4, dye};
        q = advanced(q);
        step_impl<float>(q, prec, n, pos4, vel4, false, o);
    }
    return 0;
}

// One FP64 step from FP32 inputs.  Outputs are double, original order.
// dvel_scale[i] = m*dt * sum_j |c d12| (|pterm|_abs r + |vterm| |dv_ij|): the
// absolute-value evaluation of the force sum, the scale FP32 rounding error
// of that sum is proportional to; dens_scale[i] = m Poly6 sum_j [c^3 +
// 3 c^2 (h2 + r2)] and pres_scale[i] = |p|_abs + |dp/drho| dens_scale[i] are
// the matching scales of the density and pressure.  skip_integrate != 0 treats the inputs as
// the state after integrate (hash, sort, density and force only), so the
// density and force sums can be compared on the exact inputs the FP32 path
// used.
int oracle_step_f64(const sph_params* p, uint32_t n, const float* pos4_in, const float* vel4_in,
                    int skip_integrate, double* pos4, double* vel4, double* dens, double* pres,
                    double* dvel4, double* dvel_scale, double* dens_scale, double* pres_scale) {
    if (!p || !pos4_in || !vel4_in || n != p->num_particles) return 1;
    std::vector<double> pin(pos4_in, pos4_in + 4 * (size_t)n), vin(vel4_in, vel4_in + 4 * (size_t)n);
    StepOut<double> o{pos4, vel4, nullptr, dens, pres, dvel4, dvel_scale, dens_scale, pres_scale,
                      nullptr, nullptr, nullptr, nullptr, nullptr};
    step_impl<double>(advanced(*p), SPH_PREC_REFERENCE, n, pin.data(), vin.data(), skip_integrate != 0, o);
    return 0;
}

// ---- stage functions for unit tests -------------------------------------

// integrateD on n particles (original order), in place.
void oracle_integrate_f32(const sph_params* p, uint32_t n, float* pos4, float* vel4) {
    for (uint32_t i = 0; i < n; ++i) {
        float x[3] = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2]};
        float v[3] = {vel4[4 * i], vel4[4 * i + 1], vel4[4 * i + 2]};
        integrate_one<float>(*p, x, v);
        for (int c = 0; c < 3; ++c) {
            pos4[4 * i + c] = x[c];
            vel4[4 * i + c] = v[c];
        }
    }
}

// calcHashD on n positions; returns the number of clamped cells.
uint32_t oracle_hash_f32(const sph_params* p, uint32_t n, const float* pos4, uint32_t* hash) {
    uint32_t oob = 0;
    for (uint32_t i = 0; i < n; ++i) {
        float x[3] = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2]};
        int g[3];
        grid_pos<float>(*p, x, g, &oob);
        hash[i] = grid_hash(*p, g[0], g[1], g[2]);
    }
    return oob;
}

// Density of sorted particles given sorted positions and cell ranges
// (float, chosen accumulation mode).  dens/pres in sorted order.
void oracle_density_f32(const sph_params* p, int prec, uint32_t n, const float* spos4,
                        const uint32_t* cell_start, const uint32_t* cell_end, float* dens,
                        float* pres) {
    std::vector<float> sp(spos4, spos4 + 4 * (size_t)n);
    std::vector<uint32_t> cs(cell_start, cell_start + p->num_cells),
        ce(cell_end, cell_end + p->num_cells);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float s = prec == SPH_PREC_KAHAN ? density_sum<float, true>(*p, (uint32_t)i, sp, cs, ce, p->h2)
                                         : density_sum<float, false>(*p, (uint32_t)i, sp, cs, ce, p->h2);
        float rho = (s * p->poly6_kern) * p->particle_mass;
        dens[i] = rho;
        pres[i] = eos<float>(*p, rho);
    }
}

// Neighbour counts of the density walk (computeDensityD's loops,
// source/CUDA/Kernel_Cell.cui:142-199): for each sorted slot the candidates
// j != i of the 27 cells, at most max_par_in_cell per cell, with
// r2 < h2 (pred 0, the density predicate) or max(minDist, sqrtf(r2)) < h
// (pred 1, the force predicate of compForcePair, Kernel_Cell.cui:214-218).
// counts in sorted order.
void oracle_neighbour_counts_f32(const sph_params* p, int pred, uint32_t n, const float* spos4,
                                 const uint32_t* cell_start, const uint32_t* cell_end, uint32_t* counts) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float xi[3] = {spos4[4 * i], spos4[4 * i + 1], spos4[4 * i + 2]};
        int g[3];
        grid_pos<float>(*p, xi, g, nullptr);
        uint32_t cnt = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int cx = g[0] + dx, cy = g[1] + dy, cz = g[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)p->grid_size[0] || cy >= (int)p->grid_size[1] ||
                        cz >= (int)p->grid_size[2])
                        continue;
                    const uint32_t c = grid_hash(*p, cx, cy, cz);
                    const uint32_t s = cell_start[c];
                    if (s == kEmpty) continue;
                    const uint32_t e = std::min(cell_end[c], s + p->max_par_in_cell);
                    for (uint32_t j = s; j < e; ++j) {
                        if (j == i) continue;
                        const float rx = xi[0] - spos4[4 * j], ry = xi[1] - spos4[4 * j + 1],
                                    rz = xi[2] - spos4[4 * j + 2];
                        const float r2 = rx * rx + ry * ry + rz * rz;
                        // the force predicate as compForcePair evaluates it (a NaN r2
                        // passes: max(minDist, NaN) is minDist)
                        const bool pass = pred == 0 ? r2 < p->h2 : std::fmax(p->min_dist, std::sqrt(r2)) < p->h;
                        if (pass) ++cnt;
                    }
                }
        counts[i] = cnt;
    }
}

// Brute-force O(n^2) density in double: all pairs j != i with r2 < h^2,
// no grid.  The reference for neighbour-set completeness when cell >= h and
// no cell exceeds max_par_in_cell.
void oracle_density_bruteforce_f64(const sph_params* p, uint32_t n, const float* pos4,
                                   double* dens, uint32_t* nnb) {
    const double h2 = (double)p->h * (double)p->h;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double s = 0;
        uint32_t cnt = 0;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == (uint32_t)i) continue;
            double rx = (double)pos4[4 * i] - pos4[4 * j], ry = (double)pos4[4 * i + 1] - pos4[4 * j + 1],
                   rz = (double)pos4[4 * i + 2] - pos4[4 * j + 2];
            double r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < h2) {
                double c = h2 - r2;
                s += c * c * c;
                ++cnt;
            }
        }
        dens[i] = s * (double)p->poly6_kern * (double)p->particle_mass;
        if (nnb) nnb[i] = cnt;
    }
}

}  // extern "C"
