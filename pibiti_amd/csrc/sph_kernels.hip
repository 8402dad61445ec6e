// sph_kernels.hip - HIP kernels of one WCSPH step for gfx950 (MI355X).
//
// Data layout in HBM: the particle state lives in SORTED order (the order of
// the last stable sort by (cell key, original index)), one float4 per
// particle for position (x,y,z,w) and velocity (vx,vy,vz,w), a uint32
// original index per slot, and a float2 (density, pressure) per slot.  The
// reference keeps positions in original order and scatters velocities back
// every step (source/CUDA/System.cu:402); keeping the sorted order instead
// makes integrate, reorder and force streams coalesced, and the original
// order is only materialised on readback.
//
// Every kernel serves one scene (parameters passed by value) or a batch of
// independent scenes packed into one particle array: each scene owns a
// disjoint range of global cell keys (cell_base + local hash), so one sort
// and one launch per kernel advance every scene, and a slot finds its
// scene's parameters through a per-slot scene id.
//
// Arithmetic follows the reference kernels operation by operation (no FMA
// contraction: the library is built with -ffp-contract=off and correctly
// rounded fp32 division/sqrt), so one thread per particle reproduces the
// reference's FP32 result bit for bit:
//   integrateD + boundary .......... source/CUDA/System.cu:41-159,165-205
//   calcGridPos/calcGridHash ....... source/CUDA/Kernel_Cell.cui:5-34
//   reorderD ........................ source/CUDA/Kernel_Cell.cui:40-67
//   computeDensityD ................. source/CUDA/Kernel_Cell.cui:142-199
//   computeForceD ................... source/CUDA/Kernel_Cell.cui:210-261,
//                                     source/CUDA/System.cu:228-250,402
#include <type_traits>

#include "sph_device.hpp"

namespace sph {
namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr float kBndEps = 0.00001f;  // EPS of boundary()

__device__ __forceinline__ void bnd_wall(float diff, float nx, float ny, float nz, float stiff,
                                         float damp, float dt, float& vx, float& vy, float& vz) {
    if (diff > kBndEps) {
        float dot = nx * vx + ny * vy + nz * vz;
        float acc = stiff * diff - damp * dot;
        vx = vx + (acc * nx) * dt;
        vy = vy + (acc * ny) * dt;
        vz = vz + (acc * nz) * dt;
    }
}

// Correctly rounded float atan, sin and cos for the pump's per-particle
// terms: one rounding of the double result, so the CPU oracle (libm) agrees.
__device__ __forceinline__ float atan_rn(float x) { return (float)atan((double)x); }
__device__ __forceinline__ float sin_rn(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cos_rn(float x) { return (float)cos((double)x); }

// boundary() (source/CUDA/System.cu:41-159): the soft walls of the scene's
// boundary type and z effect, in the reference's order, plus the position
// wraps of the wrap/cycle effects and the pump outlet.  The shape branches
// are uniform over a scene.
__device__ __forceinline__ void boundary(const sph_params& p, float dt, float& x, float& y, float& z,
                                         float& vx, float& vy, float& vz) {
    const float b = p.dist_bnd_soft, stiff = p.bnd_stiff, damp = p.bnd_damp, damp2 = p.bnd_damp_c;
    const float wmin_x = p.world_min[0], wmin_y = p.world_min[1];
    float wmin_z = p.world_min[2];
    const float wmax_x = p.world_max[0], wmax_y = p.world_max[1], wmax_z = p.world_max[2];
    const uint32_t t = p.bnd_type, eff = p.bnd_eff_z;
    const bool cyl_y = t == SPH_BND_CYL_Y, cyl_z = t == SPH_BND_CYL_Z;
    const bool wave = eff == SPH_BND_EFF_WAVE, no_eff = eff == SPH_BND_EFF_NONE,
               cycle = eff == SPH_BND_EFF_CYCLE;
    if (wave) {
        // sloped floor, then the moving z- wall
        const float sl = -p.rot2_angle;
        bnd_wall((b - (y - wmin_y)) - (z - wmin_z) * sl, 0.0f, 1.0f - sl, sl, stiff, damp, dt, vx, vy, vz);
        wmin_z = wmin_z + p.wave_zmin_shift;
    }
    if (t != SPH_BND_SPHERE) {
        if (!cyl_y) {
            if (no_eff || wave) bnd_wall((b - z) + wmin_z, 0.0f, 0.0f, 1.0f, stiff, damp2, dt, vx, vy, vz);
            if (!cycle) bnd_wall((b + z) - wmax_z, 0.0f, 0.0f, -1.0f, stiff, damp2, dt, vx, vy, vz);
        }
        if (!cyl_y && !cyl_z) {
            bnd_wall((b - x) + wmin_x, 1.0f, 0.0f, 0.0f, stiff, damp, dt, vx, vy, vz);
            bnd_wall((b + x) - wmax_x, -1.0f, 0.0f, 0.0f, stiff, damp, dt, vx, vy, vz);
        }
        if (!cyl_z) {
            bnd_wall((b - y) + wmin_y, 0.0f, 1.0f, 0.0f, stiff, damp, dt, vx, vy, vz);
            bnd_wall((b + y) - wmax_y, 0.0f, -1.0f, 0.0f, stiff, damp, dt, vx, vy, vz);
        }
    } else {
        const float len = sqrt_ieee((x * x + y * y) + z * z);
        const float diff = (b + len) + wmin_y;
        if (diff > kBndEps) bnd_wall(diff, -x / len, -y / len, -z / len, stiff, damp2, dt, vx, vy, vz);
    }
    if (cyl_y || t == SPH_BND_CYL_YZ) {
        const float len = sqrt_ieee(x * x + z * z);
        const float diff = (b + len) - wmax_x;
        if (diff > kBndEps) bnd_wall(diff, -x / len, 0.0f, -z / len, stiff, damp2, dt, vx, vy, vz);
    }
    if (cyl_z || t == SPH_BND_CYL_YZ) {
        const float len = sqrt_ieee(x * x + y * y);
        const float diff = (b + len) + wmin_y;
        if (diff > kBndEps) bnd_wall(diff, -x / len, -y / len, 0.0f, stiff, damp2, dt, vx, vy, vz);
    }
    if (!wave && !no_eff) {
        // wrap / cycle along z
        const float dr = 1.0f * p.particle_r;
        const float span = ((wmax_z - wmin_z) - 2.0f * b) - dr;
        if (cycle && vz > p.exit_vel && z > (wmax_z - b) - dr)
            z = z - span;
        else if (vz < -p.exit_vel && z < (wmin_z + b) + dr)
            z = z + span;
    }
    if (t == SPH_BND_PUMP_Y) {
        const float rad = wmax_x, ang = p.pump_ang_out, hc = p.h_close;
        const float rin = rad * p.pump_rad_in;
        const float len = sqrt_ieee(x * x + y * y);
        float diff = (b + len) - rad;
        if (diff > kBndEps) {
            // frame: the cylinder wall except over the outlet
            const float a = atan_rn(x / y);
            const bool wall = ang < 0.5f ? (a < -ang || a > ang || y < 0.0f)
                                         : (y < 0.0f || (len < rad * p.shape_s[4] && a < ang));
            if (wall) bnd_wall(diff, -x / len, -y / len, 0.0f, stiff, damp, dt, vx, vy, vz);
        }
        const float xs = p.pump_xs;
        if (ang < 0.5f) {
            if (y > p.pump_zs) {
                bnd_wall((b - x) - xs, 1.0f, 0.0f, 0.0f, stiff, damp, dt, vx, vy, vz);
                bnd_wall((b + x) - xs, -1.0f, 0.0f, 0.0f, stiff, damp, dt, vx, vy, vz);
            }
        } else if (len >= rad * p.shape_s[5]) {
            bnd_wall((b - x) + xs, 1.0f, 0.0f, 0.0f, stiff, damp, dt, vx, vy, vz);
        }
        // inlet hole (closed above) and the lid
        if (z > hc - b * p.shape_s[0]) {
            diff = (b + len) - rin;
            if (diff > kBndEps) bnd_wall(diff, -x / len, -y / len, 0.0f, stiff, damp, dt, vx, vy, vz);
        }
        if (z < hc - b * p.shape_s[1]) bnd_wall((b + z) - hc, 0.0f, 0.0f, -1.0f, stiff, damp, dt, vx, vy, vz);
        // outlet -> inlet: map the outlet rectangle onto the inlet circle
        diff = (y - wmax_y) + p.exit_dist;
        if (diff > kBndEps && vy > p.exit_vel) {
            const float zz = fabsf(hc - wmin_z);
            float rr, aa;
            if (ang < 0.5f) {
                const float xx = xs * 2.0f;
                rr = ((x + xx / 2.0f) / xx) * 0.7f;
                aa = ((z - zz / 2.0f) / zz) * 1.6f;
            } else {
                rr = ((wmax_x - x) / xs) * 0.45f;
                aa = ((z - zz / 2.0f) / zz) * 1.8f;
            }
            rr = rr * rin;
            aa = aa * (2.0f * 3.141592654f);
            const float nx = cos_rn(aa) * rr, ny = sin_rn(aa) * rr;
            // the 0.01 of the reference is a double literal: this term is double
            const float nz = (float)((double)(wmax_z - b) - (double)fabsf(vy - p.exit_vel) * 0.01);
            x = nx;
            y = ny;
            z = nz;
            const float nvy = vz, nvz = -vy;
            vy = nvy;
            vz = nvz;
        }
    }
}

// collideSpheresR (source/CUDA/Kernel_Cell.cui:101-115): the spring and
// damping force of one rotor or height-map sphere at offset r from the
// particle (u = -v), added to a (zero outside the sphere's reach).
__device__ __forceinline__ void collide_r(const sph_params& p, float rx, float ry, float rz, float ux,
                                          float uy, float uz, float rad, float& ax, float& ay, float& az) {
    const float dist = sqrt_ieee((rx * rx + ry * ry) + rz * rz);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (dist < rad) {
        const float inv = 1.0f / dist;  // float3 / float
        const float sp = p.coll_spring * (dist - rad);
        fx = sp * (rx * inv) + p.coll_damping * ux;
        fy = sp * (ry * inv) + p.coll_damping * uy;
        fz = sp * (rz * inv) + p.coll_damping * uz;
    }
    ax = ax + fx;
    ay = ay + fy;
    az = az + fz;
}

constexpr float kPiRef = 3.141592654f;  // PI of source/CUDA/Params.cuh:18

// The terms computeForceD adds to the velocity increment after the pair sum
// (source/CUDA/System.cu:253-399), in its order: the height map, then the
// rotor or, without one, the sphere collider (collideSpheres,
// source/CUDA/Kernel_Cell.cui:78-96), then the accelerators.  Every branch
// is uniform over a scene.
__device__ __forceinline__ void extra_dv(const sph_params& p, float dt, const float4& xi,
                                         const float4& vi, float& ax, float& ay, float& az) {
    const float ux = vi.x * -1.0f, uy = vi.y * -1.0f, uz = vi.z * -1.0f;
    if (p.physics & SPH_PHYS_HMAP) {
        // a sheet of spheres on the rot_spc lattice of the x (or y) and z
        // axes, 5 x 5 around the particle's own site, two layers deep
        const bool xz = p.hmap_type == SPH_HMAP_XZ;
        const float rr = p.particle_r + p.rot_r, spc = p.rot_spc;
        const float xa = xz ? xi.x : xi.y, wa = xz ? p.world_min[0] : p.world_min[1];
        const float wl = xz ? p.world_min[1] : p.world_min[0];
        const float sa = xz ? p.world_size_d[0] : p.world_size_d[1];
        const int ia = (int)((xa - wa) / spc);
        const int iz = (int)((xi.z - p.world_min[2]) / spc);
        const float af = (float)ia * spc + wa, zf = (float)iz * spc + p.world_min[2];
        const float ph1 = (p.shape_s[1] * kPiRef) / 180.0f, ph3 = (p.shape_s[3] * kPiRef) / 180.0f;
        for (int j = -2; j <= 2; ++j)
            for (int i = -2; i <= 2; ++i) {
                const float ah = af + (float)i * spc, zh = zf + (float)j * spc;
                const float an = ((ah / sa) * 2.0f) * kPiRef, zn = ((zh / p.world_size_d[2]) * 2.0f) * kPiRef;
                const float ss = sin_rn(p.shape_s[0] * an + ph1) * sin_rn(p.shape_s[2] * zn + ph3);
                if (!(p.shape_s[4] <= -1.0f || ss > p.shape_s[4])) continue;
                const float lf = ((ss * p.h_close) + p.rot2_angle) + wl;
                for (int k = 0; k <= 1; ++k) {
                    const float lh = lf - (float)k * spc;
                    if (xz)
                        collide_r(p, ah - xi.x, lh - xi.y, zh - xi.z, ux, uy, uz, rr, ax, ay, az);
                    else
                        collide_r(p, lh - xi.x, ah - xi.y, zh - xi.z, ux, uy, uz, rr, ax, ay, az);
                }
            }
    }
    if (p.physics & SPH_PHYS_ROTOR) {
        const int sx = p.rot_size[0], sy = p.rot_size[1], sz = p.rot_size[2], cb = p.rot_blades;
        const float sp = p.rot_spc, ca = (2.0f * kPiRef) / (float)cb;
        const float x2 = (float)sx * 0.5f, y2 = (float)sy * 0.5f;
        const float rr = p.particle_r + p.rot_r;
        const float cx = p.coll_pos[0], cy = p.coll_pos[1], cz = p.coll_pos[2];
        if (p.rot_type != SPH_ROTOR_PROPELLER_Z) {
            // pump rotor: blades of spheres along z (type 1) or y (type 2)
            const bool zax = p.rot_type == SPH_ROTOR_PUMP_Z;
            const float q = zax ? xi.z : xi.y, c0 = zax ? cz : cy;
            if (q > c0 - rr && q < (c0 + rr) + sp * (float)sy) {
                for (int z = 1; z <= sz; ++z)
                    for (int x = 0; x <= sx; ++x)
                        for (int c = 0; c < cb; ++c) {
                            const float a = ((-p.rot_angle) + (float)c * ca) + (float)z * p.rot_twist;
                            const float cs = cos_rn(a) * sp, sn = (-sin_rn(a)) * sp;
                            const float u = (float)x - x2;
                            const float pa = u * cs - (float)z * sn, pb = u * sn + (float)z * cs;
                            for (int y = 0; y <= sy; ++y) {
                                const float py = sp * (float)y;
                                const float qx = cx + pa, qy = cy + (zax ? pb : py), qz = cz + (zax ? py : pb);
                                collide_r(p, qx - xi.x, qy - xi.y, qz - xi.z, ux, uy, uz, rr, ax, ay, az);
                            }
                        }
            }
        } else {
            // propeller about z (two of them rot2_dist apart when set)
            const float half = (sp * (float)sz) / 2.0f;
            if (xi.z > (cz - rr) - half && xi.z < (cz + rr) + half) {
                float px0 = cx, aa = p.rot_angle, tw = p.rot_twist;
                if (p.rot2_dist > 0.0f) {
                    if (xi.x > 0.0f) {
                        px0 = px0 + p.rot2_dist * 0.5f;
                    } else {
                        px0 = px0 - p.rot2_dist * 0.5f;
                        aa = p.rot2_angle;
                        tw = tw * p.rot2_twist;
                    }
                }
                for (int c = 0; c < cb; ++c) {
                    const float a = aa + (float)c * ca;
                    for (int h = 0; h <= sy; ++h) {
                        const float dh = h == sy - 1 ? 0.5f : (h == sy ? 1.2f : 0.0f);
                        for (int x = 0; x <= sz; ++x) {
                            const float d = (x == sz && dh == 0.0f) ? 0.4f : dh;
                            const float k = cos_rn(((float)x - x2) * 0.2f + kPiRef * 0.6f);
                            const float at = tw * (d * -0.05f - ((float)h - y2) * k);
                            const float ac = a + at;
                            const float cs = cos_rn(ac) * sp, sn = (-sin_rn(ac)) * sp;
                            const float pz = (sp * (((-d) + (float)h) - y2)) * k;
                            const float r = p.rot_r * fabsf(1.0f - d);
                            collide_r(p, (px0 + (float)x * cs) - xi.x, (cy + (float)x * sn) - xi.y,
                                      (cz + pz) - xi.z, ux, uy, uz, p.particle_r + r, ax, ay, az);
                        }
                    }
                }
            }
        }
    } else if (p.physics & SPH_PHYS_COLLIDER) {
        const float rx = p.coll_pos[0] - xi.x, ry = p.coll_pos[1] - xi.y, rz = p.coll_pos[2] - xi.z;
        const float dist = sqrt_ieee((rx * rx + ry * ry) + rz * rz);
        const float rad = p.particle_r + p.coll_r;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        if (dist < rad) {
            const float inv = 1.0f / dist;  // float3 / float
            const float nx = rx * inv, ny = ry * inv, nz = rz * inv;
            const float d = (ux * nx + uy * ny) + uz * nz;
            const float tx = ux - d * nx, ty = uy - d * ny, tz = uz - d * nz;
            const float sp = p.coll_spring * (dist - rad);
            fx = (sp * nx + p.coll_damping * ux) + p.coll_shear * tx;
            fy = (sp * ny + p.coll_damping * uy) + p.coll_shear * ty;
            fz = (sp * nz + p.coll_damping * uz) + p.coll_shear * tz;
        }
        ax = ax + fx;
        ay = ay + fy;
        az = az + fz;
    }
    if (p.physics & SPH_PHYS_ACCEL) {
        for (int k = 0; k < SPH_NUM_ACCEL; ++k) {
            const sph_accel& c = p.acc[k];
            if (c.type == SPH_ACC_OFF) continue;
            const float rx = xi.x - c.pos[0], ry = xi.y - c.pos[1], rz = xi.z - c.pos[2];
            float g = 0.0f;
            bool in = false;
            if (c.type == SPH_ACC_BOX) {
                in = fabsf(rx) < c.size[0] && fabsf(ry) < c.size[1] && fabsf(rz) < c.size[2];
            } else if (c.type == SPH_ACC_CYL_Y) {
                const float qx = rx / c.size[0], qz = rz / c.size[2];
                in = fabsf(ry) < c.size[1] && sqrt_ieee(qx * qx + qz * qz) < 1.0f;
            } else {
                const float r = sqrt_ieee(rx * rx + rz * rz);
                in = fabsf(ry) < c.size[1] && r < c.size[0];
                g = 1.0f - r / c.size[2];
            }
            if (!in) continue;
            if (c.type == SPH_ACC_CYL_Y_SMOOTH) {
                ax = ax + (c.acc[0] * g) * dt;
                ay = ay + (c.acc[1] * g) * dt;
                az = az + (c.acc[2] * g) * dt;
            } else {
                ax = ax + c.acc[0] * dt;
                ay = ay + c.acc[1] * dt;
                az = az + c.acc[2] * dt;
            }
        }
    }
}

// The rest of computeForceD's velocity update for scenes with a height map,
// rotor, collider or accelerators: vel_out holds the pair increment the
// force kernel left there; add the other terms in the reference's order and
// store v + increment.  A kernel of its own so that the force kernels keep
// their register budget.
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_extra_terms(sph_params one, const sph_params* __restrict__ tab,
                                                        const uint32_t* __restrict__ sid, uint32_t n,
                                                        const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel,
                                                        float4* __restrict__ vel_out,
                                                        float4* __restrict__ dvel,
                                                        const float* __restrict__ dt_dev) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, i, scene);
    if (!p.physics) return;
    const float4 xi = pos[i], vi = vel[i], a = vel_out[i];
    float ax = a.x, ay = a.y, az = a.z;
    extra_dv(p, dt_dev[scene], xi, vi, ax, ay, az);
    vel_out[i] = make_float4(vi.x + ax, vi.y + ay, vi.z + az, vi.w + 0.0f);
    if (dvel) dvel[i] = make_float4(ax, ay, az, 0.0f);
}

// ---------------------------------------------------------------- colours
__device__ __forceinline__ float dot4(const float4& a, const float4& b) {
    return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;  // cutil_math.h:540-543
}
__device__ __forceinline__ float len3(float x, float y, float z) { return sqrt_ieee((x * x + y * y) + z * z); }
__device__ __forceinline__ float phi_clamp(float v, float lo, float hi) {
    return (fminf(v, hi) - fminf(v, lo)) / (hi - lo);  // phiClamp, System.cu:221-223
}

// Foam potential of the velocity colour mode (source/CUDA/System.cu:424-486):
// the trapped-air potential from the velocity differences of every other
// particle in the 27 cells (x, y, z order, at most max_par_in_cell per
// cell) weighted by W(r) = (1 - r)/3 for r <= 3, clamped into [5, 20],
// times the kinetic-energy potential clamped into [0.1, 1], times 3.
// normalize() is v * (1/sqrtf(v.v)) over all four components
// (cutil_math.h:552-556).
__device__ float foam(const sph_params& p, uint32_t i, const float4& xi, const float4& vi,
                      const float4* __restrict__ pos, const float4* __restrict__ vel,
                      const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end) {
    bool out = false;
    const int gx = cell_coord(xi.x, p.world_min[0], p.cell_size[0], p.grid_size[0], out);
    const int gy = cell_coord(xi.y, p.world_min[1], p.cell_size[1], p.grid_size[1], out);
    const int gz = cell_coord(xi.z, p.world_min[2], p.cell_size[2], p.grid_size[2], out);
    float vdiff = 0.0f;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const int cx = gx + dx, cy = gy + dy, cz = gz + dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)p.grid_size[0] || cy >= (int)p.grid_size[1] ||
                    cz >= (int)p.grid_size[2])
                    continue;
                const uint32_t c = grid_hash(p, cx, cy, cz);
                const uint32_t s = cell_start[c];
                if (s == kEmptyCell) continue;
                const uint32_t e = min(cell_end[c], s + p.max_par_in_cell);
                for (uint32_t j = s; j < e; ++j) {
                    if (j == i) continue;
                    const float4 v2 = vel[j], x2 = pos[j];
                    const float4 sv = make_float4(vi.x + v2.x, vi.y + v2.y, vi.z + v2.z, vi.w + v2.w);
                    const float4 sx = make_float4(xi.x + x2.x, xi.y + x2.y, xi.z + x2.z, xi.w + x2.w);
                    const float iv = 1.0f / sqrt_ieee(dot4(sv, sv)), ix = 1.0f / sqrt_ieee(dot4(sx, sx));
                    const float4 nv = make_float4(sv.x * iv, sv.y * iv, sv.z * iv, sv.w * iv);
                    const float4 nx = make_float4(sx.x * ix, sx.y * ix, sx.z * ix, sx.w * ix);
                    const float4 dv = make_float4(vi.x - v2.x, vi.y - v2.y, vi.z - v2.z, vi.w - v2.w);
                    const float4 dp = make_float4(xi.x - x2.x, xi.y - x2.y, xi.z - x2.z, xi.w - x2.w);
                    const float r = sqrt_ieee(dot4(dp, dp));
                    const float w = r <= 3.0f ? (1.0f - r) / 3.0f : 0.0f;
                    vdiff = vdiff + (sqrt_ieee(dot4(dv, dv)) * (1.0f - dot4(nv, nx))) * w;
                }
            }
    const float trapped = phi_clamp(vdiff, 5.0f, 20.0f);
    const float ke = ((p.particle_mass * 1000.0f) * dot4(vi, vi)) * 0.5f;
    const float kpot = phi_clamp(ke, 0.1f, 1.0f);
    return kpot * (3.0f * trapped + 3.0f * 0.0f);
}

// computeForceD's colouring and dye (source/CUDA/System.cu:405-546) for
// one sorted slot: x, v the step's state before the force update, a the
// step's whole velocity increment, rho the step's density.
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_colour(sph_params one, const sph_params* __restrict__ tab,
                                                   const uint32_t* __restrict__ sid, uint32_t n,
                                                   const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                   const float2* __restrict__ rhop,
                                                   const float4* __restrict__ dvel, const uint32_t* __restrict__ pid,
                                                   const uint32_t* __restrict__ cell_start,
                                                   const uint32_t* __restrict__ cell_end,
                                                   const float* __restrict__ dt_dev, float4* __restrict__ colour,
                                                   float* __restrict__ dye) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, i, scene);
    if (p.clr_type == SPH_CLR_NONE) return;  // no colour and no dye update
    const float dt = dt_dev[scene];
    const float4 xi = pos[i], vi = vel[i], a = dvel[i];
    const float dens = rhop[i].x;
    const float br = p.brightness, ct = p.contrast;
    float cx = 0.2f, cy = 0.5f, cz = 1.0f, intens = 0.0f;
    switch (p.clr_type) {
        case SPH_CLR_VEL_ACC: {
            const float v = 2.5f * len3(vi.x, vi.y, vi.z);
            const float f = (0.02f * len3(a.x, a.y, a.z)) / dt;
            const float kv = br + ct * v, kf = ct * f;
            cx = cx * kv + 0.7f * kf;
            cy = cy * kv + 0.35f * kf;
            cz = cz * kv + 0.0f * kf;
        } break;
        case SPH_CLR_DENS_ACC: {
            const float d = (4.0f * (dens - p.rest_density)) / p.rest_density;
            const float f = (0.02f * len3(a.x, a.y, a.z)) / dt;
            const float kd = br + ct * d, kf = ct * f;
            cx = cx * kd + 0.7f * kf;
            cy = cy * kd + 0.7f * kf;
            cz = cz * kd + 0.0f * kf;
        } break;
        case SPH_CLR_VEL: {
            const float v = 2.5f * len3(vi.x, vi.y, vi.z);
            intens = br + ct * v;
            const float nd = foam(p, i, xi, vi, pos, vel, cell_start, cell_end);
            cx = cy = cz = nd;
        } break;
        case SPH_CLR_VEL_RGB: {
            const float k = ct * 2.5f;
            cx = 0.5f + vi.x * k;
            cy = 0.5f + vi.y * k;
            cz = 0.5f + vi.z * k;
        } break;
        case SPH_CLR_ACCEL: {
            const float f = (0.02f * len3(a.x, a.y, a.z)) / dt;
            const float kf = br + ct * f;
            intens = kf;
            cx = cx * kf;
            cy = cy * kf;
            cz = cz * kf;
        } break;
        default: {  // SPH_CLR_DENS
            const float d = (4.0f * (dens - p.rest_density)) / p.rest_density;
            const float kd = br + ct * d;
            intens = kd;
            cx = cx * kd;
            cy = cy * kd;
            cz = cz * kd;
        } break;
    }
    if (p.hue == 1) {
        cx = intens;
        cy = 0.0f;
    }
    const uint32_t si = pid[i];
    if (p.dye_clear > 0) {
        dye[si] = 0.0f;
    } else if (p.dye_type > 0) {
        float dc = dye[si];
        const float rx = xi.x - p.dye_pos[0], ry = xi.y - p.dye_pos[1], rz = xi.z - p.dye_pos[2];
        if (p.dye_type == SPH_DYE_BOX) {
            if (fabsf(rx) < p.dye_size[0] && fabsf(ry) < p.dye_size[1] && fabsf(rz) < p.dye_size[2]) dc = 1.0f;
        } else if (len3(rx, ry, rz) < p.dye_size[1]) {
            dc = 1.0f;
        }
        dc = dc - dt * p.dye_fade;
        if (dc < 0.0f) dc = 0.0f;
        if (dc >= 0.0f) dye[si] = dc;
        if (p.hue == 0) {
            cx = cx + 0.9f * dc;
            cy = cy + 0.9f * dc;
            cz = cz + 1.0f * dc;
        } else {
            cy = dc;
        }
    }
    colour[i] = make_float4(cx, cy, cz, 1.0f);
}

// ------------------------------------------------------------ integrate+hash
// One slot's boundary forces, gravity, damping, position update and clamp to
// the hard walls, on the state in registers, and its new cell key.  An
// out-of-grid key is counted when counters is given.
__device__ __forceinline__ uint32_t integrate_state(const sph_params& p, float dt, float4& x4, float4& v4,
                                                    uint32_t* __restrict__ counters) {
    float x = x4.x, y = x4.y, z = x4.z;
    float vx = v4.x, vy = v4.y, vz = v4.z;
    boundary(p, dt, x, y, z, vx, vy, vz);
    vx = vx + p.gravity[0] * dt;
    vy = vy + p.gravity[1] * dt;
    vz = vz + p.gravity[2] * dt;
    vx = vx * p.global_damping;
    vy = vy * p.global_damping;
    vz = vz * p.global_damping;
    x = x + vx * dt;
    y = y + vy * dt;
    z = z + vz * dt;
    const float bh = p.dist_bnd_hard;
    {
        float hi = p.world_max[0] - bh, lo = p.world_min[0] + bh;
        if (x > hi) x = hi;
        if (x < lo) x = lo;
    }
    {
        float hi = p.world_max[1] - bh, lo = p.world_min[1] + bh;
        if (y > hi) y = hi;
        if (y < lo) y = lo;
    }
    {
        float hi = p.world_max[2] - bh, lo = p.world_min[2] + bh;
        if (z > hi) z = hi;
        if (z < lo) z = lo;
    }
    x4 = make_float4(x, y, z, x4.w);
    v4 = make_float4(vx, vy, vz, v4.w);

    bool out = false;
    const int gx = cell_coord(x, p.world_min[0], p.cell_size[0], p.grid_size[0], out);
    const int gy = cell_coord(y, p.world_min[1], p.cell_size[1], p.grid_size[1], out);
    const int gz = cell_coord(z, p.world_min[2], p.cell_size[2], p.grid_size[2], out);
    if (out && counters) atomicAdd(&counters[kCntOutOfGrid], 1u);
    if (counters && (__builtin_isnan(x) || __builtin_isnan(y) || __builtin_isnan(z)))
        atomicAdd(&counters[kCntNaN], 1u);
    return grid_hash(p, gx, gy, gz);
}

// integrate_state on slot k of pos/vel, in place.
__device__ __forceinline__ uint32_t integrate_one(const sph_params& p, float dt, float4* __restrict__ pos,
                                                  float4* __restrict__ vel, uint32_t k,
                                                  uint32_t* __restrict__ counters) {
    float4 x4 = pos[k];
    float4 v4 = vel[k];
    const uint32_t h = integrate_state(p, dt, x4, v4, counters);
    pos[k] = x4;
    vel[k] = v4;
    return h;
}

// The integration in place, for steps that keep their order (the learned
// surrogate's steps between sorts).
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_integrate(sph_params one, const sph_params* __restrict__ tab,
                                                      const uint32_t* __restrict__ sid, uint32_t n,
                                                      float4* __restrict__ pos, float4* __restrict__ vel,
                                                      uint32_t* __restrict__ counters,
                                                      const float* __restrict__ dt_dev) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, tid, scene);
    (void)integrate_one(p, dt_dev[scene], pos, vel, tid, counters);
}

// counter[key] += d (mod 2^32) for every active lane with `on`, one atomic
// per distinct key of the wave (neighbouring slots mostly share keys).
__device__ __forceinline__ void wave_add(uint32_t* __restrict__ counter, uint32_t key, uint32_t d, bool on) {
    uint64_t pending = __ballot(on);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lk = __shfl(key, leader);
        const uint64_t same = __ballot(on && key == lk) & pending;
        if ((int)(threadIdx.x & 63u) == leader) atomicAdd(&counter[lk], d * (uint32_t)__popcll(same));
        pending &= ~same;
    }
}

// The same integration for the incremental re-sort, keeping only the new
// key (in slot order; the state is integrated again where k_resort_place
// moves it).  A slot whose key changed (every slot while the order is not
// valid) is a mover: its bit is set, its new cell's mover count raised and
// its original index listed there (the first in_cap of a cell).
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_integrate_moves(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pid,
    const uint32_t* __restrict__ skey, uint32_t* __restrict__ nkey, const uint32_t* __restrict__ valid,
    uint64_t* __restrict__ moved, uint32_t* __restrict__ in_cnt, uint32_t* __restrict__ inlist, uint32_t in_cap,
    uint32_t* __restrict__ tile_delta, uint32_t* __restrict__ key_pid, uint32_t* __restrict__ val_pid,
    uint32_t* __restrict__ counters, const float* __restrict__ dt_dev, uint32_t num_cells,
    const uint32_t* __restrict__ old_start, const uint32_t* __restrict__ old_end, const uint32_t* __restrict__ err,
    uint32_t* __restrict__ repair) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    // the step's counters start here: every kernel that raises them runs later
    if (tid < kCntNum) counters[tid] = 0u;
    if (tid >= n) return;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, tid, scene);
    // the key only: k_resort_place integrates the state again as it moves it
    // (and counts an out-of-grid cell there)
    float4 x4 = pos[tid];
    float4 v4 = vel[tid];
    const uint32_t h = integrate_state(p, dt_dev[scene], x4, v4, nullptr);
    // an order left behind by a failed step is never built upon
    const bool ok = *valid != 0 && (*err & kErrInvalid) == 0;
    uint32_t h_old = ok ? skey[tid] : kEmptyCell;
    if (ok) {
        // the slot must lie in its old cell's range; otherwise the step's
        // counts cannot be trusted and it takes the full sort
        bool in_range = h_old < num_cells;
        if (in_range) {
            const uint32_t os = old_start[h_old], oe = old_end[h_old];
            in_range = os != kEmptyCell && tid >= os && tid < oe;
        }
        if (!in_range) {
            atomicOr(repair, kRepairInputs);
            h_old = kEmptyCell;
        }
    }
    const bool mv = !ok || h_old != h;
    nkey[tid] = h;
    // the radix sort's input, for a step that turns out to need it
    const uint32_t id = pid[tid];
    key_pid[id] = h;
    val_pid[id] = tid;
    // the active lanes of the last wave are a prefix of it: lane 0 writes
    const uint64_t bits = __ballot(mv);
    if ((threadIdx.x & 63) == 0) moved[tid >> 6] = bits;
    if (mv) {
        const uint32_t t = atomicAdd(&in_cnt[h], 1u);
        if (t < in_cap) inlist[(size_t)h * in_cap + t] = pid[tid];
    }
    // per scan tile: movers in less movers out (a move within a tile is none)
    const uint32_t tile_new = h / kScanTile, tile_old = h_old / kScanTile;
    const bool cross = mv && (!ok || tile_new != tile_old);
    wave_add(tile_delta, tile_new, 1u, cross);
    wave_add(tile_delta, tile_old, ~0u, cross && ok && h_old != kEmptyCell);
}

// ----------------------------------------------- incremental re-sort: apply
// The state in the step's new order, one slot per thread, after the cell
// scan (and the radix sort on a full-sort step).  Each slot's state is
// integrated again from the last step's values (the integrate kernel keeps
// only the keys; the same arithmetic gives the same bits).
//   incremental: old slot k goes to its new cell's start + the cell's
//     stayers with a smaller original index (a stayer: its rank among the
//     old range's stayers; a mover: a binary search of the old range) + the
//     cell's listed movers in with a smaller index;
//   full sort: new slot k takes the sorted pair's old slot.
// Marks the new order valid; the step's own validity was copied to
// valid_step by the cell scan.  xyz_out (null unless the pair-entry gathers
// are in use) takes the positions packed, three floats per slot, which the
// pair-entry densities read their candidates from.
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_resort_place(
    sph_params one, const sph_params* __restrict__ tab, uint32_t n, const uint32_t* __restrict__ nkey,
    const uint64_t* __restrict__ moved, const uint32_t* __restrict__ valid_step, uint32_t* __restrict__ valid,
    const uint32_t* __restrict__ old_start, const uint32_t* __restrict__ old_end,
    uint32_t* __restrict__ new_start, uint32_t* __restrict__ new_end,
    const uint32_t* __restrict__ in_cnt, const uint32_t* __restrict__ inlist, uint32_t in_cap,
    const uint32_t* __restrict__ sorted_key,
    const uint32_t* __restrict__ sorted_slot, const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,
    const uint32_t* __restrict__ pid_in, const uint32_t* __restrict__ sid_in, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ pid_out, uint32_t* __restrict__ sid_out,
    uint32_t* __restrict__ skey_out, uint32_t* __restrict__ counters, uint32_t* __restrict__ err,
    const float* __restrict__ dt_dev, uint32_t* __restrict__ repair, uint32_t* __restrict__ repairs,
    float* __restrict__ xyz_out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        // a repaired step's tile totals may be off: the next step starts
        // its order from scratch
        const uint32_t r = *repair;
        *valid = r ? 0u : 1u;
        *repair = 0u;
        if (r) atomicAdd(repairs, 1u);
    }
    if (k >= n) return;
    const bool full = counters[kCntFullSort] != 0;
    uint32_t src, dst, c, id;
    if (full) {
        src = sorted_slot[k];
        dst = k;
        c = sorted_key[k];
        id = pid_in[src];
        // the cell ranges from the sorted keys (the same values the cell scan
        // wrote, except on a repaired step)
        if (k == 0 || sorted_key[k - 1] != c) new_start[c] = k;
        if (k + 1 == n || sorted_key[k + 1] != c) new_end[c] = k + 1;
    } else {
        src = k;
        c = nkey[k];
        id = pid_in[k];
        const bool mv = is_mover(moved, k);
        const uint32_t base = new_start[c], end = new_end[c];
        bool bad = base >= end || end > n;
        uint32_t os = kEmptyCell, oe = 0, stay = 0;
        if (*valid_step) {
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
        const uint32_t nin = in_cnt[c];
        if (bad || nin > in_cap || end - base != stay + nin) {
            atomicOr(err, kErrInvalid);
            return;
        }
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
        const uint32_t* inc = inlist + (size_t)c * in_cap;
        for (uint32_t q = 0; q < nin; ++q) inr += inc[q] < id ? 1u : 0u;
        dst = base + local + inr;
        if (dst >= end) {
            atomicOr(err, kErrInvalid);
            return;
        }
    }
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid_in, src, scene);
    float4 x4 = pos_in[src];
    float4 v4 = vel_in[src];
    (void)integrate_state(p, dt_dev[scene], x4, v4, counters);
    pos_out[dst] = x4;
    vel_out[dst] = v4;
    pid_out[dst] = id;
    skey_out[dst] = c;
    if (xyz_out) {
        xyz_out[3 * (size_t)dst] = x4.x;
        xyz_out[3 * (size_t)dst + 1] = x4.y;
        xyz_out[3 * (size_t)dst + 2] = x4.z;
    }
    if constexpr (kBatch) sid_out[dst] = scene;
    // the cell's last slot counts a cell above the candidate cap
    if (full) {
        const uint32_t cap = p.max_par_in_cell;
        if ((k + 1 == n || sorted_key[k + 1] != c) && k >= cap && sorted_key[k - cap] == c)
            atomicAdd(&counters[kCntCapCells], 1u);
    } else {
        const uint32_t end = new_end[c];
        if (dst + 1 == end && end - new_start[c] > p.max_par_in_cell) atomicAdd(&counters[kCntCapCells], 1u);
    }
}

// --------------------------------------------------- candidate positions
// The gathers read candidate positions through a source: straight from the
// sorted array, or from a block's LDS stage.  Either way the walk visits
// the same candidates in the same order, so results are identical.
//
// Stage: the particles of a block (consecutive sorted slots) have cell keys
// in [kmin, kmax].  For the neighbour-row offset q = 3 (dz + 1) + (dy + 1),
// every neighbour cell of every particle lies in the key range
// [kmin - 1 + off_q, kmax + 1 + off_q] (off_q = dy gx + dz gx gy), whose
// particles are one contiguous run of sorted slots.  The block finds the 9
// runs from the cell tables, copies their positions to LDS with coalesced
// loads, and a candidate j of row q is then read at LDS index j + s_off[q].
// Blocks whose key span, staged size or scene mix exceed the stage fall
// back to global loads (a block-uniform decision).
struct PosGlobal {
    const float4* __restrict__ p;
    __device__ __forceinline__ int32_t offset(int) const { return 0; }
    // a 32-bit byte offset (slots < 2^27): one base register plus the offset
    __device__ __forceinline__ float4 at(uint32_t j, int32_t) const {
        return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(p) + (j << 4));
    }
};

// The packed array (three floats a slot), read four candidates at a time
// (packed_group); the walk's own positions still come from the float4 array.
struct PosPacked {
    const float* __restrict__ xyz;
    __device__ __forceinline__ int32_t offset(int) const { return 0; }
};

struct PosStaged {
    const float4* s;
    const int32_t* s_off;
    __device__ __forceinline__ int32_t offset(int q) const { return s_off[q]; }
    __device__ __forceinline__ float4 at(uint32_t j, int32_t off) const { return s[(int32_t)j + off]; }
};

template <int kCap>
struct Stage {
    float4 pos[kCap];
    int32_t off[9];
    uint32_t lo[9], hi[9];
};

// Largest key span (kmax - kmin) a block stages; wider blocks use global loads.
constexpr uint32_t kStageMaxSpan = 256;

template <bool kBatch, int kCap>
__device__ bool stage_block(Stage<kCap>& S, uint32_t b0, uint32_t count, const sph_params& one,
                            const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid,
                            const uint32_t* __restrict__ skey, const float4* __restrict__ pos,
                            const uint32_t* __restrict__ cell_start,
                            const uint32_t* __restrict__ cell_end) {
    const uint32_t t = threadIdx.x, nt = blockDim.x;
    const uint32_t last = b0 + count - 1;
    if constexpr (kBatch) {
        if (sid[b0] != sid[last]) return false;
    }
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, b0, scene);
    const uint32_t kmin = skey[b0], kmax = skey[last];
    if (kmax - kmin > kStageMaxSpan) return false;
    if (t < 9) {
        S.lo[t] = 0xFFFFFFFFu;
        S.hi[t] = 0u;
    }
    __syncthreads();
    const uint32_t cells = kmax - kmin + 3;
    const int64_t klo = p.cell_base, khi = (int64_t)p.cell_base + p.num_cells - 1;
    for (uint32_t c = t; c < 9 * cells; c += nt) {
        const uint32_t q = c / cells, x = c - q * cells;
        const int64_t dz = (int64_t)(q / 3) - 1, dy = (int64_t)(q % 3) - 1;
        const int64_t key = (int64_t)kmin - 1 + x + dy * p.grid_size[0] + dz * p.grid_size_yx;
        if (key < klo || key > khi) continue;
        const uint32_t s = cell_start[key];
        if (s == kEmptyCell) continue;
        atomicMin(&S.lo[q], s);
        atomicMax(&S.hi[q], cell_end[key]);
    }
    __syncthreads();
    uint32_t base[10];
    base[0] = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const uint32_t lo = S.lo[q], hi = S.hi[q];
        base[q + 1] = base[q] + (hi > lo ? hi - lo : 0u);
    }
    if (base[9] > (uint32_t)kCap) return false;
    for (uint32_t idx = t; idx < base[9]; idx += nt) {
        int q = 0;
#pragma unroll
        for (int k = 1; k < 9; ++k) q += idx >= base[k] ? 1 : 0;
        S.pos[idx] = pos[S.lo[q] + (idx - base[q])];
    }
    if (t < 9) S.off[t] = (int32_t)base[t] - (int32_t)S.lo[t];
    __syncthreads();
    return true;
}

// ------------------------------------------------------------------ density
// counters: the step's counters, never null (the walk reads kCntNaN).
//
// With kLists the density walk also records each particle's neighbours
// (the slots that pass r2 < h2, j != i, in walk order) in the group-major
// lists (ListRing), and their count in nnb[i]; a count above kNbrCap
// raises the overflow counter (entries past the cap are dropped).
//
// With kMasks it records instead, per neighbour cell in walk order, a 16-bit
// mask of the cell's slots (bit = slot - cell start) that pass the force
// predicate (r2 <= force_r2_max, j != i): the force kernel's pairs, so that
// it needs no distance tests of its own.  Cells hold at most 16 slots here.
// Masks of cells 2w and 2w+1 share word w (low, high half): kMaskWords
// words per slot, word w of slot i at nbr[w n + i].
template <bool kBatch, int kPrec, bool kLists = false, typename Src = PosGlobal, bool kCells = false,
          bool kMasks = false, bool kEntries = false>
__device__ __forceinline__ uint32_t density_one(uint32_t i, const sph_params& one,
                                            const sph_params* __restrict__ tab,
                                            const uint32_t* __restrict__ sid,
                                            const float4* __restrict__ pos,
                                            const uint32_t* __restrict__ cell_start,
                                            const uint32_t* __restrict__ cell_end,
                                            float2* __restrict__ rhop,
                                            uint32_t n = 0, uint32_t* __restrict__ nbr = nullptr,
                                            uint32_t* __restrict__ nnb = nullptr,
                                            uint32_t* __restrict__ counters = nullptr,
                                            const Src* src_in = nullptr, const CellView* cv = nullptr,
                                            float4* __restrict__ rec = nullptr,
                                            const float4* __restrict__ vel = nullptr) {
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, i, scene);
    const float4 xi = pos[i];
    Src src{};
    if constexpr (std::is_same<Src, PosGlobal>::value) {
        src.p = pos;
    } else {
        src = *src_in;
    }
    CellWalk w;
    w.init(p, xi.x, xi.y, xi.z, walk_culled(counters));
    const float h2 = p.h2;
    float sum = 0.0f;
    KahanF ks{0.0f, 0.0f};
    __shared__ uint32_t ring[kLists ? kRingSlots * kRingLanes : 1];
    ListRing lw;
    if constexpr (kLists) lw.init(ring);
    // kEntries: the pair entries pass through a four-entry ring per lane
    __shared__ uint32_t ering[kEntries ? kPairRing * kBlock : 1];
    uint32_t* const eslot = ering + threadIdx.x;
    uint32_t ecount = 0, eflushed = 0, epairs = 0;
    uint4* const entries = reinterpret_cast<uint4*>(nbr);
    uint4* const lists = reinterpret_cast<uint4*>(nbr);
    int xbase = 0;
    if constexpr (kCells) xbase = (int)(w.own_key() - cv->kmin) + 1;
    const float r2f = p.force_r2_max;
    // Position arrays in global memory are padded past their end (dalloc):
    // a group's loads past the cell's end stay in bounds and are masked out,
    // so they need no clamp.  LDS-staged positions keep it.
    constexpr bool kPacked = std::is_same<Src, PosPacked>::value;
    constexpr bool kUnclamped = std::is_same<Src, PosGlobal>::value || kPacked;
    uint32_t mword = 0;
    // the 27 cells in (dz, dy, dx) order; a row's grid test, distance and
    // key are computed once for its three cells
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy) {
            const CellWalk::Row row = w.row(dy, dz);
#pragma unroll 1
            for (int dx = -1; dx <= 1; ++dx) {
                const int ord = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1);
                uint32_t s, e;
                bool open;
                if constexpr (kCells)
                    open = w.open_staged(dx, dy, dz, cv->t, cv->cells, xbase, s, e);
                else
                    open = w.open_in(row, dx, cell_start, cell_end, s, e);
                uint32_t cmask = 0;
                if (open) {
                    float cell = 0.0f;
                    const int32_t soff = src.offset(ord / 3);
                    uint32_t fresh = kPairNewCell;  // the cell's first entry opens its partial sum
                    // candidates four at a time: the loads go out together (indices
                    // past the cell's end are masked out of the sums); the particle
                    // itself can only be a candidate of its own cell (ord 13), which
                    // takes the walk with the extra test
                    auto walk = [&](auto self_tag) {
                        constexpr bool kSelf = decltype(self_tag)::value;
                        const float* xqp = nullptr;  // kPacked: slot j0's x
                        if constexpr (kPacked) xqp = src.xyz + 3 * (size_t)s;
                        for (uint32_t j0 = s; j0 < e; j0 += 4) {
                            float cx[4], cy[4], cz[4];
                            if constexpr (kPacked) {
                                packed_group(xqp, cx, cy, cz);
                                xqp += 12;
                            } else {
#pragma unroll
                                for (int k = 0; k < 4; ++k) {
                                    const float4 xq = src.at(kUnclamped ? j0 + k : min(j0 + k, e - 1), soff);
                                    cx[k] = xq.x, cy[k] = xq.y, cz[k] = xq.z;
                                }
                            }
                            const uint32_t left = e - j0;
                            const uint32_t self = kSelf ? i - j0 : 0u;
                            uint32_t fbits = 0;
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                const uint32_t j = j0 + k;
                                const bool valid = left > (uint32_t)k && (!kSelf || self != (uint32_t)k);
                                const float rx = xi.x - cx[k], ry = xi.y - cy[k], rz = xi.z - cz[k];
                                const float r2 = rx * rx + ry * ry + rz * rz;
                                const bool hit = (r2 < h2) && valid;
                                const float c3 = h2 - r2;
                                const float wgt = (c3 * c3) * c3;
                                if (kPrec == SPH_PREC_KAHAN) {
                                    if (hit) ks.add(wgt);
                                } else {
                                    // the partial sum is >= +0, so adding +0 for a miss is exact
                                    cell = cell + (hit ? wgt : 0.0f);
                                }
                                if constexpr (kLists) lw.put(j, hit);
                                if constexpr (kMasks || kEntries)
                                    fbits |= (force_pair(r2, r2f) && valid ? 1u : 0u) << k;
                            }
                            if constexpr (kMasks) cmask |= fbits << (j0 - s);
                            if constexpr (kEntries) {
                                // the written slot is only kept when the count advances
                                eslot[(ecount & (kPairRing - 1u)) * kBlock] = (j0 << kPairSlotShift) | fresh | fbits;
                                if (fbits) {
                                    ++ecount;
                                    epairs += (uint32_t)__popc(fbits);
                                    fresh = 0;
                                }
                                if (ecount - eflushed == kPairRing) {
                                    entries[(size_t)(eflushed >> 2) * n + i] =
                                        make_uint4(eslot[0], eslot[kBlock], eslot[2 * kBlock], eslot[3 * kBlock]);
                                    eflushed += kPairRing;
                                }
                            }
                            if constexpr (kLists) lw.drain(lists, n, i);
                        }
                    };
                    if (ord == 13)
                        walk(std::true_type{});
                    else
                        walk(std::false_type{});
                    if (kPrec != SPH_PREC_KAHAN) sum = sum + cell;
                }
                if constexpr (kMasks) {
                    if (ord & 1)
                        nbr[(uint32_t)(ord >> 1) * n + i] = mword | (cmask << 16);
                    else
                        mword = cmask;
                }
            }
        }
    }
    if constexpr (kMasks) nbr[(kMaskWords - 1u) * n + i] = mword;
    if (kPrec == SPH_PREC_KAHAN) sum = ks.s;
    const float rho = (sum * p.poly6_kern) * p.particle_mass;
    const float pres = eos(p, rho);
    rhop[i] = make_float2(rho, pres);
    if constexpr (kEntries) {
        if (ecount > eflushed)
            entries[(size_t)(eflushed >> 2) * n + i] =
                make_uint4(eslot[0], eslot[kBlock], eslot[2 * kBlock], eslot[3 * kBlock]);
        nnb[i] = ecount | (epairs << 16);
        const float4 vi = vel[i];
        rec[2 * (size_t)i] = make_float4(xi.x, xi.y, xi.z, rho);
        rec[2 * (size_t)i + 1] = make_float4(vi.x, vi.y, vi.z, pres);
    }
    if constexpr (kLists) {
        lw.finish(lists, n, i);
        nnb[i] = lw.count;
        if (lw.count > kNbrCap) atomicAdd(&counters[kCntNbrOverflow], 1u);
        return lw.count;
    }
    return 0u;
}


template <bool kBatch, int kPrec>
__global__ __launch_bounds__(kBlock) void k_density(sph_params one, const sph_params* __restrict__ tab,
                                                    const uint32_t* __restrict__ sid, uint32_t n,
                                                    const float4* __restrict__ pos,
                                                    const uint32_t* __restrict__ cell_start,
                                                    const uint32_t* __restrict__ cell_end,
                                                    float2* __restrict__ rhop, uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    density_one<kBatch, kPrec>(i, one, tab, sid, pos, cell_start, cell_end, rhop, n, nullptr, nullptr, counters);
}

// Persistent form: a fixed grid pulls 256-slot tiles from a ticket counter.
// Every block ends on exactly one ticket past the last tile; the last block
// to finish resets both counters, so consecutive launches (and graph
// replays) start from zero without a separate clear.
__device__ __forceinline__ uint32_t next_tile(uint32_t* work, uint32_t* s_tile) {
    if (threadIdx.x == 0) *s_tile = atomicAdd(&work[0], 1u);
    __syncthreads();
    const uint32_t t = *s_tile;
    __syncthreads();
    return t;
}

__device__ __forceinline__ void retire_block(uint32_t* work) {
    if (threadIdx.x == 0) {
        const uint32_t d = atomicAdd(&work[1], 1u);
        if (d == gridDim.x - 1) {
            atomicExch(&work[0], 0u);
            atomicExch(&work[1], 0u);
        }
    }
}

template <bool kBatch, int kPrec>
__global__ __launch_bounds__(kBlock) void k_density_persistent(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float2* __restrict__ rhop, uint32_t* __restrict__ work,
    uint32_t* __restrict__ counters) {
    __shared__ uint32_t s_tile;
    const uint32_t ntiles = (n + kBlock - 1) / kBlock;
    for (;;) {
        const uint32_t tile = next_tile(work, &s_tile);
        if (tile >= ntiles) break;
        const uint32_t i = tile * kBlock + threadIdx.x;
        if (i < n)
            density_one<kBatch, kPrec>(i, one, tab, sid, pos, cell_start, cell_end, rhop, n, nullptr, nullptr,
                                       counters);
    }
    retire_block(work);
}

// -------------------------------------------------------------------- force
// Candidate tests and pair terms are split so that the expensive pair body
// runs with the lanes of a wave mostly active (with ~15% of candidates
// passing, a fused loop would run the body at ~15% lane occupancy):
//   test    - walk the candidates in reference order and append every slot
//             j that passes the force predicate to the lane's LDS list
//             (entry = j | neighbour-cell ordinal << 27);
//   compute - evaluate the pair terms of the listed neighbours in list
//             order.  The per-cell partial sum is flushed into the total
//             when the cell ordinal changes, which reproduces the
//             reference's per-cell sums (cells without a passing neighbour
//             contribute +0 and are not visited).
// The list is drained by the whole wave before a cell whose candidates
// could overflow some lane's list.  The list size (entries per lane) trades
// LDS per wave (occupancy) against drains at partial lane occupancy.
constexpr int kForceListDefault = 24;
// The pair-mask path appends no candidate tests, only passing pairs: a
// shorter list (more blocks per CU) is a little faster there.
constexpr int kForceListMasks = 16;
constexpr uint32_t kSlotBits = 27;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;

struct ForceAcc {
    float ax, ay, az;  // total
    float fx, fy, fz;  // current cell's partial sum
    KahanF kx, ky, kz;
    uint32_t prev_ord;
};

template <int kPrec, int kB, typename Src>
__device__ __forceinline__ void force_drain(const sph_params& p, const uint32_t* list, int t, int cnt,
                                            const float4& xi, const float4& vi, const float2& ri,
                                            const Src& src, const float4* __restrict__ vel,
                                            const float2* __restrict__ rhop, ForceAcc& a) {
    const float h = p.h, spiky = p.spiky_kern, min_dist = p.min_dist, min_dens = p.min_dens;
    const float vterm = p.lap_kern * p.viscosity;
    for (int k = 0; k < cnt; ++k) {
        const uint32_t ent = list[k * kB + t];
        const uint32_t jj = ent & kSlotMask;
        const uint32_t o = ent >> kSlotBits;
        if (kPrec != SPH_PREC_KAHAN && o != a.prev_ord) {
            a.ax = a.ax + a.fx;
            a.ay = a.ay + a.fy;
            a.az = a.az + a.fz;
            a.fx = a.fy = a.fz = 0.0f;
            a.prev_ord = o;
        }
        const float4 xj = src.at(jj, src.offset((int)(o / 3)));
        const float4 vj = vel[jj];
        const float2 rj = rhop[jj];
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = fmaxf(min_dist, sqrt_ieee(r2));
        const float d12 = fminf(min_dens, 1.0f / (ri.x * rj.x));
        const float cc = h - r;
        const float pterm = ((cc * spiky) * (ri.y + rj.y)) / r;
        const float cd = cc * d12;
        const float gxv = (pterm * rx + vterm * (vj.x - vi.x)) * cd;
        const float gyv = (pterm * ry + vterm * (vj.y - vi.y)) * cd;
        const float gzv = (pterm * rz + vterm * (vj.z - vi.z)) * cd;
        if (kPrec == SPH_PREC_KAHAN) {
            a.kx.add(gxv);
            a.ky.add(gyv);
            a.kz.add(gzv);
        } else {
            a.fx = a.fx + gxv;
            a.fy = a.fy + gyv;
            a.fz = a.fz + gzv;
        }
    }
}

// One slot per lane; every lane of the wave must call this (inactive lanes
// with active = false) because the list drains are wave-wide decisions.
//
// With kMasks the pair masks of the density kernel (density_one) replace the
// candidate tests: each neighbour cell's passing slots go straight to the
// lane's list.
template <bool kBatch, int kPrec, int kList, int kB = kBlock, typename Src = PosGlobal, bool kCells = false,
          bool kMasks = false>
__device__ __forceinline__ void force_one(uint32_t i, bool active, uint32_t* list,
                                          const sph_params& one, const sph_params* __restrict__ tab,
                                          const uint32_t* __restrict__ sid,
                                          const float4* __restrict__ pos,
                                          const float4* __restrict__ vel,
                                          const float2* __restrict__ rhop,
                                          const uint32_t* __restrict__ cell_start,
                                          const uint32_t* __restrict__ cell_end,
                                          float4* __restrict__ vel_out, float4* __restrict__ dvel,
                                          const float* __restrict__ dt_dev,
                                          const uint32_t* __restrict__ counters, const Src* src_in = nullptr,
                                          const CellView* cv = nullptr, const uint32_t* __restrict__ masks = nullptr,
                                          uint32_t n = 0) {
    const int t = threadIdx.x;
    Src src{};
    if constexpr (std::is_same<Src, PosGlobal>::value) {
        src.p = pos;
    } else {
        src = *src_in;
    }
    const uint32_t ii = active ? i : 0u;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, ii, scene);
    const float4 xi = pos[ii];
    const float4 vi = vel[ii];
    const float2 ri = rhop[ii];
    CellWalk w;
    w.init(p, xi.x, xi.y, xi.z, walk_culled(counters));
    const float r2max = p.force_r2_max;

    ForceAcc a{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, {0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f},
               0xFFFFFFFFu};
    int cnt = 0;
    int xbase = 0;
    if constexpr (kCells) xbase = (int)(w.own_key() - cv->kmin) + 1;
    uint32_t mword = 0;
    // the 27 cells in (dz, dy, dx) order; a row's grid test, distance and
    // key are computed once for its three cells
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy) {
            const CellWalk::Row row = w.row(dy, dz);
#pragma unroll 1
            for (int dx = -1; dx <= 1; ++dx) {
                const int ord = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1);
                uint32_t s = 0, e = 0;
                bool open;
                if constexpr (kCells)
                    open = active && w.open_staged(dx, dy, dz, cv->t, cv->cells, xbase, s, e);
                else
                    open = active && w.open_in(row, dx, cell_start, cell_end, s, e);
                if constexpr (kMasks) {
                    if ((ord & 1) == 0) mword = masks[(uint32_t)(ord >> 1) * n + ii];
                    uint32_t m = open ? ((ord & 1) ? (mword >> 16) : (mword & 0xFFFFu)) : 0u;
                    if (__any(cnt + __popc(m) > kList)) {
                        force_drain<kPrec, kB>(p, list, t, cnt, xi, vi, ri, src, vel, rhop, a);
                        cnt = 0;
                    }
                    const uint32_t tag = (uint32_t)ord << kSlotBits;
                    while (m != 0u) {
                        list[cnt * kB + t] = (s + (uint32_t)__builtin_ctz(m)) | tag;
                        ++cnt;
                        m &= m - 1u;
                    }
                    continue;
                }
                // every candidate of this cell could pass: make room first
                if (__any(cnt + (int)(e - s) > kList)) {
                    force_drain<kPrec, kB>(p, list, t, cnt, xi, vi, ri, src, vel, rhop, a);
                    cnt = 0;
                }
                if (!open) continue;
                const uint32_t tag = (uint32_t)ord << kSlotBits;
                const int32_t soff = src.offset(ord / 3);
                // candidates four at a time (see density_one)
                for (uint32_t j0 = s; j0 < e; j0 += 4) {
                    float4 xq[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) xq[k] = src.at(min(j0 + k, e - 1), soff);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const uint32_t j = j0 + k;
                        const float4 xj = xq[k];
                        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
                        const float r2 = rx * rx + ry * ry + rz * rz;
                        // max(minDist, sqrtf(r2)) < h  <=>  r2 <= force_r2_max.  The
                        // store is unconditional, also for the repeated slots past the
                        // cell's end (cnt <= list size holds for the whole cell, and the
                        // list has one spare row for a full list); only a hit advances cnt.
                        list[cnt * kB + t] = j | tag;
                        cnt += (force_pair(r2, r2max) && j != i && j < e) ? 1 : 0;
                    }
                }
            }
        }
    }
    force_drain<kPrec, kB>(p, list, t, cnt, xi, vi, ri, src, vel, rhop, a);
    if (!active) return;
    float ax, ay, az;
    if (kPrec == SPH_PREC_KAHAN) {
        ax = a.kx.s;
        ay = a.ky.s;
        az = a.kz.s;
    } else {
        ax = a.ax + a.fx;
        ay = a.ay + a.fy;
        az = a.az + a.fz;
    }
    const float mdt = p.particle_mass * dt_dev[scene];
    ax = ax * mdt;
    ay = ay * mdt;
    az = az * mdt;
    if (p.physics) {
        // the pair increment only: k_extra_terms adds the scene's other
        // terms and finishes the update
        vel_out[i] = make_float4(ax, ay, az, 0.0f);
        return;
    }
    vel_out[i] = make_float4(vi.x + ax, vi.y + ay, vi.z + az, vi.w + 0.0f);
    if (dvel) dvel[i] = make_float4(ax, ay, az, 0.0f);
}

template <bool kBatch, int kPrec, int kList>
__global__ __launch_bounds__(kBlock) void k_force(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float2* __restrict__ rhop, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float4* __restrict__ vel_out,
    float4* __restrict__ dvel, const float* __restrict__ dt_dev,
    const uint32_t* __restrict__ counters) {
    __shared__ uint32_t list[(kList + 1) * kBlock];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    force_one<kBatch, kPrec, kList>(i, i < n, list, one, tab, sid, pos, vel, rhop, cell_start,
                                    cell_end, vel_out, dvel, dt_dev, counters);
}

template <bool kBatch, int kPrec, int kList>
__global__ __launch_bounds__(kBlock) void k_force_persistent(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float2* __restrict__ rhop, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float4* __restrict__ vel_out,
    float4* __restrict__ dvel, const float* __restrict__ dt_dev,
    const uint32_t* __restrict__ counters, uint32_t* __restrict__ work) {
    __shared__ uint32_t list[(kList + 1) * kBlock];
    __shared__ uint32_t s_tile;
    const uint32_t ntiles = (n + kBlock - 1) / kBlock;
    for (;;) {
        const uint32_t tile = next_tile(work, &s_tile);
        if (tile >= ntiles) break;
        const uint32_t i = tile * kBlock + threadIdx.x;
        force_one<kBatch, kPrec, kList>(i, i < n, list, one, tab, sid, pos, vel, rhop, cell_start,
                                        cell_end, vel_out, dvel, dt_dev, counters);
    }
    retire_block(work);
}

// ------------------------------------------------------ staged gathers
constexpr int kDensityStage = 2816;  // positions staged per 256-slot tile (44 KB)
constexpr int kForceTile = 128;      // slots per force tile
constexpr int kForceStage = 1536;    // positions staged per force tile (24 KB)

// kOrder: 0 = tiles in launch order, 1 = XCD-ordered; kStaged: candidate
// positions staged in LDS.
template <bool kBatch, int kPrec, bool kLists, int kOrder, bool kStaged>
__global__ __launch_bounds__(kBlock) void k_density_tiled(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const uint32_t* __restrict__ skey,
    const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
    float2* __restrict__ rhop, uint32_t* __restrict__ nbr, uint32_t* __restrict__ nnb,
    uint32_t* __restrict__ counters) {
    __shared__ Stage<kStaged ? kDensityStage : 1> S;
    const uint32_t b0 = (kOrder ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x) * kBlock;
    const uint32_t count = min((uint32_t)kBlock, n - b0);
    bool staged = false;
    if constexpr (kStaged)
        staged = stage_block<kBatch>(S, b0, count, one, tab, sid, skey, pos, cell_start, cell_end);
    const uint32_t i = b0 + threadIdx.x;
    uint32_t listed = 0;
    if (i < n) {
        uint32_t c;
        if (kStaged && staged) {
            const PosStaged src{S.pos, S.off};
            c = density_one<kBatch, kPrec, kLists, PosStaged>(i, one, tab, sid, pos, cell_start, cell_end,
                                                              rhop, n, nbr, nnb, counters, &src);
        } else {
            c = density_one<kBatch, kPrec, kLists>(i, one, tab, sid, pos, cell_start, cell_end, rhop, n, nbr,
                                                   nnb, counters);
        }
        listed = min(c, kNbrCap);
    }
    if constexpr (kLists) {
        for (int off = 32; off > 0; off >>= 1) listed += __shfl_xor(listed, off);
        if ((threadIdx.x & 63) == 0 && listed) atomicAdd(&counters[kCntPairs], listed);
    }
}

// Density plus the force kernel's pair masks (density_one with kMasks), one
// slot per lane in XCD-ordered tiles.
template <bool kBatch, int kPrec>
__global__ __launch_bounds__(kBlock) void k_density_masks(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float2* __restrict__ rhop, uint32_t* __restrict__ masks,
    uint32_t* __restrict__ counters) {
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    density_one<kBatch, kPrec, false, PosGlobal, false, true>(i, one, tab, sid, pos, cell_start, cell_end, rhop,
                                                             n, masks, nullptr, counters);
}

// Density plus the force kernel's pair entries and force records
// (density_one with kEntries), one slot per lane in XCD-ordered tiles.
template <bool kBatch, int kPrec>
__global__ __launch_bounds__(kBlock) void k_density_entries(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const float* __restrict__ xyz, const float4* __restrict__ vel,
    const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end, float2* __restrict__ rhop,
    uint4* __restrict__ entries, uint32_t* __restrict__ nnb, float4* __restrict__ rec,
    uint32_t* __restrict__ counters) {
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const PosPacked src{xyz};
    density_one<kBatch, kPrec, false, PosPacked, false, false, true>(
        i, one, tab, sid, pos, cell_start, cell_end, rhop, n, reinterpret_cast<uint32_t*>(entries), nnb, counters,
        &src, nullptr, rec, vel);
}

// Cell-table stage: the (start, end) pairs of the block's 9 neighbour rows
// (the key ranges of stage_block), so that opening a neighbour cell reads
// LDS instead of two dependent global loads.
constexpr uint32_t kCellSpanMax = 96;
constexpr int kCellRow = (int)kCellSpanMax + 3;

struct CellStage {
    uint2 t[9 * kCellRow];
};

template <bool kBatch>
__device__ bool stage_cells(CellStage& S, uint32_t b0, uint32_t count, const sph_params& one,
                            const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid,
                            const uint32_t* __restrict__ skey, const uint32_t* __restrict__ cell_start,
                            const uint32_t* __restrict__ cell_end, CellView& cv) {
    const uint32_t t = threadIdx.x, nt = blockDim.x;
    const uint32_t last = b0 + count - 1;
    if constexpr (kBatch) {
        if (sid[b0] != sid[last]) return false;
    }
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, b0, scene);
    const uint32_t kmin = skey[b0], kmax = skey[last];
    if (kmax - kmin > kCellSpanMax) return false;
    const uint32_t cells = kmax - kmin + 3;
    const int64_t klo = p.cell_base, khi = (int64_t)p.cell_base + p.num_cells - 1;
    for (uint32_t c = t; c < 9u * cells; c += nt) {
        const uint32_t q = c / cells, x = c - q * cells;
        const int64_t dz = (int64_t)(q / 3) - 1, dy = (int64_t)(q % 3) - 1;
        const int64_t key = (int64_t)kmin - 1 + x + dy * p.grid_size[0] + dz * p.grid_size_yx;
        uint2 v = make_uint2(kEmptyCell, 0u);
        if (key >= klo && key <= khi) {
            v.x = cell_start[key];
            v.y = cell_end[key];
        }
        S.t[c] = v;
    }
    __syncthreads();
    cv.t = S.t;
    cv.cells = (int)cells;
    cv.kmin = kmin;
    return true;
}

template <bool kBatch, int kPrec, bool kLists>
__global__ __launch_bounds__(kBlock) void k_density_cells(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const uint32_t* __restrict__ skey,
    const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
    float2* __restrict__ rhop, uint32_t* __restrict__ nbr, uint32_t* __restrict__ nnb,
    uint32_t* __restrict__ counters) {
    __shared__ CellStage S;
    const uint32_t b0 = xcd_tile(blockIdx.x, gridDim.x) * kBlock;
    const uint32_t count = min((uint32_t)kBlock, n - b0);
    CellView cv{};
    const bool staged = stage_cells<kBatch>(S, b0, count, one, tab, sid, skey, cell_start, cell_end, cv);
    const uint32_t i = b0 + threadIdx.x;
    uint32_t listed = 0;
    if (i < n) {
        uint32_t c;
        if (staged)
            c = density_one<kBatch, kPrec, kLists, PosGlobal, true>(i, one, tab, sid, pos, cell_start, cell_end,
                                                                    rhop, n, nbr, nnb, counters, nullptr, &cv);
        else
            c = density_one<kBatch, kPrec, kLists>(i, one, tab, sid, pos, cell_start, cell_end, rhop, n, nbr,
                                                   nnb, counters);
        listed = min(c, kNbrCap);
    }
    if constexpr (kLists) {
        for (int off = 32; off > 0; off >>= 1) listed += __shfl_xor(listed, off);
        if ((threadIdx.x & 63) == 0 && listed) atomicAdd(&counters[kCntPairs], listed);
    }
}

template <bool kBatch, int kPrec, int kList>
__global__ __launch_bounds__(kBlock) void k_force_cells(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const uint32_t* __restrict__ skey, const float4* __restrict__ vel,
    const float2* __restrict__ rhop, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float4* __restrict__ vel_out,
    float4* __restrict__ dvel, const float* __restrict__ dt_dev,
    const uint32_t* __restrict__ counters) {
    __shared__ uint32_t list[(kList + 1) * kBlock];
    __shared__ CellStage S;
    const uint32_t b0 = xcd_tile(blockIdx.x, gridDim.x) * kBlock;
    const uint32_t count = min((uint32_t)kBlock, n - b0);
    CellView cv{};
    const bool staged = stage_cells<kBatch>(S, b0, count, one, tab, sid, skey, cell_start, cell_end, cv);
    const uint32_t i = b0 + threadIdx.x;
    if (staged)
        force_one<kBatch, kPrec, kList, kBlock, PosGlobal, true>(i, i < n, list, one, tab, sid, pos, vel, rhop,
                                                                 cell_start, cell_end, vel_out, dvel, dt_dev, counters,
                                                                 nullptr, &cv);
    else
        force_one<kBatch, kPrec, kList>(i, i < n, list, one, tab, sid, pos, vel, rhop, cell_start, cell_end,
                                        vel_out, dvel, dt_dev, counters);
}

template <bool kBatch, int kPrec, int kList, int kOrder, bool kMasks = false>
__global__ __launch_bounds__(kBlock) void k_force_ordered(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float2* __restrict__ rhop, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float4* __restrict__ vel_out,
    float4* __restrict__ dvel, const float* __restrict__ dt_dev,
    const uint32_t* __restrict__ counters, const uint32_t* __restrict__ masks = nullptr) {
    __shared__ uint32_t list[(kList + 1) * kBlock];
    const uint32_t i = (kOrder ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x) * kBlock + threadIdx.x;
    if constexpr (kMasks)
        force_one<kBatch, kPrec, kList, kBlock, PosGlobal, false, true>(
            i, i < n, list, one, tab, sid, pos, vel, rhop, cell_start, cell_end, vel_out, dvel, dt_dev, counters, nullptr,
            nullptr, masks, n);
    else
        force_one<kBatch, kPrec, kList>(i, i < n, list, one, tab, sid, pos, vel, rhop, cell_start, cell_end,
                                        vel_out, dvel, dt_dev, counters);
}

template <bool kBatch, int kPrec, int kList>
__global__ __launch_bounds__(kForceTile) void k_force_tiled(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const uint32_t* __restrict__ skey, const float4* __restrict__ vel,
    const float2* __restrict__ rhop, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float4* __restrict__ vel_out,
    float4* __restrict__ dvel, const float* __restrict__ dt_dev,
    const uint32_t* __restrict__ counters) {
    __shared__ uint32_t list[(kList + 1) * kForceTile];
    __shared__ Stage<kForceStage> S;
    const uint32_t b0 = xcd_tile(blockIdx.x, gridDim.x) * kForceTile;
    const uint32_t count = min((uint32_t)kForceTile, n - b0);
    const bool staged = stage_block<kBatch>(S, b0, count, one, tab, sid, skey, pos, cell_start, cell_end);
    const uint32_t i = b0 + threadIdx.x;
    if (staged) {
        const PosStaged src{S.pos, S.off};
        force_one<kBatch, kPrec, kList, kForceTile, PosStaged>(i, i < n, list, one, tab, sid, pos, vel, rhop,
                                                               cell_start, cell_end, vel_out, dvel, dt_dev, counters,
                                                               &src);
    } else {
        force_one<kBatch, kPrec, kList, kForceTile>(i, i < n, list, one, tab, sid, pos, vel, rhop, cell_start,
                                                    cell_end, vel_out, dvel, dt_dev, counters);
    }
}

// Force over the density kernel's pair entries (k_density_entries, or
// k_density_fast in FAST mode), one slot per lane: each lane walks only its
// own pairs, in the reference's order (cells in walk order, slots
// ascending), with the reference's pair arithmetic.  In the reference order
// (kPrec 0) with per-cell partial sums (a new-cell entry adds the open
// partial sum to the total first), so the result equals force_one's bit for
// bit; Kahan: compensated sums; FAST: the entries carry no new-cell flags,
// so one running sum.  Four pairs per iteration: they are decoded first, so
// that their eight record loads go out together.  A record is (x, y, z, rho)
// and (vx, vy, vz, p), one aligned 32-byte block.
constexpr int kForcePairs = 4;  // pairs decoded and loaded per iteration

template <bool kBatch, int kPrec>
__global__ __launch_bounds__(kBlock) void k_force_entries(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ vel, const float4* __restrict__ rec, const uint4* __restrict__ entries,
    const uint32_t* __restrict__ nnb, float4* __restrict__ vel_out, float4* __restrict__ dvel,
    const float* __restrict__ dt_dev) {
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, i, scene);
    const float4 xi = rec[2 * (size_t)i];       // (x, y, z, rho)
    const float4 vi = rec[2 * (size_t)i + 1];   // (vx, vy, vz, p)
    const float h = p.h, spiky = p.spiky_kern, min_dist = p.min_dist, min_dens = p.min_dens;
    const float vterm = p.lap_kern * p.viscosity;
    const uint32_t npairs = nnb[i] >> 16;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;  // total
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;  // the current cell's partial sum
    KahanF kx{0.0f, 0.0f}, ky{0.0f, 0.0f}, kz{0.0f, 0.0f};
    uint4 buf = make_uint4(0, 0, 0, 0);
    uint32_t bits = 0, jb = 0, next = 0, fresh = 0;
    // the next pair's slot, and whether it opens a cell
    auto next_pair = [&](bool& opens) {
        opens = false;
        if (bits == 0u) {
            if ((next & 3u) == 0u) buf = entries[(size_t)(next >> 2) * n + i];
            const uint32_t q = next & 3u;
            const uint32_t e = q == 0u ? buf.x : (q == 1u ? buf.y : (q == 2u ? buf.z : buf.w));
            ++next;
            bits = e & 15u;
            fresh = e & kPairNewCell;
            jb = e >> kPairSlotShift;
        }
        opens = fresh != 0u;
        fresh = 0u;
        const uint32_t k = (uint32_t)__builtin_ctz(bits);
        bits &= bits - 1u;
        return jb + k;
    };
    for (uint32_t t = 0; t < npairs; t += kForcePairs) {
        uint32_t j[kForcePairs];
        bool opens[kForcePairs];
        j[0] = next_pair(opens[0]);
#pragma unroll
        for (int k = 1; k < kForcePairs; ++k) {
            opens[k] = false;
            j[k] = t + (uint32_t)k < npairs ? next_pair(opens[k]) : j[0];
        }
        float4 xq[kForcePairs], vq[kForcePairs];
#pragma unroll
        for (int k = 0; k < kForcePairs; ++k) {
            xq[k] = rec[2 * (size_t)j[k]];
            vq[k] = rec[2 * (size_t)j[k] + 1];
        }
#pragma unroll
        for (int k = 0; k < kForcePairs; ++k) {
            // past the last pair (a repeat of the batch's first): nothing added
            const bool on = t + (uint32_t)k < npairs;
            const float4 xj = xq[k], vj = vq[k];
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            const float r = fmaxf(min_dist, sqrt_ieee(r2));
            const float d12 = fminf(min_dens, 1.0f / (xi.w * xj.w));
            const float cc = h - r;
            const float pterm = ((cc * spiky) * (vi.w + vj.w)) / r;
            const float cd = cc * d12;
            const float gxv = (pterm * rx + vterm * (vj.x - vi.x)) * cd;
            const float gyv = (pterm * ry + vterm * (vj.y - vi.y)) * cd;
            const float gzv = (pterm * rz + vterm * (vj.z - vi.z)) * cd;
            if (kPrec == SPH_PREC_KAHAN) {
                if (on) {
                    kx.add(gxv);
                    ky.add(gyv);
                    kz.add(gzv);
                }
            } else {
                if (on && opens[k]) {
                    ax = ax + fx;
                    ay = ay + fy;
                    az = az + fz;
                    fx = fy = fz = 0.0f;
                }
                if (on) {
                    fx = fx + gxv;
                    fy = fy + gyv;
                    fz = fz + gzv;
                }
            }
        }
    }
    if (kPrec == SPH_PREC_KAHAN) {
        ax = kx.s;
        ay = ky.s;
        az = kz.s;
    } else {
        ax = ax + fx;
        ay = ay + fy;
        az = az + fz;
    }
    const float mdt = p.particle_mass * dt_dev[scene];
    ax = ax * mdt;
    ay = ay * mdt;
    az = az * mdt;
    if (p.physics) {
        // the pair increment only: k_extra_terms adds the scene's other
        // terms and finishes the update
        vel_out[i] = make_float4(ax, ay, az, 0.0f);
        return;
    }
    const float4 v = vel[i];
    vel_out[i] = make_float4(v.x + ax, v.y + ay, v.z + az, v.w + 0.0f);
    if (dvel) dvel[i] = make_float4(ax, ay, az, 0.0f);
}

// Force kernel for max_par_in_cell above the list capacity: the pair body
// sits directly in the candidate loop (same order, same sums).
template <bool kBatch, int kPrec>
__global__ __launch_bounds__(kBlock) void k_force_fused(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ sid, uint32_t n,
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float2* __restrict__ rhop, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float4* __restrict__ vel_out,
    float4* __restrict__ dvel, const float* __restrict__ dt_dev,
    const uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t scene;
    const sph_params& p = scene_params<kBatch>(one, tab, sid, i, scene);
    const float4 xi = pos[i];
    const float4 vi = vel[i];
    const float2 ri = rhop[i];
    CellWalk w;
    w.init(p, xi.x, xi.y, xi.z, walk_culled(counters));
    const float h = p.h, spiky = p.spiky_kern, min_dist = p.min_dist, min_dens = p.min_dens;
    const float vterm = p.lap_kern * p.viscosity;
    const float r2max = p.force_r2_max;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    KahanF kx{0.0f, 0.0f}, ky{0.0f, 0.0f}, kz{0.0f, 0.0f};
    for (int ord = 0; ord < 27; ++ord) {
        uint32_t s, e;
        if (!w.open(ord % 3 - 1, (ord / 3) % 3 - 1, ord / 9 - 1, cell_start, cell_end, s, e))
            continue;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        for (uint32_t j = s; j < e; ++j) {
            if (j == i) continue;
            const float4 xj = pos[j];
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (force_pair(r2, r2max)) {
                const float4 vj = vel[j];
                const float2 rj = rhop[j];
                const float r = fmaxf(min_dist, sqrt_ieee(r2));
                const float d12 = fminf(min_dens, 1.0f / (ri.x * rj.x));
                const float cc = h - r;
                const float pterm = ((cc * spiky) * (ri.y + rj.y)) / r;
                const float cd = cc * d12;
                const float gxv = (pterm * rx + vterm * (vj.x - vi.x)) * cd;
                const float gyv = (pterm * ry + vterm * (vj.y - vi.y)) * cd;
                const float gzv = (pterm * rz + vterm * (vj.z - vi.z)) * cd;
                if (kPrec == SPH_PREC_KAHAN) {
                    kx.add(gxv);
                    ky.add(gyv);
                    kz.add(gzv);
                } else {
                    fx = fx + gxv;
                    fy = fy + gyv;
                    fz = fz + gzv;
                }
            }
        }
        if (kPrec != SPH_PREC_KAHAN) {
            ax = ax + fx;
            ay = ay + fy;
            az = az + fz;
        }
    }
    if (kPrec == SPH_PREC_KAHAN) {
        ax = kx.s;
        ay = ky.s;
        az = kz.s;
    }
    const float mdt = p.particle_mass * dt_dev[scene];
    ax = ax * mdt;
    ay = ay * mdt;
    az = az * mdt;
    if (p.physics) {
        // the pair increment only: k_extra_terms adds the scene's other
        // terms and finishes the update
        vel_out[i] = make_float4(ax, ay, az, 0.0f);
        return;
    }
    vel_out[i] = make_float4(vi.x + ax, vi.y + ay, vi.z + az, vi.w + 0.0f);
    if (dvel) dvel[i] = make_float4(ax, ay, az, 0.0f);
}

// -------------------------------------------------------------------- stats
constexpr int kStatBlock = 256;

__device__ __forceinline__ void block_reduce4(double v[4], double* sh) {
    // sh: 4 * kStatBlock doubles.  Sums for 0..2, max for 3; fixed order.
    const int t = threadIdx.x;
    for (int k = 0; k < 4; ++k) sh[k * kStatBlock + t] = v[k];
    __syncthreads();
    for (int w = kStatBlock / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 3; ++k) sh[k * kStatBlock + t] += sh[k * kStatBlock + t + w];
            sh[3 * kStatBlock + t] = fmax(sh[3 * kStatBlock + t], sh[3 * kStatBlock + t + w]);
        }
        __syncthreads();
    }
    for (int k = 0; k < 4; ++k) v[k] = sh[k * kStatBlock];
}

// grid (chunks, nscenes): block (b, s) reduces chunk b of scene s's slots.
template <bool kBatch>
__global__ __launch_bounds__(kStatBlock) void k_stats_partials(
    sph_params one, const sph_params* __restrict__ tab, const uint32_t* __restrict__ scene_off,
    const float4* __restrict__ pos, const float4* __restrict__ vel, double* __restrict__ partials) {
    __shared__ double sh[4 * kStatBlock];
    const uint32_t s = blockIdx.y;
    uint32_t unused;
    const sph_params& p = scene_params<kBatch>(one, tab, &s, 0, unused);
    const uint32_t lo = scene_off[s], hi = scene_off[s + 1];
    const double m = (double)p.particle_mass;
    const double g[3] = {(double)p.gravity[0], (double)p.gravity[1], (double)p.gravity[2]};
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = lo + blockIdx.x * kStatBlock + threadIdx.x; i < hi; i += gridDim.x * kStatBlock) {
        const float4 x = pos[i];
        const float4 v = vel[i];
        const double v2 = (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z;
        acc[0] += m;
        acc[1] += 0.5 * m * v2;
        acc[2] += -m * (g[0] * x.x + g[1] * x.y + g[2] * x.z);
        acc[3] = fmax(acc[3], v2);
    }
    block_reduce4(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partials[((size_t)s * gridDim.x + blockIdx.x) * kPNum + k] = acc[k];
}

// grid (nscenes): one block per scene.
template <bool kBatch>
__global__ __launch_bounds__(kStatBlock) void k_stats_finalize(
    sph_params one, const sph_params* __restrict__ tab, uint32_t chunks,
    const double* __restrict__ partials, double* __restrict__ stats, float* __restrict__ dt_dev,
    int advance_clock) {
    __shared__ double sh[4 * kStatBlock];
    const uint32_t s = blockIdx.x;
    uint32_t unused;
    const sph_params& p = scene_params<kBatch>(one, tab, &s, 0, unused);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < chunks; b += kStatBlock) {
        const double* q = partials + ((size_t)s * chunks + b) * kPNum;
        for (int k = 0; k < 3; ++k) acc[k] += q[k];
        acc[3] = fmax(acc[3], q[3]);
    }
    block_reduce4(acc, sh);
    if (threadIdx.x == 0) {
        double* st = stats + (size_t)s * kNumStats;
        const double vmax = sqrt(acc[3]);
        st[kPMass] = acc[0];
        st[kPKin] = acc[1];
        st[kPPot] = acc[2];
        st[kPVmax] = vmax;
        double dt = (double)p.time_step;
        if (p.cfl > 0.0f) {
            const double dcfl = (double)p.cfl * (double)p.h / ((double)p.sound_speed + vmax);
            if (dcfl < dt) dt = dcfl;
        }
        if (p.t_end > 0.0f) {
            const double rem = (double)p.t_end - st[kSClock];
            // a remainder below 1e-4 of a nominal step ends the run instead
            if (rem < dt) dt = rem > 1e-4 * (double)p.time_step ? rem : 0.0;
        }
        st[kSDt] = dt;
        dt_dev[s] = (float)dt;
        if (advance_clock) st[kSClock] += (double)(float)dt;
    }
}

__global__ __launch_bounds__(kBlock) void k_cell_stats(uint32_t num_cells,
                                                      const uint32_t* __restrict__ cell_start,
                                                      const uint32_t* __restrict__ cell_end,
                                                      uint32_t* __restrict__ max_count) {
    uint32_t m = 0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < num_cells;
         c += gridDim.x * blockDim.x) {
        const uint32_t s = cell_start[c];
        if (s != kEmptyCell) m = max(m, cell_end[c] - s);
    }
    if (m) atomicMax(max_count, m);
}

// ---------------------------------------------------------------- transfers
__global__ __launch_bounds__(kBlock) void k_scatter_to_original(uint32_t n,
                                                               const uint32_t* __restrict__ pid,
                                                               const float* __restrict__ src,
                                                               float* __restrict__ dst, int rec,
                                                               int stride) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t id = pid[k];
    for (int c = 0; c < rec; ++c) dst[(size_t)id * rec + c] = src[(size_t)k * stride + c];
}

__global__ __launch_bounds__(kBlock) void k_update_from_original(uint32_t n,
                                                                const uint32_t* __restrict__ pid,
                                                                const float4* __restrict__ staging,
                                                                uint32_t first, uint32_t count,
                                                                float4* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t id = pid[k];
    if (id >= first && id - first < count) dst[k] = staging[id - first];
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ dst, uint32_t v,
                                                    size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_iota_u32(uint32_t* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_scene_ids(uint32_t n, uint32_t nscenes,
                                                     const uint32_t* __restrict__ scene_off,
                                                     const uint32_t* __restrict__ pid,
                                                     uint32_t* __restrict__ sid) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t id = pid[k];
    uint32_t lo = 0, hi = nscenes;  // last s with scene_off[s] <= id
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (scene_off[mid] <= id)
            lo = mid;
        else
            hi = mid;
    }
    sid[k] = lo;
}

inline uint32_t blocks_for(uint64_t n, uint32_t block) {
    const uint32_t b = (uint32_t)((n + block - 1) / block);
    return b == 0 ? 1 : b;
}

// ------------------------------------------------------------- dispatchers
template <bool kBatch>
hipError_t resort_apply_impl(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, int parity,
                             const RadixSortPlan& sort, hipStream_t s) {
    // the full sort runs only when the cell scan flagged the step
    uint32_t *skey = nullptr, *oslot = nullptr;
    hipError_t e = radix_sort_pairs(sort, c.key_bits, b.key_pid, b.val_pid, b.key_tmp, b.val_tmp, &skey, &oslot, s,
                                    b.counters + kCntFullSort);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_resort_place<kBatch>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock), 0, s, c.one, c.tab,
                       c.n, b.key_tmp, rp.moved, rp.valid_step, rp.valid, b.cell_start, b.cell_end, b.cell_start_alt,
                       b.cell_end_alt, rp.in_cnt[parity], rp.inlist, rp.in_cap, skey, oslot, b.pos_cur, b.vel_cur,
                       b.pid_cur,
                       b.sid_cur, b.pos_alt, b.vel_alt, b.pid_alt, b.sid_alt, b.skey_alt, b.counters, rp.err,
                       b.dt_dev, rp.repair, rp.repairs, b.xyz);
    return hipGetLastError();
}

// The pair-mask path (density writes the force's pair masks, the force
// fills its lists from them): tune[0] == 2, and the default launch
// (tune all 0) whenever cells hold at most 16 candidates.
// tune[0]: 0 = default (pair entries where the lists exist, else pair
// masks), 1 = fused force, 2 = pair masks, 3 = pair entries, 16-48 = the
// candidate-testing force's list size.
bool use_entries(const LaunchCfg& c, const StepBuffers& b) {
    const bool asked = c.tune[0] == 3 || (c.tune[0] == 0 && c.tune[1] == 0 && c.tune[3] == 0);
    return asked && b.pairs && b.xyz && c.max_cap <= kFastMaxCap;
}

bool use_masks(const LaunchCfg& c, const StepBuffers& b) {
    const bool asked = c.tune[0] == 2 || (c.tune[0] == 0 && c.tune[1] == 0 && c.tune[3] == 0);
    return asked && !use_entries(c, b) && b.masks && c.max_cap <= kMaskCap;
}

template <bool kBatch, int kPrec>
void density_impl(const LaunchCfg& c, const StepBuffers& b, hipStream_t s) {
    const uint32_t grid = c.tune[1];
    if (use_entries(c, b))
        hipLaunchKernelGGL((k_density_entries<kBatch, kPrec>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock), 0, s,
                           c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.xyz, b.vel_cur, b.cell_start, b.cell_end,
                           b.rhop, b.pairs, b.pair_cnt, b.rec, b.counters);
    else if (use_masks(c, b))
        hipLaunchKernelGGL((k_density_masks<kBatch, kPrec>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock), 0, s,
                           c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.cell_start, b.cell_end, b.rhop, b.masks, b.counters);
    else if (!grid && c.tune[3] == 2 && b.skey)
        hipLaunchKernelGGL((k_density_cells<kBatch, kPrec, false>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock),
                           0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.skey, b.cell_start, b.cell_end,
                           b.rhop, nullptr, nullptr, b.counters);
    else if (!grid && c.tune[3] == 1 && b.skey)
        hipLaunchKernelGGL((k_density_tiled<kBatch, kPrec, false, 1, true>), dim3(blocks_for(c.n, kBlock)),
                           dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.skey, b.cell_start,
                           b.cell_end, b.rhop, nullptr, nullptr, b.counters);
    else if (!grid && c.tune[3] == 0)
        hipLaunchKernelGGL((k_density_tiled<kBatch, kPrec, false, 1, false>), dim3(blocks_for(c.n, kBlock)),
                           dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.skey, b.cell_start,
                           b.cell_end, b.rhop, nullptr, nullptr, b.counters);
    else if (grid)
        hipLaunchKernelGGL((k_density_persistent<kBatch, kPrec>), dim3(grid), dim3(kBlock), 0, s,
                           c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.cell_start, b.cell_end, b.rhop,
                           b.work, b.counters);
    else
        hipLaunchKernelGGL((k_density<kBatch, kPrec>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock), 0,
                           s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.cell_start, b.cell_end,
                           b.rhop, b.counters);
}

template <bool kBatch, int kPrec, int kList>
void force_list_impl(const LaunchCfg& c, const StepBuffers& b, float4* vel_out, float4* dvel,
                     hipStream_t s) {
    const uint32_t grid = c.tune[1];
    if (!grid && c.tune[3] == 2 && b.skey)
        hipLaunchKernelGGL((k_force_cells<kBatch, kPrec, kList>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock), 0,
                           s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.skey, b.vel_cur, b.rhop, b.cell_start,
                           b.cell_end, vel_out, dvel, b.dt_dev, b.counters);
    else if (!grid && c.tune[3] == 1 && b.skey)
        hipLaunchKernelGGL((k_force_tiled<kBatch, kPrec, kList>), dim3(blocks_for(c.n, kForceTile)),
                           dim3(kForceTile), 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.skey, b.vel_cur,
                           b.rhop, b.cell_start, b.cell_end, vel_out, dvel, b.dt_dev, b.counters);
    else if (!grid && c.tune[3] == 0)
        hipLaunchKernelGGL((k_force_ordered<kBatch, kPrec, kList, 1>), dim3(blocks_for(c.n, kBlock)),
                           dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.vel_cur, b.rhop,
                           b.cell_start, b.cell_end, vel_out, dvel, b.dt_dev, b.counters);
    else if (grid)
        hipLaunchKernelGGL((k_force_persistent<kBatch, kPrec, kList>), dim3(grid), dim3(kBlock), 0, s,
                           c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.vel_cur, b.rhop, b.cell_start,
                           b.cell_end, vel_out, dvel, b.dt_dev, b.counters, b.work + 2);
    else
        hipLaunchKernelGGL((k_force<kBatch, kPrec, kList>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock),
                           0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.vel_cur, b.rhop,
                           b.cell_start, b.cell_end, vel_out, dvel, b.dt_dev, b.counters);
}

template <bool kBatch, int kPrec>
void force_impl(const LaunchCfg& c, const StepBuffers& b, float4* vel_out, float4* dvel,
                hipStream_t s) {
    if (use_entries(c, b)) {
        hipLaunchKernelGGL((k_force_entries<kBatch, kPrec>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock), 0, s, c.one,
                           c.tab, b.sid_cur, c.n, b.vel_cur, b.rec, b.pairs, b.pair_cnt, vel_out, dvel, b.dt_dev);
        return;
    }
    if (use_masks(c, b)) {
        hipLaunchKernelGGL((k_force_ordered<kBatch, kPrec, kForceListMasks, 1, true>), dim3(blocks_for(c.n, kBlock)),
                           dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.vel_cur, b.rhop,
                           b.cell_start, b.cell_end, vel_out, dvel, b.dt_dev, b.counters, b.masks);
        return;
    }
    uint32_t list = c.tune[0] ? c.tune[0] : (uint32_t)kForceListDefault;
    if (list != 1 && list != 16 && list != 24 && list != 32 && list != 48) list = kForceListDefault;
    if (list != 1 && c.max_cap > list) list = 48;
    if (list == 1 || c.max_cap > 48) {
        // fused test + pair loop: chosen by tune[0] == 1, or forced when one
        // cell could overflow any neighbour list
        hipLaunchKernelGGL((k_force_fused<kBatch, kPrec>), dim3(blocks_for(c.n, kBlock)), dim3(kBlock),
                           0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.vel_cur, b.rhop,
                           b.cell_start, b.cell_end, vel_out, dvel, b.dt_dev, b.counters);
        return;
    }
    switch (list) {
        case 16: force_list_impl<kBatch, kPrec, 16>(c, b, vel_out, dvel, s); break;
        case 24: force_list_impl<kBatch, kPrec, 24>(c, b, vel_out, dvel, s); break;
        case 32: force_list_impl<kBatch, kPrec, 32>(c, b, vel_out, dvel, s); break;
        default: force_list_impl<kBatch, kPrec, 48>(c, b, vel_out, dvel, s); break;
    }
}

template <bool kBatch>
void stats_impl(const LaunchCfg& c, const StepBuffers& b, bool advance, hipStream_t s) {
    const uint32_t chunks = stats_chunks(c);
    hipLaunchKernelGGL((k_stats_partials<kBatch>), dim3(chunks, c.nscenes), dim3(kStatBlock), 0, s,
                       c.one, c.tab, c.scene_off, b.pos_cur, b.vel_cur, b.partials);
    hipLaunchKernelGGL((k_stats_finalize<kBatch>), dim3(c.nscenes), dim3(kStatBlock), 0, s, c.one,
                       c.tab, chunks, b.partials, b.stats, b.dt_dev, advance ? 1 : 0);
}

}  // namespace

bool pair_entries_in_use(const LaunchCfg& c, const StepBuffers& b) { return use_entries(c, b); }

uint32_t stats_chunks(const LaunchCfg& c) {
    uint32_t ch = (c.max_scene_n + 2047) / 2048;
    if (ch < 1) ch = 1;
    if (ch > 1024) ch = 1024;
    return ch;
}

size_t stats_partials_count(const LaunchCfg& c) {
    return (size_t)stats_chunks(c) * c.nscenes * kPNum;
}

hipError_t launch_integrate(const LaunchCfg& c, const StepBuffers& b, hipStream_t s) {
    const dim3 grid(blocks_for(c.n, kBlock)), block(kBlock);
    if (c.tab)
        hipLaunchKernelGGL((k_integrate<true>), grid, block, 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur, b.vel_cur,
                           b.counters, b.dt_dev);
    else
        hipLaunchKernelGGL((k_integrate<false>), grid, block, 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur,
                           b.vel_cur, b.counters, b.dt_dev);
    return hipGetLastError();
}

hipError_t launch_integrate_moves(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, int parity,
                                  uint32_t* nkey, hipStream_t s) {
    const dim3 grid(blocks_for(c.n, kBlock)), block(kBlock);
    if (c.tab)
        hipLaunchKernelGGL((k_integrate_moves<true>), grid, block, 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur,
                           b.vel_cur, b.pid_cur, b.skey, nkey, rp.valid, rp.moved, rp.in_cnt[parity], rp.inlist,
                           rp.in_cap, rp.tile_delta[parity], b.key_pid, b.val_pid, b.counters, b.dt_dev,
                           c.num_cells, b.cell_start, b.cell_end, rp.err, rp.repair);
    else
        hipLaunchKernelGGL((k_integrate_moves<false>), grid, block, 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur,
                           b.vel_cur, b.pid_cur, b.skey, nkey, rp.valid, rp.moved, rp.in_cnt[parity], rp.inlist,
                           rp.in_cap, rp.tile_delta[parity], b.key_pid, b.val_pid, b.counters, b.dt_dev,
                           c.num_cells, b.cell_start, b.cell_end, rp.err, rp.repair);
    return hipGetLastError();
}

hipError_t resort_apply(const LaunchCfg& c, const StepBuffers& b, const ResortPlan& rp, int parity,
                        const RadixSortPlan& sort, hipStream_t s) {
    return c.tab ? resort_apply_impl<true>(c, b, rp, parity, sort, s)
                 : resort_apply_impl<false>(c, b, rp, parity, sort, s);
}

hipError_t launch_density_lists(const LaunchCfg& c, const StepBuffers& b, uint32_t* nbr, uint32_t* nnb,
                                hipStream_t s) {
    if (!b.skey) return hipErrorInvalidValue;
    const bool kahan = c.precision == SPH_PREC_KAHAN;
    const dim3 grid(blocks_for(c.n, kBlock)), block(kBlock);
#define SPH_DENS_LISTS(BATCH, PREC)                                                                   \
    if (c.tune[3] == 2)                                                                               \
        hipLaunchKernelGGL((k_density_cells<BATCH, PREC, true>), grid, block, 0, s, c.one, c.tab,          \
                           b.sid_cur, c.n, b.pos_cur, b.skey, b.cell_start, b.cell_end, b.rhop, nbr, nnb,  \
                           b.counters);                                                                   \
    else if (c.tune[3] == 1)                                                                          \
        hipLaunchKernelGGL((k_density_tiled<BATCH, PREC, true, 1, true>), grid, block, 0, s, c.one, c.tab, \
                           b.sid_cur, c.n, b.pos_cur, b.skey, b.cell_start, b.cell_end, b.rhop, nbr, nnb,  \
                           b.counters);                                                                   \
    else if (c.tune[3] == 0)                                                                          \
        hipLaunchKernelGGL((k_density_tiled<BATCH, PREC, true, 1, false>), grid, block, 0, s, c.one,       \
                           c.tab, b.sid_cur, c.n, b.pos_cur, b.skey, b.cell_start, b.cell_end, b.rhop, nbr, \
                           nnb, b.counters);                                                              \
    else                                                                                              \
        hipLaunchKernelGGL((k_density_tiled<BATCH, PREC, true, 0, false>), grid, block, 0, s, c.one,       \
                           c.tab, b.sid_cur, c.n, b.pos_cur, b.skey, b.cell_start, b.cell_end, b.rhop, nbr, \
                           nnb, b.counters)
    if (c.tab) {
        if (kahan)
            SPH_DENS_LISTS(true, SPH_PREC_KAHAN);
        else
            SPH_DENS_LISTS(true, SPH_PREC_REFERENCE);
    } else {
        if (kahan)
            SPH_DENS_LISTS(false, SPH_PREC_KAHAN);
        else
            SPH_DENS_LISTS(false, SPH_PREC_REFERENCE);
    }
#undef SPH_DENS_LISTS
    return hipGetLastError();
}

hipError_t launch_density(const LaunchCfg& c, const StepBuffers& b, hipStream_t s) {
    const bool kahan = c.precision == SPH_PREC_KAHAN;
    if (c.tab) {
        if (kahan)
            density_impl<true, SPH_PREC_KAHAN>(c, b, s);
        else
            density_impl<true, SPH_PREC_REFERENCE>(c, b, s);
    } else {
        if (kahan)
            density_impl<false, SPH_PREC_KAHAN>(c, b, s);
        else
            density_impl<false, SPH_PREC_REFERENCE>(c, b, s);
    }
    return hipGetLastError();
}

hipError_t launch_force(const LaunchCfg& c, const StepBuffers& b, float4* vel_out, bool keep_dvel,
                        hipStream_t s) {
    float4* dvel = keep_dvel ? b.dvel : nullptr;
    const bool kahan = c.precision == SPH_PREC_KAHAN;
    if (c.precision == SPH_PREC_FAST) {
        // the pair entries and records of k_density_fast
        if (!b.pairs || !b.pair_cnt || !b.rec) return hipErrorInvalidValue;
        const dim3 grid(blocks_for(c.n, kBlock)), block(kBlock);
        if (c.tab)
            hipLaunchKernelGGL((k_force_entries<true, SPH_PREC_FAST>), grid, block, 0, s, c.one, c.tab, b.sid_cur, c.n,
                               b.vel_cur, b.rec, b.pairs, b.pair_cnt, vel_out, dvel, b.dt_dev);
        else
            hipLaunchKernelGGL((k_force_entries<false, SPH_PREC_FAST>), grid, block, 0, s, c.one, c.tab, b.sid_cur,
                               c.n, b.vel_cur, b.rec, b.pairs, b.pair_cnt, vel_out, dvel, b.dt_dev);
    } else if (c.tab) {
        if (kahan)
            force_impl<true, SPH_PREC_KAHAN>(c, b, vel_out, dvel, s);
        else
            force_impl<true, SPH_PREC_REFERENCE>(c, b, vel_out, dvel, s);
    } else {
        if (kahan)
            force_impl<false, SPH_PREC_KAHAN>(c, b, vel_out, dvel, s);
        else
            force_impl<false, SPH_PREC_REFERENCE>(c, b, vel_out, dvel, s);
    }
    if (c.physics) return launch_extra_terms(c, b, vel_out, keep_dvel, s);
    return hipGetLastError();
}

hipError_t launch_extra_terms(const LaunchCfg& c, const StepBuffers& b, float4* vel_out, bool keep_dvel,
                              hipStream_t s) {
    float4* dvel = keep_dvel ? b.dvel : nullptr;
    const dim3 grid(blocks_for(c.n, kBlock));
    if (c.tab)
        hipLaunchKernelGGL(k_extra_terms<true>, grid, dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n,
                           b.pos_cur, b.vel_cur, vel_out, dvel, b.dt_dev);
    else
        hipLaunchKernelGGL(k_extra_terms<false>, grid, dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n,
                           b.pos_cur, b.vel_cur, vel_out, dvel, b.dt_dev);
    return hipGetLastError();
}

hipError_t launch_colour(const LaunchCfg& c, const StepBuffers& b, float4* colour, float* dye, hipStream_t s) {
    const dim3 grid(blocks_for(c.n, kBlock));
    if (c.tab)
        hipLaunchKernelGGL(k_colour<true>, grid, dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur,
                           b.vel_cur, b.rhop, b.dvel, b.pid_cur, b.cell_start, b.cell_end, b.dt_dev, colour, dye);
    else
        hipLaunchKernelGGL(k_colour<false>, grid, dim3(kBlock), 0, s, c.one, c.tab, b.sid_cur, c.n, b.pos_cur,
                           b.vel_cur, b.rhop, b.dvel, b.pid_cur, b.cell_start, b.cell_end, b.dt_dev, colour, dye);
    return hipGetLastError();
}

hipError_t launch_stats(const LaunchCfg& c, const StepBuffers& b, bool advance_clock,
                        hipStream_t s) {
    if (c.tab)
        stats_impl<true>(c, b, advance_clock, s);
    else
        stats_impl<false>(c, b, advance_clock, s);
    return hipGetLastError();
}

hipError_t launch_cell_stats(const LaunchCfg& c, const StepBuffers& b, uint32_t* max_count,
                             hipStream_t s) {
    uint32_t nb = blocks_for(c.num_cells, kBlock);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_cell_stats, dim3(nb), dim3(kBlock), 0, s, c.num_cells, b.cell_start,
                       b.cell_end, max_count);
    return hipGetLastError();
}

hipError_t launch_scatter_to_original(uint32_t n, const uint32_t* pid, const void* src, void* dst,
                                      int record_floats, int src_stride_floats, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_to_original, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, n,
                       pid, (const float*)src, (float*)dst, record_floats, src_stride_floats);
    return hipGetLastError();
}

hipError_t launch_update_from_original(uint32_t n, const uint32_t* pid, const float4* staging,
                                       uint32_t first, uint32_t count, float4* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_update_from_original, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, n,
                       pid, staging, first, count, dst);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* dst, uint32_t value, size_t count, hipStream_t s) {
    uint32_t nb = blocks_for(count, kBlock);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(nb), dim3(kBlock), 0, s, dst, value, count);
    return hipGetLastError();
}

hipError_t launch_iota_u32(uint32_t* dst, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_iota_u32, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, dst, n);
    return hipGetLastError();
}

hipError_t launch_scene_ids(uint32_t n, uint32_t nscenes, const uint32_t* scene_off,
                            const uint32_t* pid, uint32_t* sid, hipStream_t s) {
    hipLaunchKernelGGL(k_scene_ids, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, n, nscenes,
                       scene_off, pid, sid);
    return hipGetLastError();
}

}  // namespace sph
