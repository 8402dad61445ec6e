// sph_scene.cpp - host-only part of the C ABI: scene defaults, derivation of
// the kernel constants, and the seeded lattice initialiser.  Needs no device.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/pibiti_sph.h"
#include "sph_derive.hpp"

namespace {

// The reference's PI (source/CUDA/Params.cuh:18).
constexpr float kPi = 3.141592654f;

// SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014), used as a counter-based generator:
// draw(seed, stream, counter) = mix(mix(seed ^ stream*K1) + (counter+1)*K2).
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t counter) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + (counter + 1) * 0x9E3779B97F4A7C15ull);
}

// Uniform double in [0, 1) from the top 53 bits.
inline double u01(uint64_t seed, uint64_t stream, uint64_t counter) {
    return (double)(draw(seed, stream, counter) >> 11) * (1.0 / 9007199254740992.0);
}

// Largest float r2 with fmaxf(min_dist, sqrtf(r2)) < h (sqrtf correctly
// rounded, hence monotone).  -1 when no r2 qualifies.
float force_threshold(float h, float min_dist) {
    if (!(min_dist < h)) return -1.0f;
    float t = h * h;
    while (t > 0.0f && !(std::sqrt(t) < h)) t = std::nextafter(t, 0.0f);
    for (;;) {
        float up = std::nextafter(t, INFINITY);
        if (std::sqrt(up) < h)
            t = up;
        else
            break;
    }
    return t;
}

// Correctly rounded float sine and cosine (double evaluation, one rounding).
inline float sin_rn(float x) { return (float)std::sin((double)x); }
inline float cos_rn(float x) { return (float)std::cos((double)x); }

}  // namespace

namespace sph {

void derive_shape(sph_params& p) {
    // wmin.z += rTwist*(1 + sinf(rAngle))  (source/CUDA/System.cu:60)
    p.wave_zmin_shift = p.rot_twist * (1.0f + sin_rn(p.rot_angle));
    // pump outlet (source/CUDA/System.cu:104-117): rad = worldMax.x
    const float rad = p.world_max[0], ang = p.pump_ang_out;
    if (ang < 0.5f) {
        p.pump_xs = sin_rn(ang * p.shape_s[2]) * rad;
        p.pump_zs = (cos_rn(ang * p.shape_s[2]) * rad) * p.shape_s[3];
    } else {
        p.pump_xs = 0.09f * p.shape_s[3];
        p.pump_zs = 0.0f;
    }
    // worldSizeD (source/SPH/Scene.cpp:99-103): b = soft band - r per side
    const float bd = p.dist_bnd_soft - p.particle_r;
    for (int c = 0; c < 3; ++c) p.world_size_d[c] = (p.world_max[c] - bd) - (p.world_min[c] + bd);
    uint32_t phys = 0;
    if (p.rot_type > 0)
        phys |= SPH_PHYS_ROTOR;
    else if (p.collider)
        phys |= SPH_PHYS_COLLIDER;
    if (p.hmap_type > 0) phys |= SPH_PHYS_HMAP;
    for (int k = 0; k < SPH_NUM_ACCEL; ++k)
        if (p.acc[k].type != SPH_ACC_OFF) phys |= SPH_PHYS_ACCEL;
    p.physics = phys;
}

int check_shape(const sph_params& p) {
    if (p.bnd_type > SPH_BND_PUMP_Y || p.bnd_eff_z > SPH_BND_EFF_WAVE || p.collider > 1u)
        return SPH_ERR_INVALID_PARAMS;
    for (int k = 0; k < SPH_NUM_ACCEL; ++k)
        if (p.acc[k].type > SPH_ACC_CYL_Y_SMOOTH) return SPH_ERR_INVALID_PARAMS;
    if (p.rot_type < 0 || p.rot_type > SPH_ROTOR_PROPELLER_Z || p.hmap_type < 0 || p.hmap_type > SPH_HMAP_YZ)
        return SPH_ERR_INVALID_PARAMS;
    if (p.rot_type > 0) {
        // the rotor's loops run over these counts: bound them
        if (p.rot_blades < 1 || p.rot_blades > 64) return SPH_ERR_INVALID_PARAMS;
        for (int c = 0; c < 3; ++c)
            if (p.rot_size[c] < 0 || p.rot_size[c] > 256) return SPH_ERR_INVALID_PARAMS;
    }
    if (p.hmap_type > 0 && !(p.rot_spc > 0.0f)) return SPH_ERR_INVALID_PARAMS;
    if (p.clr_type > SPH_CLR_NONE || p.dye_type < 0 || p.dye_type > SPH_DYE_SPHERE || p.hue < 0 || p.hue > 1)
        return SPH_ERR_INVALID_PARAMS;
    return SPH_OK;
}

bool changes_per_step(const sph_params& p) { return p.rot_vel != 0.0f || p.dye_clear > 0; }

void advance_frame(sph_params& p) {
    // App::UpdateEmitter counts the dye clear down before each step
    // (source/App/Update.cpp:23)
    if (p.dye_clear > 0) p.dye_clear = p.dye_clear - 1;
    advance_motion(p);
}

void advance_motion(sph_params& p) {
    // App::UpdateEmitter before each cSPH::Update (source/App/Update.cpp:15-20,
    // source/App/Render.cpp:10-11): the rotor angle (also the wave phase)
    // advances by rVel dt, the second propeller's by rVel r2Vel dt.
    if (p.rot_vel != 0.0f) {
        p.rot_angle = p.rot_angle + p.rot_vel * p.time_step;
        if (p.rot2_dist > 0.0f) {
            p.rot2_angle = p.rot2_angle + (p.rot_vel * p.rot2_vel) * p.time_step;
            p.rot2_twist = p.rot2_vel < 0.0f ? -1.0f : 1.0f;
        }
        derive_shape(p);
    }
}

}  // namespace sph

extern "C" {

uint32_t sph_abi_version(void) { return SPH_ABI_VERSION; }

const char* sph_strerror(int status) {
    switch (status) {
        case SPH_OK: return "ok";
        case SPH_ERR_INVALID_ARGUMENT: return "invalid argument";
        case SPH_ERR_INVALID_PARAMS: return "invalid simulation parameters";
        case SPH_ERR_NO_DEVICE: return "no HIP device available";
        case SPH_ERR_OUT_OF_MEMORY: return "device out of memory";
        case SPH_ERR_DEVICE: return "HIP runtime error";
        case SPH_ERR_NOT_INITIALISED: return "particle state not initialised";
        case SPH_ERR_UNSUPPORTED: return "unsupported feature";
        case SPH_ERR_CAPACITY: return "size exceeds a fixed capacity";
        case SPH_ERR_ORDER: return "a step's particle order could not be built (results invalid; upload the state again)";
        default: return "unknown status";
    }
}

int sph_scene_default(sph_scene* s) {
    if (!s) return SPH_ERR_INVALID_ARGUMENT;
    std::memset(s, 0, sizeof(*s));
    // source/SPH/Scene.cpp:11-31
    s->num_particles = 7 * 8 * 1024;
    s->max_par_in_cell = 16;
    s->time_step = 0.0026f;
    s->global_damping = 1.0f;
    s->gravity[0] = 0.0f;
    s->gravity[1] = -9.81f;
    s->gravity[2] = 0.0f;
    s->particle_r = 0.004f;
    s->min_dist = 1.0f;
    s->h = 0.01f;
    s->rest_density = 1000.0f;
    s->min_dens = 1.0f;
    s->stiffness = 3.0f;
    s->viscosity = 0.5f;
    const float w[3] = {0.2f, 0.25f, 0.2f};
    for (int c = 0; c < 3; ++c) {
        s->world_min[c] = -w[c];
        s->world_max[c] = w[c];
    }
    s->cell_size = s->particle_r * 2.0f;
    s->particle_mass = 0.0f;  // reference mass formula
    s->dist_bnd_soft = 8.0f;
    s->dist_bnd_hard = 1.0f;
    s->bnd_stiff = 30000.0f;
    s->bnd_damp = 256.0f;
    s->bnd_damp_c = 60.0f;
    s->eos = SPH_EOS_LINEAR;
    s->precision = SPH_PREC_REFERENCE;
    s->cfl = 0.0f;
    // source/SPH/Scene.cpp:31-48 (boundary, collider, pump, rotor twist)
    s->bnd_type = SPH_BND_BOX;
    s->bnd_eff_z = SPH_BND_EFF_NONE;
    s->rot2_angle = 0.0f;
    s->rot_twist = 0.04f;
    s->rot_angle = 0.0f;
    s->rot_vel = 0.0f;
    s->exit_vel = 0.03f;
    s->exit_dist = 0.04f;
    s->pump_ang_out = 0.7f;
    s->h_close = -0.03f;
    s->pump_rad_in = 0.4f;
    const float ps[6] = {0.90f, 0.90f, 1.198f, 0.7782f, 0.8696f, 0.843f};
    for (int k = 0; k < 6; ++k) s->shape_s[k] = ps[k];
    // The collider's values are the reference's, but it starts switched off:
    // the hot path's workloads run without it.  The scene-file loader turns
    // it on, as the reference applies it in every scene without a rotor.
    s->collider = 0;
    s->coll_pos[0] = -77.0f;
    s->coll_r = 0.05f;
    s->coll_spring = 80.0f;
    s->coll_damping = 0.02f;
    s->coll_shear = 0.1f;
    // rotor (source/SPH/Scene.cpp:40-48)
    s->rot_type = SPH_ROTOR_NONE;
    s->rot_size[0] = 0;
    s->rot_size[1] = 4;
    s->rot_size[2] = 12;
    s->rot_blades = 4;
    s->rot_r = 3.0f;
    s->rot_spc = 0.7f;
    s->rot2_dist = 0.0f;
    s->rot2_vel = 1.0f;
    s->rot2_twist = 1.0f;
    s->hmap_type = SPH_HMAP_NONE;
    // colours and dye (source/SPH/Scene.cpp:61-64).  No colour output by
    // default (the reference's app starts in CLR_VelAcc, which the scene-file
    // loader keeps); the dye-clear count starts at 0 for the same reason.
    s->clr_type = SPH_CLR_NONE;
    s->hue = 0;
    s->brightness = 0.3f;
    s->contrast = 0.3f;
    s->dye_type = SPH_DYE_OFF;
    s->dye_clear = 0;
    s->dye_fade = 1.0f;
    s->dye_pos[0] = 0.0f;
    s->dye_pos[1] = -0.21f;
    s->dye_pos[2] = 0.0f;
    s->dye_size[0] = 0.008f;
    s->dye_size[1] = 0.012f;
    s->dye_size[2] = 0.008f;
    // source/SPH/Scene.cpp:54-57
    for (int k = 0; k < SPH_NUM_ACCEL; ++k) {
        sph_accel& a = s->acc[k];
        a.size[0] = 0.005f;
        a.size[1] = 0.010f;
        a.size[2] = 0.005f;
        a.acc[1] = 10.0f;
        a.type = SPH_ACC_OFF;
    }
    return SPH_OK;
}

int sph_params_from_scene(const sph_scene* s, sph_params* p) {
    if (!s || !p) return SPH_ERR_INVALID_ARGUMENT;
    if (s->num_particles == 0 || !(s->h > 0.0f) || !(s->cell_size > 0.0f) ||
        !(s->time_step > 0.0f) || !(s->rest_density > 0.0f) || !(s->particle_r > 0.0f) ||
        s->max_par_in_cell == 0 || s->eos > SPH_EOS_TAIT || s->precision > SPH_PREC_FAST ||
        !(s->cfl >= 0.0f) || !(s->stiffness >= 0.0f))
        return SPH_ERR_INVALID_PARAMS;
    for (int c = 0; c < 3; ++c)
        if (!(s->world_max[c] > s->world_min[c])) return SPH_ERR_INVALID_PARAMS;

    std::memset(p, 0, sizeof(*p));
    p->time_step = s->time_step;
    p->num_particles = s->num_particles;
    p->max_par_in_cell = s->max_par_in_cell;
    for (int c = 0; c < 3; ++c) p->gravity[c] = s->gravity[c];
    p->global_damping = s->global_damping;

    // Scene::_UpdatePar (source/SPH/Scene.cpp:76-94).  The reference's
    // pow(float, int) calls resolve to the double overload, so Poly6, Spiky
    // and Lap are formed in double and rounded once to float.
    const float r = s->particle_r, h = s->h;
    p->particle_r = r;
    p->h = h;
    p->min_dist = s->min_dist * r;
    p->h2 = h * h;
    p->poly6_kern = (float)(315.0f / ((double)(64.0f * kPi) * std::pow((double)h, 9)));
    p->spiky_kern = (float)((double)(-0.5f * -45.0f) / ((double)kPi * std::pow((double)h, 6)));
    p->lap_kern = (float)(45.0f / ((double)kPi * std::pow((double)h, 6)));
    const float md = s->min_dens * s->rest_density;
    p->min_dens = 1.0f / (md * md);
    p->rest_density = s->rest_density;
    p->particle_mass = s->particle_mass > 0.0f
                           ? s->particle_mass
                           : ((s->rest_density * 4.0f) / 3.0f * kPi) * ((r * r) * r);
    p->stiffness = s->stiffness;
    p->viscosity = s->viscosity;
    p->dist_bnd_soft = s->dist_bnd_soft * r;
    p->dist_bnd_hard = s->dist_bnd_hard * r;
    p->bnd_stiff = s->bnd_stiff;
    p->bnd_damp = s->bnd_damp;
    p->bnd_damp_c = s->bnd_damp_c;
    p->bnd_type = s->bnd_type;
    p->bnd_eff_z = s->bnd_eff_z;
    p->rot2_angle = s->rot2_angle;
    p->rot_twist = s->rot_twist;
    p->rot_angle = s->rot_angle;
    p->rot_vel = s->rot_vel;
    p->exit_vel = s->exit_vel;
    p->exit_dist = s->exit_dist;
    p->pump_ang_out = s->pump_ang_out;
    p->h_close = s->h_close;
    p->pump_rad_in = s->pump_rad_in;
    for (int k = 0; k < 6; ++k) p->shape_s[k] = s->shape_s[k];
    p->collider = s->collider;
    p->coll_r = s->coll_r;
    p->coll_spring = s->coll_spring;
    p->coll_damping = s->coll_damping;
    p->coll_shear = s->coll_shear;
    for (int k = 0; k < SPH_NUM_ACCEL; ++k) p->acc[k] = s->acc[k];
    p->rot_type = s->rot_type;
    for (int c = 0; c < 3; ++c) p->rot_size[c] = s->rot_size[c];
    p->rot_blades = s->rot_blades;
    p->rot_r = s->rot_r * r;               // rotR *= particleR
    p->rot_spc = s->rot_spc * p->rot_r;    // rotSpc *= rotR (source/SPH/Scene.cpp:93)
    p->rot2_dist = s->rot2_dist;
    p->rot2_vel = s->rot2_vel;
    p->rot2_twist = s->rot2_twist;
    p->hmap_type = s->hmap_type;
    p->clr_type = s->clr_type;
    p->hue = s->hue;
    p->brightness = s->brightness;
    p->contrast = s->contrast;
    p->dye_type = s->dye_type;
    p->dye_clear = s->dye_clear;
    p->dye_fade = s->dye_fade;
    for (int c = 0; c < 3; ++c) {
        p->dye_pos[c] = s->dye_pos[c];
        p->dye_size[c] = s->dye_size[c];
    }

    // Scene::_UpdateGrid (source/SPH/Scene.cpp:98-109).
    uint64_t cells = 1;
    for (int c = 0; c < 3; ++c) {
        p->world_min[c] = s->world_min[c];
        p->world_max[c] = s->world_max[c];
        p->cell_size[c] = s->cell_size;
        const float size = s->world_max[c] - s->world_min[c];
        const float g = std::ceil(size / s->cell_size);
        if (!(g >= 1.0f) || g > 16777216.0f) return SPH_ERR_INVALID_PARAMS;
        p->grid_size[c] = (uint32_t)g;
        cells *= p->grid_size[c];
    }
    if (cells >= (1ull << 31)) return SPH_ERR_INVALID_PARAMS;
    p->grid_size_yx = p->grid_size[1] * p->grid_size[0];
    p->num_cells = (uint32_t)cells;
    uint32_t bits = 0;
    while ((1ull << bits) < cells) ++bits;
    p->key_bits = bits;

    p->eos = s->eos;
    p->precision = s->precision;
    p->cfl = s->cfl;
    p->sound_speed = std::sqrt(s->stiffness);
    p->force_r2_max = force_threshold(h, p->min_dist);

    // Collider placement of Scene::_UpdateGrid (source/SPH/Scene.cpp:100,111-112).
    if (s->coll_pos[0] == -77.0f) {
        const float b = p->dist_bnd_soft - r;
        p->coll_pos[0] = (p->world_min[0] + b) - p->coll_r * 1.2f;
        p->coll_pos[1] = (p->world_min[1] + b) + p->coll_r * 1.0f;
        p->coll_pos[2] = 0.0f;
    } else {
        for (int c = 0; c < 3; ++c) p->coll_pos[c] = s->coll_pos[c];
    }
    const int st = sph::check_shape(*p);
    if (st != SPH_OK) return st;
    sph::derive_shape(*p);
    return SPH_OK;
}

int sph_scene_lattice(uint32_t nx, uint32_t ny, uint32_t nz, const float region_min[3],
                      float spacing, float jitter, float vel_sigma, const float vel_offset[3],
                      uint64_t seed, uint64_t stream, float* pos4, float* vel4) {
    if (!region_min || !pos4 || !vel4 || nx == 0 || ny == 0 || nz == 0 || !(spacing > 0.0f))
        return SPH_ERR_INVALID_ARGUMENT;
    const double s = spacing, jit = jitter, sig = vel_sigma;
    const double off[3] = {vel_offset ? vel_offset[0] : 0.0, vel_offset ? vel_offset[1] : 0.0,
                           vel_offset ? vel_offset[2] : 0.0};
    // x fastest, then z, then y: the reference's default fill order
    // (initLast = 1, Inc3(x,z,y), source/SPH/SPH_Init.cpp:41-46).
    for (uint32_t iy = 0; iy < ny; ++iy)
        for (uint32_t iz = 0; iz < nz; ++iz)
            for (uint32_t ix = 0; ix < nx; ++ix) {
                const uint64_t n = ((uint64_t)iy * nz + iz) * nx + ix;
                const uint32_t idx[3] = {ix, iy, iz};
                for (int c = 0; c < 3; ++c) {
                    const double u = u01(seed, stream, n * 8 + (uint64_t)c);
                    pos4[4 * n + c] =
                        (float)((double)region_min[c] + ((double)idx[c] + 0.5) * s + (2.0 * u - 1.0) * jit * s);
                }
                pos4[4 * n + 3] = 1.0f;
                // Box-Muller from two pairs of uniforms.
                double g[4];
                for (int k = 0; k < 2; ++k) {
                    const double u1 = 1.0 - u01(seed, stream, n * 8 + 3 + 2 * (uint64_t)k);
                    const double u2 = u01(seed, stream, n * 8 + 4 + 2 * (uint64_t)k);
                    const double rr = std::sqrt(-2.0 * std::log(u1));
                    g[2 * k] = rr * std::cos(2.0 * M_PI * u2);
                    g[2 * k + 1] = rr * std::sin(2.0 * M_PI * u2);
                }
                for (int c = 0; c < 3; ++c) vel4[4 * n + c] = (float)(off[c] + sig * g[c]);
                vel4[4 * n + 3] = 0.0f;
            }
    return SPH_OK;
}

}  // extern "C"
