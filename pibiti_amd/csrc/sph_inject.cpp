// sph_inject.cpp - host-only particle injection of the C ABI: the drops,
// rain and emitter blocks the reference's app writes into the particle
// arrays between steps through cSPH::setArray.  Needs no device; the caller
// uploads the records with sph_set_array.
//
//   sph_scene_drop ... cSPH::Drop (source/SPH/SPH_Init.cpp:82-118)
//   sph_scene_emit ... the emitter block of App::UpdateEmitter
//                      (source/App/Update.cpp:62-92, App::mulTr source/App/Input.cpp:588-593)
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/pibiti_sph.h"

namespace {

constexpr float kPi = 3.141592654f;  // source/CUDA/Params.cuh:18

inline float sin_rn(float x) { return (float)std::sin((double)x); }
inline float cos_rn(float x) { return (float)std::cos((double)x); }

// Counter-based uniform in [0, 1] standing in for rand() / RAND_MAX
// (SplitMix64 finaliser, as sph_scene_lattice).
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline float unit(uint64_t seed, uint64_t counter) {
    const uint64_t d = mix64(mix64(seed ^ 0x5DEECE66Dull) + (counter + 1) * 0x9E3779B97F4A7C15ull);
    return (float)(d >> 40) * (1.0f / 16777215.0f);
}

// The modelview matrix of glLoadIdentity; glRotatef(ax, 1,0,0);
// glRotatef(-ay, 0,1,0) (angles in degrees), column-major as
// glGetFloatv returns it: M = Rx(ax) Ry(-ay), m[col*4 + row].
void emitter_matrix(float ax_deg, float ay_deg, float m[16]) {
    const float a = (float)((double)ax_deg * (3.14159265358979323846 / 180.0));
    const float b = (float)((double)-ay_deg * (3.14159265358979323846 / 180.0));
    const float ca = cos_rn(a), sa = sin_rn(a), cb = cos_rn(b), sb = sin_rn(b);
    const float rx[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, ca, -sa}, {0.0f, sa, ca}};
    const float ry[3][3] = {{cb, 0.0f, sb}, {0.0f, 1.0f, 0.0f}, {-sb, 0.0f, cb}};
    std::memset(m, 0, 16 * sizeof(float));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[c * 4 + r] = (rx[r][0] * ry[0][c] + rx[r][1] * ry[1][c]) + rx[r][2] * ry[2][c];
    m[15] = 1.0f;
}

// App::mulTr: v times the transpose of the 3x3 part.
void mul_tr(const float v[3], const float m[16], float r[3]) {
    r[0] = v[0] * m[0] + v[1] * m[1] + v[2] * m[2];
    r[1] = v[0] * m[4] + v[1] * m[5] + v[2] * m[6];
    r[2] = v[0] * m[8] + v[1] * m[9] + v[2] * m[10];
}

}  // namespace

extern "C" {

int sph_scene_drop(const sph_scene* scene, const sph_scene_info* info, int random, uint64_t seed,
                   uint64_t draw, uint32_t first, float* pos4, float* vel4, uint32_t* count) {
    if (!scene || !info || !pos4 || !vel4 || !count) return SPH_ERR_INVALID_ARGUMENT;
    sph_params p;
    const int st = sph_params_from_scene(scene, &p);
    if (st != SPH_OK) return st;
    const uint32_t n = scene->num_particles;
    if (first > n) return SPH_ERR_INVALID_ARGUMENT;
    const float r = info->drop_r, spc = info->spacing, r2 = r * r, db = (r + 1.0f) * spc, d = 0.5f;
    if (!(r >= 0.0f) || !(r <= 1000.0f)) return SPH_ERR_INVALID_ARGUMENT;
    float pos[3] = {0.5f, 0.5f, 0.5f};
    if (random) {
        // the drop's centre in the inner world: high above a box or
        // cylinder-Y floor, anywhere across a box / cylinder Z, else inside
        // the unit disc (source/SPH/SPH_Init.cpp:89-97)
        const uint32_t t = scene->bnd_type;
        pos[1] = (t == SPH_BND_BOX || t == SPH_BND_CYL_Y) ? 1.0f : 0.75f;
        if (t == SPH_BND_BOX || t == SPH_BND_CYL_Z) {
            pos[0] = unit(seed, 4 * draw);
            pos[2] = unit(seed, 4 * draw + 1);
        } else {
            const float rr = unit(seed, 4 * draw + 2);
            const float k = (2.0f * kPi - 0.0f) * unit(seed, 4 * draw + 3) + 0.0f;
            pos[0] = d + (d * cos_rn(k)) * rr;
            pos[2] = d - (d * sin_rn(k)) * rr;
        }
    }
    // posw = worldMinD + r3 + pos * (worldSizeD - r3*2)
    const float bd = p.dist_bnd_soft - p.particle_r;
    float posw[3];
    for (int c = 0; c < 3; ++c) posw[c] = ((p.world_min[c] + bd) + db) + pos[c] * (p.world_size_d[c] - db * 2.0f);
    const int lo = (int)(-r);
    uint32_t a = first;
    for (int z = lo; (float)z <= r; ++z)
        for (int y = lo; (float)y <= r; ++y)
            for (int x = lo; (float)x <= r; ++x)
                if ((float)(x * x + y * y + z * z) <= r2 && a < n) {
                    float* q = pos4 + 4 * (size_t)(a - first);
                    q[0] = (float)x * spc + posw[0];
                    q[1] = (float)y * spc + posw[1];
                    q[2] = (float)z * spc + posw[2];
                    q[3] = 1.0f;
                    float* w = vel4 + 4 * (size_t)(a - first);
                    w[0] = w[1] = w[2] = w[3] = 0.0f;
                    ++a;
                }
    *count = a - first;
    return SPH_OK;
}

int sph_scene_emit(const sph_scene_info* info, uint32_t emitter, float* pos4, float* vel4, uint32_t* count) {
    if (!info || !pos4 || !vel4 || !count || emitter >= 4) return SPH_ERR_INVALID_ARGUMENT;
    const sph_emitter& em = info->emit[emitter];
    *count = 0;
    if (em.size <= 0) return SPH_OK;
    const int ex = em.size, ey = em.size2 == 0 ? ex : em.size2;
    if (ey <= 0 || (int64_t)ex * ey > SPH_MAX_EMIT) return SPH_ERR_CAPACITY;
    const float spc = info->spacing;
    float m[16];
    emitter_matrix(em.rot[0], em.rot[1], m);
    const float ev[3] = {0.0f, 0.0f, em.vel};
    float v[3];
    mul_tr(ev, m, v);
    const float z = (float)(ex - 1) * 0.5f, z2 = (float)(ey - 1) * 0.5f;
    uint32_t i = 0;
    for (int y = 0; y < ey; ++y)
        for (int x = 0; x < ex; ++x, ++i) {
            const float pp[3] = {((float)x - z) * spc, ((float)y - z2) * spc, -spc};
            float q[3];
            mul_tr(pp, m, q);
            float* o = pos4 + 4 * (size_t)i;
            for (int c = 0; c < 3; ++c) o[c] = q[c] + em.pos[c];
            o[3] = 1.0f;
            float* w = vel4 + 4 * (size_t)i;
            for (int c = 0; c < 3; ++c) w[c] = v[c];
            w[3] = 0.0f;
        }
    *count = i;
    return SPH_OK;
}

}  // extern "C"
