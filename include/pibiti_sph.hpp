/* pibiti_sph.hpp - C++ host class over the C ABI of pibiti_sph.h.
 *
 * cSPH mirrors the reference's host class (source/SPH/SPH.h:9-50) for the
 * step path, without GL buffers: a Scene (scn) whose derived constants go
 * to the device context, InitScene/_InitMem/_FreeMem, Reset, Update,
 * getArray/setArray with the reference's 0 = position / 1 = velocity flag
 * values, and a density getter.  Every call goes through the C ABI; errors
 * throw pibiti::Error instead of exiting the process.
 */
#ifndef PIBITI_SPH_HPP
#define PIBITI_SPH_HPP

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "pibiti_sph.h"

namespace pibiti {

class Error : public std::runtime_error {
public:
    Error(int st, const char* what)
        : std::runtime_error(std::string(what) + ": " + sph_strerror(st)), status(st) {}
    int status;
};

inline void check(int st, const char* what) {
    if (st != SPH_OK) throw Error(st, what);
}

/* The block cSPH::Reset(0) fills (source/SPH/SPH_Init.cpp:40-76): a lattice
 * from `min` in steps of `spacing`, x fastest, then z, then y (the
 * reference's default initLast), optionally jittered and with random
 * velocities.                                                               */
struct InitVolume {
    uint32_t nx = 0, ny = 0, nz = 0;
    float min[3] = {0.0f, 0.0f, 0.0f};
    float spacing = 0.0f;
    float jitter = 0.0f;     /* fraction of spacing                          */
    float vel_sigma = 0.0f;  /* m/s per axis                                 */
    uint64_t seed = 0;
};

class cSPH {
public:
    sph_scene scn{};    /* raw scene values (Scene)                          */
    sph_scene_info info{}; /* the scene's init volume, emitters, ... (file scenes) */
    sph_params par{};   /* derived constants (Scene::_UpdatePar/_UpdateGrid) */
    InitVolume vol{};
    int device = 0;
    uint32_t flags = 0; /* SPH_FLAG_*                                        */
    bool bInitialized = false;
    int curScene = -1;  /* index into the loaded scene file, -1: scn set by hand */
    uint64_t seed = 0;  /* random fill (initType 1) and random drops         */
    uint32_t emitId = 0; /* next particle the emitters and drops overwrite (App::emitId) */
    int cntRain = 0;     /* frames since the last rain drop (App::cntRain)   */
    uint64_t drops = 0;  /* random drops so far (the generator's draw number) */

    cSPH() { check(sph_scene_default(&scn), "sph_scene_default"); }
    ~cSPH() { _FreeMem(); }
    cSPH(const cSPH&) = delete;
    cSPH& operator=(const cSPH&) = delete;

    /* cSPH::LoadScenes (source/SPH/SPH_Scenes.cpp:53-83): read a scene file
     * and select its startup scene (not yet initialised).                  */
    void LoadScenes(const std::string& path) {
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) throw Error(SPH_ERR_INVALID_ARGUMENT, ("cSPH::LoadScenes " + path).c_str());
        std::string text;
        char buf[65536];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
        std::fclose(f);
        uint32_t count = 0, def = 0;
        check(sph_xml_scene_count(text.data(), text.size(), &count, &def), "sph_xml_scene_count");
        xml_ = std::move(text);
        count_ = count;
        SetScene(def);
    }

    uint32_t SceneCount() const { return count_; }

    /* scn = scenes[index] (the value part of cSPH::UpdScene). */
    void SetScene(uint32_t index) {
        check(sph_xml_scene(xml_.data(), xml_.size(), index, &scn, &info), "sph_xml_scene");
        curScene = (int)index;
    }

    /* cSPH::NextScene / PrevScene (source/SPH/SPH_Scenes.cpp:36-47). */
    void NextScene(bool chapter = false) { Step(+1, chapter); }
    void PrevScene(bool chapter = false) { Step(-1, chapter); }

    /* cSPH::InitScene: derive the constants, allocate, fill the init volume
     * (a file scene with cSPH::Reset(initType), else the `vol` lattice). */
    void InitScene() {
        check(sph_params_from_scene(&scn, &par), "sph_params_from_scene");
        _FreeMem();
        _InitMem();
        Reset(curScene >= 0 ? info.init_type : 0);
    }

    void _InitMem() {
        check(sph_create(&par, device, flags, &ctx_), "sph_create");
        const size_t n = par.num_particles;
        hPos_.assign(4 * n, 0.0f);
        hVel_.assign(4 * n, 0.0f);
        hDens_.assign(n, 0.0f);
        bInitialized = true;
    }

    void _FreeMem() {
        if (ctx_) sph_destroy(ctx_);
        ctx_ = nullptr;
        bInitialized = false;
    }

    /* A file scene: its init volume or random fill (sph_scene_fill, any
     * type).  Otherwise type 0: the lattice block of `vol` (its size must be
     * the scene's particle count).                                         */
    void Reset(int type) {
        require();
        if (curScene >= 0) {
            sph_scene_info in = info;
            in.init_type = type;
            check(sph_scene_fill(&scn, &in, seed, hPos_.data(), hVel_.data()), "sph_scene_fill");
            setArray(SPH_ARRAY_POSITION, hPos_.data(), 0, (int)par.num_particles);
            setArray(SPH_ARRAY_VELOCITY, hVel_.data(), 0, (int)par.num_particles);
            return;
        }
        if (type != 0) throw Error(SPH_ERR_UNSUPPORTED, "cSPH::Reset");
        if ((uint64_t)vol.nx * vol.ny * vol.nz != par.num_particles)
            throw Error(SPH_ERR_INVALID_ARGUMENT, "cSPH::Reset (volume size)");
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        check(sph_scene_lattice(vol.nx, vol.ny, vol.nz, vol.min, vol.spacing, vol.jitter, vol.vel_sigma,
                                zero, vol.seed, 0, hPos_.data(), hVel_.data()),
              "sph_scene_lattice");
        setArray(SPH_ARRAY_POSITION, hPos_.data(), 0, (int)par.num_particles);
        setArray(SPH_ARRAY_VELOCITY, hVel_.data(), 0, (int)par.num_particles);
    }

    /* The particle part of App::UpdateEmitter (source/App/Update.cpp:62-97),
     * run before each frame's Update: every emitter writes its block from
     * emitId on, and every info.rain frames a drop falls at a random place.
     * emitId wraps at the particle count; a block that crosses the end is
     * split there (the reference writes past the end of its arrays).      */
    void UpdateEmitter() {
        require();
        for (uint32_t e = 0; e < 4; ++e) {
            const sph_emitter& em = info.emit[e];
            if (em.size <= 0) continue;
            const int64_t need = (int64_t)em.size * (em.size2 == 0 ? em.size : em.size2);
            if (need <= 0 || need > SPH_MAX_EMIT) throw Error(SPH_ERR_CAPACITY, "cSPH::UpdateEmitter");
            emPos_.resize(4 * (size_t)need);
            emVel_.resize(4 * (size_t)need);
            uint32_t cnt = 0;
            check(sph_scene_emit(&info, e, emPos_.data(), emVel_.data(), &cnt), "sph_scene_emit");
            PutWrapped(emPos_.data(), emVel_.data(), cnt);
        }
        if (info.rain > 0 && ++cntRain >= info.rain) {
            cntRain = 0;
            Drop(true);
        }
    }

    /* cSPH::Drop (source/SPH/SPH_Init.cpp:82-118): a ball of particles at
     * rest from emitId on (those below the particle count).               */
    void Drop(bool random) {
        require();
        uint32_t cnt = 0;
        check(sph_scene_drop(&scn, &info, random ? 1 : 0, seed, random ? drops++ : 0, emitId, hPos_.data(),
                             hVel_.data(), &cnt),
              "sph_scene_drop");
        if (cnt == 0) return;
        setArray(SPH_ARRAY_POSITION, hPos_.data(), (int)emitId, (int)cnt);
        setArray(SPH_ARRAY_VELOCITY, hVel_.data(), (int)emitId, (int)cnt);
        emitId += cnt;
        if (emitId >= par.num_particles) emitId -= par.num_particles;
    }

    /* One step (cSPH::Update, source/SPH/SPH_Update.cpp:12-81), or several. */
    void Update(uint32_t nsteps = 1) {
        require();
        check(sph_step(ctx_, nsteps), "sph_step");
    }

    /* float4 per particle in original order (source/SPH/SPH_Util.cpp:44-57):
     * which = 0 positions, 1 velocities.                                   */
    float* getArray(int which) {
        require();
        std::vector<float>& h = which == SPH_ARRAY_POSITION ? hPos_ : hVel_;
        check(sph_get_array(ctx_, which, h.data()), "sph_get_array");
        return h.data();
    }

    /* source/SPH/SPH_Util.cpp:59-71: count float4 records from `start`. */
    void setArray(int which, const float* data, int start, int count) {
        require();
        check(sph_set_array(ctx_, which, data, (uint32_t)start, (uint32_t)count), "sph_set_array");
    }

    /* Density per particle (original order) of the last step. */
    float* getDensity() {
        require();
        check(sph_get_array(ctx_, SPH_ARRAY_DENSITY, hDens_.data()), "sph_get_array");
        return hDens_.data();
    }

    sph_stats Stats() {
        require();
        sph_stats s{};
        check(sph_get_stats(ctx_, &s), "sph_get_stats");
        return s;
    }

    sph_ctx* context() const { return ctx_; }

private:
    void require() const {
        if (!bInitialized) throw Error(SPH_ERR_NOT_INITIALISED, "cSPH");
    }

    // cnt records at emitId, wrapping at the particle count; emitId advances.
    void PutWrapped(const float* pos, const float* vel, uint32_t cnt) {
        const uint32_t n = par.num_particles;
        uint32_t done = 0;
        while (done < cnt) {
            const uint32_t part = std::min(cnt - done, n - emitId);
            setArray(SPH_ARRAY_POSITION, pos + 4 * (size_t)done, (int)emitId, (int)part);
            setArray(SPH_ARRAY_VELOCITY, vel + 4 * (size_t)done, (int)emitId, (int)part);
            done += part;
            emitId += part;
            if (emitId >= n) emitId -= n;
        }
    }

    void Step(int dir, bool chapter) {
        if (count_ == 0) throw Error(SPH_ERR_NOT_INITIALISED, "cSPH: no scene file");
        int k = curScene < 0 ? 0 : curScene;
        sph_scene s{};
        sph_scene_info in{};
        for (uint32_t tries = 0; tries < count_; ++tries) {
            k = (k + dir + (int)count_) % (int)count_;
            check(sph_xml_scene(xml_.data(), xml_.size(), (uint32_t)k, &s, &in), "sph_xml_scene");
            if (!chapter || in.chapter) break;
        }
        SetScene((uint32_t)k);
        InitScene();
    }

    sph_ctx* ctx_ = nullptr;
    std::vector<float> hPos_, hVel_, hDens_, emPos_, emVel_;
    std::string xml_;
    uint32_t count_ = 0;
};

}  // namespace pibiti

#endif /* PIBITI_SPH_HPP */
