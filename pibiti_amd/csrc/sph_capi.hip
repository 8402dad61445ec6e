// sph_capi.hip - extern "C" entry points of include/pibiti_sph.h for single
// scenes and batches; the device context is sph::Engine (sph_engine.hpp).
#include <atomic>
#include <cstring>

#include "sph_engine.hpp"

using namespace sph;

namespace sph {
namespace {
// sph_debug_fail_alloc: the allocation that takes the count from 1 to 0 fails.
std::atomic<uint32_t> fail_alloc_countdown{0};
}  // namespace

hipError_t device_malloc(void** ptr, size_t bytes) {
    uint32_t k = fail_alloc_countdown.load();
    while (k > 0 && !fail_alloc_countdown.compare_exchange_weak(k, k - 1)) {
    }
    if (k == 1) {
        *ptr = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(ptr, bytes);
}
}  // namespace sph

extern "C" {

// ---------------------------------------------------------------- one scene
int sph_create(const sph_params* params, int device, uint32_t flags, sph_ctx** out) {
    if (!out) return SPH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int st = validate(params);
    if (st != SPH_OK) return st;
    if ((1ull << params->key_bits) < params->num_cells || params->key_bits > 31)
        return SPH_ERR_INVALID_PARAMS;
    sph_ctx* c = new (std::nothrow) sph_ctx();
    if (!c) return SPH_ERR_OUT_OF_MEMORY;
    sph_params p = *params;
    p.cell_base = 0;
    st = c->eng.init(&p, 1, device, flags, false);
    if (st != SPH_OK) {
        {
            DeviceGuard guard(c->eng.device);
            c->eng.release();
        }
        delete c;
        return st;
    }
    *out = c;
    return SPH_OK;
}

int sph_destroy(sph_ctx* c) {
    if (!c) return SPH_OK;
    {
        DeviceGuard guard(c->eng.device);
        if (c->eng.stream) (void)hipStreamSynchronize(c->eng.stream);
        c->eng.release();
    }
    delete c;
    return SPH_OK;
}

int sph_set_params(sph_ctx* c, const sph_params* params) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    if (!params) return SPH_ERR_INVALID_ARGUMENT;
    sph_params p = *params;
    p.cell_base = 0;
    return c->eng.set_params(0, &p);
}

int sph_set_array(sph_ctx* c, int which, const float* data, uint32_t first, uint32_t count) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.set_array(0, which, data, first, count);
}

int sph_get_array(sph_ctx* c, int which, float* out) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.get_array(0, which, out);
}

int sph_step(sph_ctx* c, uint32_t nsteps) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.step(nsteps);
}

int sph_get_grid(sph_ctx* c, uint32_t* sorted_hash, uint32_t* perm, uint32_t* cell_start,
                 uint32_t* cell_end) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.get_grid(0, sorted_hash, perm, cell_start, cell_end);
}

int sph_get_stats(sph_ctx* c, sph_stats* out) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.get_stats(0, out);
}

int sph_get_device_view(sph_ctx* c, sph_device_view* out) {
    if (!c || !out) return SPH_ERR_INVALID_ARGUMENT;
    const Engine& e = c->eng;
    if (!e.ready()) return SPH_ERR_NOT_INITIALISED;
    std::memset(out, 0, sizeof(*out));
    out->pos4 = reinterpret_cast<const float*>(e.pos[e.cur_pos]);
    out->vel4 = reinterpret_cast<const float*>(e.vel[e.cur_vel]);
    out->colour4 = reinterpret_cast<const float*>(e.colour);
    out->index = e.pid[e.cur_pid];
    out->stream = e.stream;
    out->n = e.cfg.n;
    out->device = e.device;
    return SPH_OK;
}

int sph_synchronize(sph_ctx* c) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.synchronize();
}

int sph_debug_poke(sph_ctx* c, int what, uint32_t index, uint32_t value) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.debug_poke(what, index, value);
}

int sph_debug_fail_alloc(uint32_t after) {
    fail_alloc_countdown.store(after);
    return SPH_OK;
}

int sph_set_timing(sph_ctx* c, int enable) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    c->eng.timing = enable != 0;
    return SPH_OK;
}

int sph_get_timing(sph_ctx* c, float* ms6) {
    if (!c || !ms6) return SPH_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < 6; ++k) ms6[k] = c->eng.last_ms[k];
    return SPH_OK;
}

int sph_capture(sph_ctx* c, uint32_t nsteps) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.capture(nsteps);
}

int sph_replay(sph_ctx* c, uint32_t times) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return c->eng.replay(times);
}

// -------------------------------------------------------------------- batch
int sph_batch_create(const sph_params* scenes, uint32_t nscenes, int device, uint32_t flags,
                     sph_batch** out) {
    if (!out || !scenes || nscenes == 0) return SPH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    for (uint32_t s = 0; s < nscenes; ++s) {
        int st = validate(&scenes[s]);
        if (st != SPH_OK) return st;
    }
    sph_batch* b = new (std::nothrow) sph_batch();
    if (!b) return SPH_ERR_OUT_OF_MEMORY;
    int st = b->eng.init(scenes, nscenes, device, flags, true);
    if (st != SPH_OK) {
        {
            DeviceGuard guard(b->eng.device);
            b->eng.release();
        }
        delete b;
        return st;
    }
    *out = b;
    return SPH_OK;
}

int sph_batch_destroy(sph_batch* b) {
    if (!b) return SPH_OK;
    {
        DeviceGuard guard(b->eng.device);
        if (b->eng.stream) (void)hipStreamSynchronize(b->eng.stream);
        b->eng.release();
    }
    delete b;
    return SPH_OK;
}

int sph_batch_get_params(sph_batch* b, uint32_t scene, sph_params* out) {
    if (!b || !out || scene >= b->eng.cfg.nscenes) return SPH_ERR_INVALID_ARGUMENT;
    *out = b->eng.scenes[scene];
    return SPH_OK;
}

int sph_batch_set_params(sph_batch* b, uint32_t scene, const sph_params* params) {
    if (!b || !params || scene >= b->eng.cfg.nscenes) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.set_params(scene, params);
}

int sph_batch_set_array(sph_batch* b, uint32_t scene, int which, const float* data, uint32_t first,
                        uint32_t count) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.set_array(scene, which, data, first, count);
}

int sph_batch_get_array(sph_batch* b, uint32_t scene, int which, float* out) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.get_array(scene, which, out);
}

int sph_batch_get_grid(sph_batch* b, uint32_t scene, uint32_t* sorted_hash, uint32_t* perm, uint32_t* cell_start,
                       uint32_t* cell_end) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.get_grid(scene, sorted_hash, perm, cell_start, cell_end);
}

int sph_batch_step(sph_batch* b, uint32_t nsteps) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.step(nsteps);
}

int sph_batch_capture(sph_batch* b, uint32_t nsteps) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.capture(nsteps);
}

int sph_batch_replay(sph_batch* b, uint32_t times) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.replay(times);
}

int sph_batch_get_stats(sph_batch* b, uint32_t scene, sph_stats* out) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.get_stats(scene, out);
}

int sph_batch_synchronize(sph_batch* b) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.synchronize();
}

int sph_batch_set_tune(sph_batch* b, const uint32_t tune[4]) {
    if (!b || !tune) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.set_tune(tune);
}

int sph_batch_autotune(sph_batch* b, const char* cache_path, uint32_t tune_out[4], float* best_ms) {
    if (!b || !tune_out) return SPH_ERR_INVALID_ARGUMENT;
    return b->eng.autotune(cache_path, tune_out, best_ms);
}

int sph_batch_info(sph_batch* b, uint32_t* nscenes, uint32_t* total_particles,
                   uint32_t* total_cells, uint32_t* key_bits) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    if (nscenes) *nscenes = b->eng.cfg.nscenes;
    if (total_particles) *total_particles = b->eng.cfg.n;
    if (total_cells) *total_cells = b->eng.cfg.num_cells;
    if (key_bits) *key_bits = b->eng.cfg.key_bits;
    return SPH_OK;
}

}  // extern "C"
