// sph_capi.hip - device context behind the C ABI of include/pibiti_sph.h.
//
// The context replaces what cSPH::_InitMem/_FreeMem (source/SPH/SPH_Mem.cpp:11-82)
// allocate and what cSPH::Update (source/SPH/SPH_Update.cpp:12-81) launches,
// without GL buffers and without a device-wide sync after every kernel: a
// step is a chain of kernels on the context's own stream.
#include <cstring>
#include <new>

#include "sph_internal.hpp"

struct sph_ctx {
    int device = 0;
    uint32_t flags = 0;
    sph_params p{};
    hipStream_t stream = nullptr;

    float4* pos[2] = {nullptr, nullptr};
    float4* vel[2] = {nullptr, nullptr};
    uint32_t* pid[2] = {nullptr, nullptr};
    float2* rhop = nullptr;
    float4* dvel = nullptr;
    uint32_t* key_pid = nullptr;
    uint32_t* val_pid = nullptr;
    uint32_t* key_tmp = nullptr;
    uint32_t* val_tmp = nullptr;
    uint32_t* cell_start = nullptr;
    uint32_t* cell_end = nullptr;
    uint32_t* counters = nullptr;
    uint32_t* cell_max = nullptr;
    double* partials = nullptr;
    double* stats = nullptr;
    float* dt_dev = nullptr;
    float4* staging = nullptr;
    void* sort_scratch = nullptr;
    sph::RadixSortPlan sort{};

    int cur_pos = 0, cur_vel = 0, cur_pid = 0;
    const uint32_t* last_skey = nullptr;
    bool have_pos = false, have_vel = false, stepped = false;
    uint64_t steps = 0;
    double host_clock = 0.0;

    bool timing = false;
    hipEvent_t ev[6] = {};
    float last_ms[6] = {0, 0, 0, 0, 0, 0};

    sph::StepBuffers buffers() const {
        sph::StepBuffers b{};
        b.pos_cur = pos[cur_pos];
        b.vel_cur = vel[cur_vel];
        b.pid_cur = pid[cur_pid];
        b.pos_alt = pos[cur_pos ^ 1];
        b.vel_alt = vel[cur_vel ^ 1];
        b.pid_alt = pid[cur_pid ^ 1];
        b.rhop = rhop;
        b.dvel = dvel;
        b.key_pid = key_pid;
        b.val_pid = val_pid;
        b.key_tmp = key_tmp;
        b.val_tmp = val_tmp;
        b.cell_start = cell_start;
        b.cell_end = cell_end;
        b.counters = counters;
        b.partials = partials;
        b.stats = stats;
        b.dt_dev = dt_dev;
        return b;
    }
};

namespace {

constexpr int kNumStats = 8;

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

int to_status(hipError_t e) {
    if (e == hipSuccess) return SPH_OK;
    if (e == hipErrorOutOfMemory) return SPH_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SPH_ERR_NO_DEVICE;
    return SPH_ERR_DEVICE;
}

#define SPH_TRY(expr)                              \
    do {                                           \
        hipError_t e_ = (expr);                    \
        if (e_ != hipSuccess) return to_status(e_); \
    } while (0)

template <typename T>
hipError_t dalloc(T** ptr, size_t count) {
    return hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T) + 16);
}

void free_all(sph_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        (void)hipFree(c->pos[k]);
        (void)hipFree(c->vel[k]);
        (void)hipFree(c->pid[k]);
    }
    void* ptrs[] = {c->rhop,      c->dvel,       c->key_pid,  c->val_pid,  c->key_tmp,
                    c->val_tmp,   c->cell_start, c->cell_end, c->counters, c->cell_max,
                    c->partials,  c->stats,      c->dt_dev,   c->staging,  c->sort_scratch};
    for (void* q : ptrs) (void)hipFree(q);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

bool params_compatible(const sph_params& a, const sph_params& b) {
    return a.num_particles == b.num_particles && a.num_cells == b.num_cells &&
           a.grid_size[0] == b.grid_size[0] && a.grid_size[1] == b.grid_size[1] &&
           a.grid_size[2] == b.grid_size[2];
}

int validate(const sph_params* p) {
    if (!p) return SPH_ERR_INVALID_ARGUMENT;
    if (p->num_particles == 0 || p->num_cells == 0 || p->max_par_in_cell == 0 || !(p->h > 0.0f) ||
        !(p->time_step > 0.0f) || p->key_bits > 31 || p->bnd_type != SPH_BND_BOX ||
        p->bnd_eff_z != SPH_BND_EFF_NONE || p->eos > SPH_EOS_TAIT || p->precision > SPH_PREC_KAHAN)
        return SPH_ERR_INVALID_PARAMS;
    if ((uint64_t)p->grid_size[0] * p->grid_size[1] * p->grid_size[2] != p->num_cells ||
        p->grid_size_yx != p->grid_size[0] * p->grid_size[1])
        return SPH_ERR_INVALID_PARAMS;
    if ((1ull << p->key_bits) < p->num_cells) return SPH_ERR_INVALID_PARAMS;
    return SPH_OK;
}

}  // namespace

extern "C" {

int sph_create(const sph_params* params, int device, uint32_t flags, sph_ctx** out) {
    if (!out) return SPH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int st = validate(params);
    if (st != SPH_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPH_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return SPH_ERR_NO_DEVICE;
    DeviceGuard guard(device);

    sph_ctx* c = new (std::nothrow) sph_ctx();
    if (!c) return SPH_ERR_OUT_OF_MEMORY;
    c->device = device;
    c->flags = flags;
    c->p = *params;
    const size_t n = params->num_particles;
    const size_t ncell = params->num_cells;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = dalloc(&c->pos[k], n);
        if (e == hipSuccess) e = dalloc(&c->vel[k], n);
        if (e == hipSuccess) e = dalloc(&c->pid[k], n);
    }
    if (e == hipSuccess) e = dalloc(&c->rhop, n);
    if (e == hipSuccess && (flags & SPH_FLAG_KEEP_DVEL)) e = dalloc(&c->dvel, n);
    if (e == hipSuccess) e = dalloc(&c->key_pid, n);
    if (e == hipSuccess) e = dalloc(&c->val_pid, n);
    if (e == hipSuccess) e = dalloc(&c->key_tmp, n);
    if (e == hipSuccess) e = dalloc(&c->val_tmp, n);
    if (e == hipSuccess) e = dalloc(&c->cell_start, ncell);
    if (e == hipSuccess) e = dalloc(&c->cell_end, ncell);
    if (e == hipSuccess) e = dalloc(&c->counters, (size_t)sph::kCntNum);
    if (e == hipSuccess) e = dalloc(&c->cell_max, (size_t)1);
    if (e == hipSuccess) e = dalloc(&c->partials, (size_t)sph::kPNum * sph::stats_num_blocks((uint32_t)n));
    if (e == hipSuccess) e = dalloc(&c->stats, (size_t)kNumStats);
    if (e == hipSuccess) e = dalloc(&c->dt_dev, (size_t)1);
    if (e == hipSuccess) e = dalloc(&c->staging, n);
    if (e == hipSuccess) e = hipMalloc(&c->sort_scratch, sph::radix_scratch_bytes((uint32_t)n));
    if (e == hipSuccess) e = sph::radix_plan_init(c->sort, (uint32_t)n, c->sort_scratch, c->stream);
    for (int k = 0; k < 6 && e == hipSuccess; ++k) e = hipEventCreate(&c->ev[k]);
    if (e == hipSuccess) e = sph::launch_iota_u32(c->pid[0], (uint32_t)n, c->stream);
    if (e == hipSuccess) e = sph::launch_fill_u32(c->cell_start, sph::kEmptyCell, ncell, c->stream);
    if (e == hipSuccess) e = sph::launch_fill_u32(c->cell_end, sph::kEmptyCell, ncell, c->stream);
    if (e == hipSuccess) e = sph::launch_fill_u32(c->counters, 0u, sph::kCntNum, c->stream);
    if (e == hipSuccess) e = sph::launch_fill_u32(reinterpret_cast<uint32_t*>(c->rhop), 0u, 2 * n, c->stream);
    if (e == hipSuccess) e = sph::launch_fill_u32(reinterpret_cast<uint32_t*>(c->stats), 0u, 2 * kNumStats, c->stream);
    if (e == hipSuccess && c->dvel) e = sph::launch_fill_u32(reinterpret_cast<uint32_t*>(c->dvel), 0u, 4 * n, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->dt_dev, &c->p.time_step, sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        free_all(c);
        delete c;
        return to_status(e);
    }
    *out = c;
    return SPH_OK;
}

int sph_destroy(sph_ctx* c) {
    if (!c) return SPH_OK;
    DeviceGuard guard(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return SPH_OK;
}

int sph_set_params(sph_ctx* c, const sph_params* params) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    int st = validate(params);
    if (st != SPH_OK) return st;
    if (!params_compatible(c->p, *params)) return SPH_ERR_INVALID_PARAMS;
    DeviceGuard guard(c->device);
    c->p = *params;
    SPH_TRY(hipMemcpyAsync(c->dt_dev, &c->p.time_step, sizeof(float), hipMemcpyHostToDevice, c->stream));
    SPH_TRY(hipStreamSynchronize(c->stream));
    return SPH_OK;
}

int sph_set_array(sph_ctx* c, int which, const float* data, uint32_t first, uint32_t count) {
    if (!c || !data) return SPH_ERR_INVALID_ARGUMENT;
    if (which != SPH_ARRAY_POSITION && which != SPH_ARRAY_VELOCITY) return SPH_ERR_INVALID_ARGUMENT;
    const uint32_t n = c->p.num_particles;
    if (first > n || count > n - first) return SPH_ERR_INVALID_ARGUMENT;
    if (count == 0) return SPH_OK;
    DeviceGuard guard(c->device);
    SPH_TRY(hipMemcpyAsync(c->staging, data, (size_t)count * sizeof(float4), hipMemcpyHostToDevice,
                           c->stream));
    float4* dst = which == SPH_ARRAY_POSITION ? c->pos[c->cur_pos] : c->vel[c->cur_vel];
    SPH_TRY(sph::launch_update_from_original(n, c->pid[c->cur_pid], c->staging, first, count, dst,
                                             c->stream));
    SPH_TRY(hipStreamSynchronize(c->stream));
    if (first == 0 && count == n) {
        if (which == SPH_ARRAY_POSITION)
            c->have_pos = true;
        else
            c->have_vel = true;
    }
    return SPH_OK;
}

int sph_get_array(sph_ctx* c, int which, float* out) {
    if (!c || !out) return SPH_ERR_INVALID_ARGUMENT;
    const uint32_t n = c->p.num_particles;
    DeviceGuard guard(c->device);
    const uint32_t* pid = c->pid[c->cur_pid];
    hipError_t e = hipSuccess;
    size_t bytes = 0;
    switch (which) {
        case SPH_ARRAY_POSITION:
            e = sph::launch_scatter_to_original(n, pid, c->pos[c->cur_pos], c->staging, 4, 4, c->stream);
            bytes = (size_t)n * 16;
            break;
        case SPH_ARRAY_VELOCITY:
            e = sph::launch_scatter_to_original(n, pid, c->vel[c->cur_vel], c->staging, 4, 4, c->stream);
            bytes = (size_t)n * 16;
            break;
        case SPH_ARRAY_DENSITY:
            e = sph::launch_scatter_to_original(n, pid, c->rhop, c->staging, 1, 2, c->stream);
            bytes = (size_t)n * 4;
            break;
        case SPH_ARRAY_PRESSURE:
            e = sph::launch_scatter_to_original(n, pid, reinterpret_cast<float*>(c->rhop) + 1,
                                                c->staging, 1, 2, c->stream);
            bytes = (size_t)n * 4;
            break;
        case SPH_ARRAY_DVEL:
            if (!c->dvel) return SPH_ERR_UNSUPPORTED;
            e = sph::launch_scatter_to_original(n, pid, c->dvel, c->staging, 4, 4, c->stream);
            bytes = (size_t)n * 16;
            break;
        default:
            return SPH_ERR_INVALID_ARGUMENT;
    }
    SPH_TRY(e);
    SPH_TRY(hipMemcpyAsync(out, c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
    SPH_TRY(hipStreamSynchronize(c->stream));
    return SPH_OK;
}

int sph_step(sph_ctx* c, uint32_t nsteps) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    if (!c->have_pos || !c->have_vel) return SPH_ERR_NOT_INITIALISED;
    DeviceGuard guard(c->device);
    const sph_params& p = c->p;
    const bool keep = (c->flags & SPH_FLAG_KEEP_DVEL) != 0;
    const bool cfl = p.cfl > 0.0f;
    for (uint32_t s = 0; s < nsteps; ++s) {
        const bool tm = c->timing && s + 1 == nsteps;
        if (tm) SPH_TRY(hipEventRecord(c->ev[0], c->stream));
        if (cfl) SPH_TRY(sph::launch_stats(p, c->buffers(), true, c->stream));
        SPH_TRY(sph::launch_fill_u32(c->counters, 0u, sph::kCntNum, c->stream));
        SPH_TRY(sph::launch_integrate_hash(p, c->buffers(), c->stream));
        if (tm) SPH_TRY(hipEventRecord(c->ev[1], c->stream));
        uint32_t *skey = nullptr, *oslot = nullptr;
        SPH_TRY(sph::radix_sort_pairs(c->sort, p.key_bits, c->key_pid, c->val_pid, c->key_tmp,
                                      c->val_tmp, &skey, &oslot, c->stream));
        if (tm) SPH_TRY(hipEventRecord(c->ev[2], c->stream));
        SPH_TRY(sph::launch_reorder(p, c->buffers(), skey, oslot, c->stream));
        c->cur_pos ^= 1;
        c->cur_vel ^= 1;
        c->cur_pid ^= 1;
        c->last_skey = skey;
        if (tm) SPH_TRY(hipEventRecord(c->ev[3], c->stream));
        SPH_TRY(sph::launch_density(p, c->buffers(), c->stream));
        if (tm) SPH_TRY(hipEventRecord(c->ev[4], c->stream));
        sph::StepBuffers b = c->buffers();
        SPH_TRY(sph::launch_force(p, b, b.vel_alt, keep, c->stream));
        c->cur_vel ^= 1;
        if (tm) SPH_TRY(hipEventRecord(c->ev[5], c->stream));
        ++c->steps;
        if (!cfl) c->host_clock += (double)p.time_step;
        c->stepped = true;
        if (tm) {
            SPH_TRY(hipEventSynchronize(c->ev[5]));
            for (int k = 0; k < 5; ++k) SPH_TRY(hipEventElapsedTime(&c->last_ms[k], c->ev[k], c->ev[k + 1]));
            SPH_TRY(hipEventElapsedTime(&c->last_ms[5], c->ev[0], c->ev[5]));
        }
    }
    return SPH_OK;
}

int sph_get_grid(sph_ctx* c, uint32_t* sorted_hash, uint32_t* perm, uint32_t* cell_start,
                 uint32_t* cell_end) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    if (!c->stepped) return SPH_ERR_NOT_INITIALISED;
    DeviceGuard guard(c->device);
    const size_t n = c->p.num_particles, ncell = c->p.num_cells;
    if (sorted_hash)
        SPH_TRY(hipMemcpyAsync(sorted_hash, c->last_skey, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (perm) SPH_TRY(hipMemcpyAsync(perm, c->pid[c->cur_pid], n * 4, hipMemcpyDeviceToHost, c->stream));
    if (cell_start)
        SPH_TRY(hipMemcpyAsync(cell_start, c->cell_start, ncell * 4, hipMemcpyDeviceToHost, c->stream));
    if (cell_end)
        SPH_TRY(hipMemcpyAsync(cell_end, c->cell_end, ncell * 4, hipMemcpyDeviceToHost, c->stream));
    SPH_TRY(hipStreamSynchronize(c->stream));
    // cellEnd is only maintained for occupied cells; report empty ones as the
    // reference marks an empty cellStart.
    if (cell_end) {
        if (cell_start) {
            for (size_t k = 0; k < ncell; ++k)
                if (cell_start[k] == sph::kEmptyCell) cell_end[k] = sph::kEmptyCell;
        } else {
            uint32_t* tmp = new (std::nothrow) uint32_t[ncell];
            if (!tmp) return SPH_ERR_OUT_OF_MEMORY;
            hipError_t e = hipMemcpy(tmp, c->cell_start, ncell * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                for (size_t k = 0; k < ncell; ++k)
                    if (tmp[k] == sph::kEmptyCell) cell_end[k] = sph::kEmptyCell;
            delete[] tmp;
            SPH_TRY(e);
        }
    }
    return SPH_OK;
}

int sph_get_stats(sph_ctx* c, sph_stats* out) {
    if (!c || !out) return SPH_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->device);
    std::memset(out, 0, sizeof(*out));
    double st[kNumStats];
    uint32_t cnt[sph::kCntNum];
    uint32_t cmax = 0;
    // Keep the step's time step and clock: recompute only the reductions.
    double saved[2];
    SPH_TRY(hipMemcpyAsync(saved, c->stats + sph::kPNum, sizeof(saved), hipMemcpyDeviceToHost, c->stream));
    SPH_TRY(hipStreamSynchronize(c->stream));
    SPH_TRY(sph::launch_stats(c->p, c->buffers(), false, c->stream));
    SPH_TRY(sph::launch_fill_u32(c->cell_max, 0u, 1, c->stream));
    if (c->stepped) SPH_TRY(sph::launch_cell_stats(c->p, c->buffers(), c->cell_max, c->stream));
    SPH_TRY(hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    SPH_TRY(hipMemcpyAsync(cnt, c->counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    SPH_TRY(hipMemcpyAsync(&cmax, c->cell_max, sizeof(cmax), hipMemcpyDeviceToHost, c->stream));
    uint32_t sort_err = 0;
    SPH_TRY(hipMemcpyAsync(&sort_err, c->sort.err, sizeof(sort_err), hipMemcpyDeviceToHost, c->stream));
    // Restore the time step the kernels use (the finalisation rewrote it from
    // the current state, which is what the next CFL step computes anyway).
    if (c->p.cfl <= 0.0f)
        SPH_TRY(hipMemcpyAsync(c->dt_dev, &c->p.time_step, sizeof(float), hipMemcpyHostToDevice, c->stream));
    SPH_TRY(hipMemcpyAsync(c->stats + sph::kPNum, saved, sizeof(saved), hipMemcpyHostToDevice, c->stream));
    SPH_TRY(hipStreamSynchronize(c->stream));
    out->mass = st[sph::kPMass];
    out->kinetic_energy = st[sph::kPKin];
    out->potential_energy = st[sph::kPPot];
    out->max_speed = st[sph::kPVmax];
    out->dt_cfl = c->p.cfl > 0.0f ? st[sph::kPNum] : (double)c->p.time_step;
    out->sim_time = c->p.cfl > 0.0f ? saved[1] : c->host_clock;
    out->steps = c->steps;
    out->cap_cells = cnt[sph::kCntCapCells];
    out->out_of_grid = cnt[sph::kCntOutOfGrid];
    out->max_cell_count = cmax;
    out->sort_errors = sort_err;
    return SPH_OK;
}

int sph_synchronize(sph_ctx* c) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->device);
    SPH_TRY(hipStreamSynchronize(c->stream));
    return SPH_OK;
}

int sph_set_timing(sph_ctx* c, int enable) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    c->timing = enable != 0;
    return SPH_OK;
}

int sph_get_timing(sph_ctx* c, float* ms6) {
    if (!c || !ms6) return SPH_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < 6; ++k) ms6[k] = c->last_ms[k];
    return SPH_OK;
}

}  // extern "C"
