// sph_engine.hpp - the device context behind the C ABI (one particle array:
// a single scene or a batch of scenes) and the host helpers its entry
// points share.
//
// An Engine owns every device buffer of one particle array and the stream
// its kernels run on.  A single-scene context (sph_ctx) is an Engine with
// one scene whose parameters travel by value; a batch (sph_batch) is an
// Engine whose scenes are packed one after the other in original-index
// order, each with a disjoint cell-key range, parameters in a device table
// and a per-slot scene id.
//
// The Engine replaces what cSPH::_InitMem/_FreeMem (source/SPH/SPH_Mem.cpp:11-82)
// allocate and what cSPH::Update (source/SPH/SPH_Update.cpp:12-81) launches,
// without GL buffers and without a device-wide sync after every kernel: a
// step is a chain of kernels on the engine's own stream, and a run of steps
// can be recorded once into a hipGraph and replayed.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sph_internal.hpp"

namespace sph {
namespace detail {

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

inline int to_status(hipError_t e) {
    if (e == hipSuccess) return SPH_OK;
    if (e == hipErrorOutOfMemory) return SPH_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SPH_ERR_NO_DEVICE;
    return SPH_ERR_DEVICE;
}

#define SPH_TRY(expr)                               \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return to_status(e_); \
    } while (0)

#define SPH_TRY_ST(expr)                 \
    do {                                 \
        int st_ = (expr);                \
        if (st_ != SPH_OK) return st_;   \
    } while (0)

// Device arrays carry 128 bytes past their end: the density kernels' groups
// of position loads (at most eight float4) may run past the last slot.
template <typename T>
inline hipError_t dalloc(T** ptr, size_t count) {
    return device_malloc(reinterpret_cast<void**>(ptr), count * sizeof(T) + 128);
}

// Slot ranges of the sort's 27-bit slot field bound the particle count.
constexpr uint64_t kMaxParticles = 1ull << 27;

inline int validate(const sph_params* p) {
    if (!p) return SPH_ERR_INVALID_ARGUMENT;
    if (p->num_particles == 0 || p->num_particles > kMaxParticles || p->num_cells == 0 ||
        p->max_par_in_cell == 0 || !(p->h > 0.0f) || !(p->time_step > 0.0f) ||
        p->eos > SPH_EOS_TAIT || p->precision > SPH_PREC_FAST || !(p->t_end >= 0.0f))
        return SPH_ERR_INVALID_PARAMS;
    if (p->precision == SPH_PREC_FAST && p->max_par_in_cell > kFastMaxCap) return SPH_ERR_UNSUPPORTED;
    if (check_shape(*p) != SPH_OK) return SPH_ERR_INVALID_PARAMS;
    if ((uint64_t)p->grid_size[0] * p->grid_size[1] * p->grid_size[2] != p->num_cells ||
        p->grid_size_yx != p->grid_size[0] * p->grid_size[1])
        return SPH_ERR_INVALID_PARAMS;
    return SPH_OK;
}

inline uint32_t bits_for(uint64_t cells) {
    uint32_t b = 0;
    while ((1ull << b) < cells) ++b;
    return b;
}

}  // namespace detail

using namespace detail;

struct LearnState;
void learn_release(struct Engine& e);
// Neighbour counts of the last surrogate step into staging (original order).
int learn_neighbour_counts(struct Engine& e);

struct Engine {
    int device = 0;
    uint32_t flags = 0;
    bool batch = false;
    hipStream_t stream = nullptr;
    LaunchCfg cfg{};
    std::vector<sph_params> scenes;     // host copies (cell_base assigned)
    std::vector<uint32_t> scene_off;    // first original index (= slot) of each scene
    sph_params* d_tab = nullptr;
    uint32_t* d_scene_off = nullptr;

    float4* pos[2] = {nullptr, nullptr};
    float4* vel[2] = {nullptr, nullptr};
    uint32_t* pid[2] = {nullptr, nullptr};
    uint32_t* sid[2] = {nullptr, nullptr};
    float2* rhop = nullptr;
    float4* dvel = nullptr;
    uint32_t* masks = nullptr;  // force pair masks (kMaskWords per slot) while cells hold <= 16 candidates
    uint4* fast_nbr = nullptr;  // force pair entries (fast_groups uint4 per slot, group-major), max_cap <= 64
    uint32_t* fast_nnb = nullptr;
    float4* fast_rec = nullptr;  // force records, 2 per slot
    float* fast_xyz = nullptr;   // packed positions, 3 floats per slot
    uint32_t fast_groups = 0;
    float4* colour = nullptr;   // per sorted slot, scenes with clr_type != NONE
    float* dye = nullptr;       // per ORIGINAL index (it follows the particle)
    uint32_t *key_pid = nullptr, *val_pid = nullptr, *key_tmp = nullptr, *val_tmp = nullptr;
    // cell ranges and slot keys, two of each: a step builds the next order
    // from the current one (cur_cell / cur_key pick the current)
    uint32_t* cell_start[2] = {nullptr, nullptr};
    uint32_t* cell_end[2] = {nullptr, nullptr};
    uint32_t* skey_buf[2] = {nullptr, nullptr};
    uint32_t *counters = nullptr, *cell_max = nullptr;
    uint32_t* work = nullptr;
    double *partials = nullptr, *stats = nullptr;
    float* dt_dev = nullptr;
    float4* staging = nullptr;
    void* sort_scratch = nullptr;
    RadixSortPlan sort{};
    void* resort_scratch = nullptr;
    ResortPlan resort{};

    int cur_pos = 0, cur_vel = 0, cur_pid = 0, cur_cell = 0, cur_key = 0;
    const uint32_t* last_skey = nullptr;
    std::vector<uint8_t> have_pos, have_vel;  // per scene: full upload seen
    bool stepped = false;
    bool order_failed = false;  // a step's order could not be built: the state must be uploaded again
    int last_kind = 0;          // what produced the current state: 0 upload, 1 plain step, 2 surrogate step
    bool last_pair_counts = false;  // the last plain step's gathers used the pair entries (fast_nnb holds its counts)
    uint64_t steps = 0;
    double host_clock = 0.0;
    bool device_clock = false;  // clock and time step advanced on the device each step

    bool timing = false;
    hipEvent_t ev[8] = {};
    float last_ms[6] = {0, 0, 0, 0, 0, 0};

    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_steps = 0;

    LearnState* learn = nullptr;  // surrogate training / learned stepping buffers

    StepBuffers buffers() const {
        StepBuffers b{};
        b.pos_cur = pos[cur_pos];
        b.vel_cur = vel[cur_vel];
        b.pid_cur = pid[cur_pid];
        b.sid_cur = sid[cur_pid];
        b.pos_alt = pos[cur_pos ^ 1];
        b.vel_alt = vel[cur_vel ^ 1];
        b.pid_alt = pid[cur_pid ^ 1];
        b.sid_alt = sid[cur_pid ^ 1];
        b.rhop = rhop;
        b.dvel = dvel;
        b.skey = last_skey;
        b.skey_alt = skey_buf[cur_key ^ 1];
        b.key_pid = key_pid;
        b.val_pid = val_pid;
        b.key_tmp = key_tmp;
        b.val_tmp = val_tmp;
        b.cell_start = cell_start[cur_cell];
        b.cell_end = cell_end[cur_cell];
        b.cell_start_alt = cell_start[cur_cell ^ 1];
        b.cell_end_alt = cell_end[cur_cell ^ 1];
        b.counters = counters;
        b.work = work;
        b.partials = partials;
        b.stats = stats;
        b.dt_dev = dt_dev;
        b.masks = cfg.max_cap <= kMaskCap ? masks : nullptr;
        const bool lists = fast_nbr && fast_nnb && fast_rec && fast_list_groups(cfg.max_cap) <= fast_groups;
        b.pairs = lists ? fast_nbr : nullptr;
        b.pair_cnt = lists ? fast_nnb : nullptr;
        b.rec = lists ? fast_rec : nullptr;
        // the place kernel packs the positions for the pair-entry densities
        b.xyz = lists && fast_xyz ? fast_xyz : nullptr;
        return b;
    }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_steps = 0;
    }

    void release() {
        drop_graph();
        learn_release(*this);
        for (int k = 0; k < 2; ++k) {
            (void)hipFree(pos[k]);
            (void)hipFree(vel[k]);
            (void)hipFree(pid[k]);
            (void)hipFree(sid[k]);
            (void)hipFree(cell_start[k]);
            (void)hipFree(cell_end[k]);
            (void)hipFree(skey_buf[k]);
        }
        void* ptrs[] = {rhop,     dvel,     masks,    fast_nbr, fast_nnb, fast_rec, fast_xyz, colour,   dye,      key_pid,  val_pid,    key_tmp, val_tmp,
                        counters, cell_max, work,    partials,
                        stats,    dt_dev,   staging,  sort_scratch, resort_scratch, d_tab, d_scene_off};
        for (void* q : ptrs) (void)hipFree(q);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Scenes must be validated; cell_base / key_bits are assigned here.
    int init(const sph_params* in, uint32_t nscenes, int dev, uint32_t fl, bool as_batch) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPH_ERR_NO_DEVICE;
        if (dev < 0 || dev >= ndev) return SPH_ERR_NO_DEVICE;
        device = dev;
        flags = fl;
        batch = as_batch;
        scenes.assign(in, in + nscenes);
        for (auto& sc : scenes) {
            // a turning rotor or moving wave needs new parameters every step:
            // single scenes only
            if (as_batch && changes_per_step(sc)) return SPH_ERR_UNSUPPORTED;
            derive_shape(sc);
        }
        scene_off.assign(nscenes + 1, 0);
        uint64_t n = 0, cells = 0;
        uint32_t max_cap = 0, max_n = 0;
        for (uint32_t s = 0; s < nscenes; ++s) {
            scene_off[s] = (uint32_t)n;
            scenes[s].cell_base = (uint32_t)cells;
            n += scenes[s].num_particles;
            cells += scenes[s].num_cells;
            max_cap = std::max(max_cap, scenes[s].max_par_in_cell);
            max_n = std::max(max_n, scenes[s].num_particles);
            if (scenes[s].precision != scenes[0].precision) return SPH_ERR_INVALID_PARAMS;
            if (scenes[s].cfl > 0.0f || scenes[s].t_end > 0.0f) device_clock = true;
        }
        if (n > kMaxParticles || cells >= (1ull << 31)) return SPH_ERR_CAPACITY;
        scene_off[nscenes] = (uint32_t)n;
        const uint32_t kb = bits_for(cells);
        if (!as_batch) {
            // a single scene keeps its own key width (it may ask for more bits)
            if ((1ull << scenes[0].key_bits) < cells || scenes[0].key_bits > 31)
                return SPH_ERR_INVALID_PARAMS;
        } else {
            for (auto& s : scenes) s.key_bits = kb;
        }
        cfg.n = (uint32_t)n;
        cfg.num_cells = (uint32_t)cells;
        cfg.key_bits = as_batch ? kb : scenes[0].key_bits;
        cfg.nscenes = nscenes;
        cfg.max_cap = max_cap;
        cfg.max_scene_n = max_n;
        cfg.precision = scenes[0].precision;
        for (int k = 0; k < 4; ++k) cfg.tune[k] = scenes[0].tune[k];
        cfg.one = scenes[0];
        cfg.physics = 0;
        for (const auto& sc : scenes) cfg.physics |= sc.physics;
        have_pos.assign(nscenes, 0);
        have_vel.assign(nscenes, 0);
        batch = as_batch;
        device_clock = device_clock || as_batch;

        DeviceGuard guard(device);
        const size_t N = n, NC = cells;
        hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        for (int k = 0; k < 2 && e == hipSuccess; ++k) {
            e = dalloc(&pos[k], N);
            if (e == hipSuccess) e = dalloc(&vel[k], N);
            if (e == hipSuccess) e = dalloc(&pid[k], N);
            if (e == hipSuccess && as_batch) e = dalloc(&sid[k], N);
        }
        if (e == hipSuccess) e = dalloc(&rhop, N);
        if (e == hipSuccess && max_cap <= kMaskCap) e = dalloc(&masks, N * kMaskWords);
        if (e == hipSuccess && (flags & SPH_FLAG_KEEP_DVEL)) e = dalloc(&dvel, N);
        if (e == hipSuccess) e = dalloc(&key_pid, N);
        if (e == hipSuccess) e = dalloc(&val_pid, N);
        if (e == hipSuccess) e = dalloc(&key_tmp, N);
        if (e == hipSuccess) e = dalloc(&val_tmp, N);
        // cell tables padded to a multiple of four cells (16-byte accesses)
        for (int k = 0; k < 2 && e == hipSuccess; ++k) {
            e = dalloc(&cell_start[k], (NC + 3) & ~(size_t)3);
            if (e == hipSuccess) e = dalloc(&cell_end[k], (NC + 3) & ~(size_t)3);
            if (e == hipSuccess) e = dalloc(&skey_buf[k], N);
        }
        if (e == hipSuccess) e = dalloc(&counters, (size_t)kCntNum);
        if (e == hipSuccess) e = dalloc(&cell_max, (size_t)1);
        if (e == hipSuccess) e = dalloc(&work, (size_t)4);
        if (e == hipSuccess) e = dalloc(&partials, stats_partials_count(cfg));
        if (e == hipSuccess) e = dalloc(&stats, (size_t)kNumStats * nscenes);
        if (e == hipSuccess) e = dalloc(&dt_dev, (size_t)nscenes);
        if (e == hipSuccess) e = dalloc(&staging, N);
        if (e == hipSuccess) e = dalloc(&d_scene_off, (size_t)nscenes + 1);
        if (e == hipSuccess && as_batch) e = dalloc(&d_tab, (size_t)nscenes);
        if (e == hipSuccess) e = device_malloc(&sort_scratch, radix_scratch_bytes((uint32_t)N));
        if (e == hipSuccess) e = radix_plan_init(sort, (uint32_t)N, sort_scratch, stream);
        if (e == hipSuccess) e = device_malloc(&resort_scratch, resort_scratch_bytes((uint32_t)N, (uint32_t)NC));
        if (e == hipSuccess) e = resort_plan_init(resort, (uint32_t)N, (uint32_t)NC, resort_scratch, stream);
        // one error word for the sort and the re-sort
        sort.err = resort.err;
        for (int k = 0; k < 8 && e == hipSuccess; ++k) e = hipEventCreate(&ev[k]);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_scene_off, scene_off.data(), (nscenes + 1) * sizeof(uint32_t),
                               hipMemcpyHostToDevice, stream);
        if (e == hipSuccess && as_batch)
            e = hipMemcpyAsync(d_tab, scenes.data(), nscenes * sizeof(sph_params),
                               hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = launch_iota_u32(pid[0], (uint32_t)N, stream);
        if (e == hipSuccess && as_batch)
            e = launch_scene_ids((uint32_t)N, nscenes, d_scene_off, pid[0], sid[0], stream);
        for (int k = 0; k < 2 && e == hipSuccess; ++k) {
            e = launch_fill_u32(cell_start[k], kEmptyCell, NC, stream);
            if (e == hipSuccess) e = launch_fill_u32(cell_end[k], kEmptyCell, NC, stream);
            if (e == hipSuccess) e = launch_fill_u32(skey_buf[k], 0u, N, stream);
        }
        last_skey = skey_buf[0];
        if (e == hipSuccess) e = launch_fill_u32(counters, 0u, kCntNum, stream);
        if (e == hipSuccess) e = launch_fill_u32(work, 0u, 4, stream);
        if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(rhop), 0u, 2 * N, stream);
        if (e == hipSuccess)
            e = launch_fill_u32(reinterpret_cast<uint32_t*>(stats), 0u, 2 * (size_t)kNumStats * nscenes,
                                stream);
        if (e == hipSuccess && dvel)
            e = launch_fill_u32(reinterpret_cast<uint32_t*>(dvel), 0u, 4 * N, stream);
        if (e == hipSuccess) e = ensure_colour();
        if (e == hipSuccess) e = ensure_fast();
        if (e == hipSuccess) {
            std::vector<float> dts(nscenes);
            for (uint32_t s = 0; s < nscenes; ++s) dts[s] = scenes[s].time_step;
            e = hipMemcpyAsync(dt_dev, dts.data(), nscenes * sizeof(float), hipMemcpyHostToDevice,
                               stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
        }
        return to_status(e);
    }

    // Colour output: the colour and dye buffers, and the force increment the
    // colours read, for any scene with a colour mode.
    hipError_t ensure_colour() {
        cfg.colour = 0;
        for (const auto& sc : scenes)
            if (sc.clr_type != SPH_CLR_NONE) cfg.colour = 1;
        if (!cfg.colour || (colour && dye && dvel)) return hipSuccess;
        const size_t N = cfg.n;
        hipError_t e = colour ? hipSuccess : dalloc(&colour, N);
        if (e == hipSuccess && !dye) e = dalloc(&dye, N);
        if (e == hipSuccess && !dvel) e = dalloc(&dvel, N);
        if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(colour), 0u, 4 * N, stream);
        if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(dye), 0u, N, stream);
        if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(dvel), 0u, 4 * N, stream);
        return e;
    }

    // The force pair lists and records (FAST precision, and the reference
    // order's default gathers), sized for the largest cap; none above
    // kFastMaxCap (the reference order then takes other gathers).
    hipError_t ensure_fast() {
        if (cfg.max_cap > kFastMaxCap) return hipSuccess;
        const uint32_t g = fast_list_groups(cfg.max_cap);
        if (fast_nbr && g <= fast_groups) return hipSuccess;
        (void)hipFree(fast_nbr);
        fast_nbr = nullptr;
        fast_groups = 0;
        hipError_t e = dalloc(&fast_nbr, (size_t)g * cfg.n);
        if (e == hipSuccess && !fast_nnb) e = dalloc(&fast_nnb, (size_t)cfg.n);
        if (e == hipSuccess && !fast_rec) e = dalloc(&fast_rec, 2 * (size_t)cfg.n);
        if (e == hipSuccess && !fast_xyz) e = dalloc(&fast_xyz, 3 * (size_t)cfg.n);
        if (e == hipSuccess) fast_groups = g;
        return e;
    }

    LaunchCfg launch_cfg() const {
        LaunchCfg c = cfg;
        c.tab = batch ? d_tab : nullptr;
        c.scene_off = d_scene_off;
        return c;
    }

    // One step's kernels (timed: record events between the stages).
    int enqueue_step(bool timed) {
        const LaunchCfg c = launch_cfg();
        const bool keep = (flags & SPH_FLAG_KEEP_DVEL) != 0;
        if (timed) SPH_TRY(hipEventRecord(ev[0], stream));
        if (device_clock) SPH_TRY(launch_stats(c, buffers(), true, stream));
        // (k_integrate_moves clears the step's counters)
        // the mover counts alternate with the cell tables
        SPH_TRY(launch_integrate_moves(c, buffers(), resort, cur_cell, key_tmp, stream));
        if (timed) SPH_TRY(hipEventRecord(ev[1], stream));
        SPH_TRY(resort_order(c, buffers(), resort, cur_cell, stream));
        if (timed) SPH_TRY(hipEventRecord(ev[2], stream));
        SPH_TRY(resort_apply(c, buffers(), resort, cur_cell, sort, stream));
        cur_pos ^= 1;
        cur_vel ^= 1;
        cur_pid ^= 1;
        cur_cell ^= 1;
        cur_key ^= 1;
        last_skey = skey_buf[cur_key];
        if (timed) SPH_TRY(hipEventRecord(ev[3], stream));
        const bool fast = cfg.precision == SPH_PREC_FAST;
        last_pair_counts = fast || pair_entries_in_use(c, buffers());
        if (fast)
            SPH_TRY(launch_density_fast(c, buffers(), fast_rec, fast_nbr, fast_nnb, stream));
        else
            SPH_TRY(launch_density(c, buffers(), stream));
        if (timed) SPH_TRY(hipEventRecord(ev[4], stream));
        StepBuffers b = buffers();
        // (the FAST force walks k_density_fast's pair entries)
        SPH_TRY(launch_force(c, b, b.vel_alt, keep || cfg.colour, stream));
        if (cfg.colour) SPH_TRY(launch_colour(c, b, colour, dye, stream));
        cur_vel ^= 1;
        last_kind = 1;
        if (timed) SPH_TRY(hipEventRecord(ev[5], stream));
        return SPH_OK;
    }

    bool ready() const {
        for (size_t s = 0; s < have_pos.size(); ++s)
            if (!have_pos[s] || !have_vel[s]) return false;
        return true;
    }

    // The buffers the configured step needs exist: a lazy allocation that
    // failed (set_params to FAST or to a colour mode under a full device)
    // leaves the context unable to step until a later call allocates them.
    bool buffers_ok() const {
        if (cfg.max_cap <= kFastMaxCap && (!fast_nbr || !fast_nnb || !fast_rec || !fast_xyz)) return false;
        if (cfg.colour && (!colour || !dye || !dvel)) return false;
        return true;
    }

    void after_steps(uint32_t k) {
        steps += k;
        if (!device_clock) host_clock += (double)scenes[0].time_step * k;
        stepped = true;
    }

    int step(uint32_t nsteps) {
        if (!ready()) return SPH_ERR_NOT_INITIALISED;
        if (!buffers_ok()) return SPH_ERR_OUT_OF_MEMORY;
        DeviceGuard guard(device);
        for (uint32_t s = 0; s < nsteps; ++s) {
            const bool tm = timing && s + 1 == nsteps;
            if (!batch && changes_per_step(scenes[0])) {
                // the rotor angle (wave phase) advances before every step; the
                // kernels take the single scene's parameters by value
                advance_frame(scenes[0]);
                cfg.one = scenes[0];
            }
            SPH_TRY_ST(enqueue_step(tm));
            after_steps(1);
            if (tm) {
                SPH_TRY(hipEventSynchronize(ev[5]));
                for (int k = 0; k < 5; ++k) SPH_TRY(hipEventElapsedTime(&last_ms[k], ev[k], ev[k + 1]));
                SPH_TRY(hipEventElapsedTime(&last_ms[5], ev[0], ev[5]));
            }
        }
        return SPH_OK;
    }

    // Record `nsteps` (even, so the ping-pong buffers end where they began)
    // steps into a graph.
    int capture(uint32_t nsteps) {
        if (nsteps == 0 || (nsteps & 1u)) return SPH_ERR_INVALID_ARGUMENT;
        if (!ready()) return SPH_ERR_NOT_INITIALISED;
        if (!buffers_ok()) return SPH_ERR_OUT_OF_MEMORY;
        // a recording replays fixed parameters: no turning rotor or moving wave
        if (changes_per_step(scenes[0])) return SPH_ERR_UNSUPPORTED;
        DeviceGuard guard(device);
        drop_graph();
        SPH_TRY(hipStreamSynchronize(stream));
        const int p0 = cur_pos, v0 = cur_vel, q0 = cur_pid, k0 = last_kind, c0 = cur_cell, key0 = cur_key;
        const uint32_t* s0 = last_skey;
        SPH_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int st = SPH_OK;
        for (uint32_t s = 0; s < nsteps && st == SPH_OK; ++s) st = enqueue_step(false);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(stream, &g);
        cur_pos = p0;
        cur_vel = v0;
        cur_pid = q0;
        cur_cell = c0;
        cur_key = key0;
        last_skey = s0;
        last_kind = k0;  // recording ran nothing
        if (st != SPH_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        SPH_TRY(ee);
        graph = g;
        SPH_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
        graph_steps = nsteps;
        return SPH_OK;
    }

    int replay(uint32_t times) {
        if (!graph_exec) return SPH_ERR_NOT_INITIALISED;
        DeviceGuard guard(device);
        for (uint32_t t = 0; t < times; ++t) {
            SPH_TRY(hipGraphLaunch(graph_exec, stream));
            after_steps(graph_steps);
            last_kind = 1;
        }
        return SPH_OK;
    }

    // The error word of the step's order, read at every synchronising call.
    // A failed order is reported once as SPH_ERR_ORDER; the state then
    // counts as not uploaded (steps refuse to run, reads report the error)
    // until both arrays of every scene are uploaded in full again.
    int check_order() {
        DeviceGuard guard(device);
        uint32_t err = 0;
        SPH_TRY(hipMemcpyAsync(&err, resort.err, sizeof(err), hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        if (err & kErrInvalid) {
            SPH_TRY(launch_fill_u32(resort.err, 0u, 1, stream));
            SPH_TRY(resort_invalidate(resort, stream));
            // the slot -> index map may no longer be a permutation: back to
            // the identity, so that the uploads that must follow place every
            // particle
            SPH_TRY(launch_iota_u32(pid[cur_pid], cfg.n, stream));
            if (batch) SPH_TRY(launch_scene_ids(cfg.n, cfg.nscenes, d_scene_off, pid[cur_pid], sid[cur_pid], stream));
            SPH_TRY(hipStreamSynchronize(stream));
            order_failed = true;
            std::fill(have_pos.begin(), have_pos.end(), 0);
            std::fill(have_vel.begin(), have_vel.end(), 0);
        }
        return order_failed ? SPH_ERR_ORDER : SPH_OK;
    }

    int synchronize() {
        DeviceGuard guard(device);
        SPH_TRY(hipStreamSynchronize(stream));
        return check_order();
    }

    // Testing hook: perturb one word of the order state (sph_debug_poke).
    int debug_poke(int what, uint32_t index, uint32_t value) {
        DeviceGuard guard(device);
        uint32_t* word = nullptr;
        bool add = true;
        switch (what) {
            case SPH_DEBUG_MOVER_COUNT:
                if (index >= cfg.num_cells) return SPH_ERR_INVALID_ARGUMENT;
                word = resort.in_cnt[cur_cell] + index;  // the next step counts its movers here
                break;
            case SPH_DEBUG_CELL_END:
                if (index >= cfg.num_cells) return SPH_ERR_INVALID_ARGUMENT;
                word = cell_end[cur_cell] + index;
                break;
            case SPH_DEBUG_ORDER_ERROR:
                word = resort.err;
                value = kErrInvalid;
                add = false;
                break;
            default:
                return SPH_ERR_INVALID_ARGUMENT;
        }
        uint32_t v = 0;
        SPH_TRY(hipMemcpyAsync(&v, word, sizeof(v), hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        v = add ? v + value : (v | value);
        SPH_TRY(hipMemcpyAsync(word, &v, sizeof(v), hipMemcpyHostToDevice, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        return SPH_OK;
    }

    int set_array(uint32_t scene, int which, const float* data, uint32_t first, uint32_t count) {
        if (!data || scene >= cfg.nscenes) return SPH_ERR_INVALID_ARGUMENT;
        if (which != SPH_ARRAY_POSITION && which != SPH_ARRAY_VELOCITY) return SPH_ERR_INVALID_ARGUMENT;
        const uint32_t ns = scenes[scene].num_particles;
        if (first > ns || count > ns - first) return SPH_ERR_INVALID_ARGUMENT;
        if (count == 0) return SPH_OK;
        DeviceGuard guard(device);
        SPH_TRY(hipMemcpyAsync(staging, data, (size_t)count * sizeof(float4), hipMemcpyHostToDevice,
                               stream));
        float4* dst = which == SPH_ARRAY_POSITION ? pos[cur_pos] : vel[cur_vel];
        SPH_TRY(launch_update_from_original(cfg.n, pid[cur_pid], staging, scene_off[scene] + first,
                                            count, dst, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        if (first == 0 && count == ns) (which == SPH_ARRAY_POSITION ? have_pos : have_vel)[scene] = 1;
        if (order_failed && ready()) order_failed = false;
        last_kind = 0;
        return SPH_OK;
    }

    int get_array(uint32_t scene, int which, float* out) {
        if (!out || scene >= cfg.nscenes) return SPH_ERR_INVALID_ARGUMENT;
        SPH_TRY_ST(check_order());
        DeviceGuard guard(device);
        const uint32_t n = cfg.n;
        const uint32_t* p = pid[cur_pid];
        int rec = 4;
        bool pair_counts = false;  // FAST lists: the count word holds entries | pairs << 16
        switch (which) {
            case SPH_ARRAY_POSITION:
                SPH_TRY(launch_scatter_to_original(n, p, pos[cur_pos], staging, 4, 4, stream));
                break;
            case SPH_ARRAY_VELOCITY:
                SPH_TRY(launch_scatter_to_original(n, p, vel[cur_vel], staging, 4, 4, stream));
                break;
            case SPH_ARRAY_DENSITY:
                SPH_TRY(launch_scatter_to_original(n, p, rhop, staging, 1, 2, stream));
                rec = 1;
                break;
            case SPH_ARRAY_PRESSURE:
                SPH_TRY(launch_scatter_to_original(n, p, reinterpret_cast<float*>(rhop) + 1, staging, 1,
                                                   2, stream));
                rec = 1;
                break;
            case SPH_ARRAY_DVEL:
                if (!dvel) return SPH_ERR_UNSUPPORTED;
                SPH_TRY(launch_scatter_to_original(n, p, dvel, staging, 4, 4, stream));
                break;
            case SPH_ARRAY_COLOR:
                if (!colour) return SPH_ERR_UNSUPPORTED;
                SPH_TRY(launch_scatter_to_original(n, p, colour, staging, 4, 4, stream));
                break;
            case SPH_ARRAY_DYE:
                if (!dye) return SPH_ERR_UNSUPPORTED;
                SPH_TRY(hipMemcpyAsync(staging, dye, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, stream));
                rec = 1;
                break;
            case SPH_ARRAY_NEIGHBOURS:
                if (last_kind == 1 && last_pair_counts && fast_nnb) {
                    // the last plain step's force pairs per particle
                    SPH_TRY(launch_scatter_to_original(n, p, fast_nnb, staging, 1, 1, stream));
                    pair_counts = true;
                } else {
                    if (last_kind != 2) return SPH_ERR_UNSUPPORTED;
                    SPH_TRY_ST(learn_neighbour_counts(*this));
                }
                rec = 1;
                break;
            default:
                return SPH_ERR_INVALID_ARGUMENT;
        }
        const float* src = reinterpret_cast<const float*>(staging) + (size_t)scene_off[scene] * rec;
        SPH_TRY(hipMemcpyAsync(out, src, (size_t)scenes[scene].num_particles * rec * sizeof(float),
                               hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        if (pair_counts) {
            uint32_t* w = reinterpret_cast<uint32_t*>(out);
            for (uint32_t k = 0; k < scenes[scene].num_particles; ++k) w[k] >>= 16;
        }
        return SPH_OK;
    }

    int get_stats(uint32_t scene, sph_stats* out) {
        if (!out || scene >= cfg.nscenes) return SPH_ERR_INVALID_ARGUMENT;
        SPH_TRY_ST(check_order());
        DeviceGuard guard(device);
        std::memset(out, 0, sizeof(*out));
        const uint32_t ns = cfg.nscenes;
        std::vector<double> keep(2 * (size_t)ns), st((size_t)kNumStats * ns);
        std::vector<float> dts(ns);
        // keep each scene's time step and clock: recompute only the reductions
        for (uint32_t s = 0; s < ns; ++s)
            SPH_TRY(hipMemcpyAsync(&keep[2 * s], stats + (size_t)s * kNumStats + kSDt, 2 * sizeof(double),
                                   hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipMemcpyAsync(dts.data(), dt_dev, ns * sizeof(float), hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        const LaunchCfg c = launch_cfg();
        SPH_TRY(launch_stats(c, buffers(), false, stream));
        SPH_TRY(launch_fill_u32(cell_max, 0u, 1, stream));
        if (stepped) SPH_TRY(launch_cell_stats(c, buffers(), cell_max, stream));
        uint32_t cnt[kCntNum], cmax = 0, repairs = 0;
        SPH_TRY(hipMemcpyAsync(st.data(), stats, st.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipMemcpyAsync(cnt, counters, sizeof(cnt), hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipMemcpyAsync(&cmax, cell_max, sizeof(cmax), hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipMemcpyAsync(&repairs, resort.repairs, sizeof(repairs), hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        SPH_TRY(launch_fill_u32(resort.repairs, 0u, 1, stream));
        // the next step's time step (as the next step will compute it) for reporting
        const double dt_next = st[(size_t)scene * kNumStats + kSDt];
        for (uint32_t s = 0; s < ns; ++s)
            SPH_TRY(hipMemcpyAsync(stats + (size_t)s * kNumStats + kSDt, &keep[2 * s], 2 * sizeof(double),
                                   hipMemcpyHostToDevice, stream));
        SPH_TRY(hipMemcpyAsync(dt_dev, dts.data(), ns * sizeof(float), hipMemcpyHostToDevice, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        const double* q = &st[(size_t)scene * kNumStats];
        out->mass = q[kPMass];
        out->kinetic_energy = q[kPKin];
        out->potential_energy = q[kPPot];
        out->max_speed = q[kPVmax];
        out->dt_cfl = dt_next;
        out->sim_time = device_clock ? keep[2 * scene + 1] : host_clock;
        out->steps = steps;
        out->cap_cells = cnt[kCntCapCells];
        out->out_of_grid = cnt[kCntOutOfGrid];
        out->max_cell_count = cmax;
        out->sort_errors = repairs;
        out->full_sort = cnt[kCntFullSort] != 0;
        return SPH_OK;
    }

    // The grid of one scene: its slots are one contiguous run of the sorted
    // order (its cell keys are a range of their own), reported in scene-local
    // keys, indices and slots.
    int get_grid(uint32_t scene, uint32_t* sorted_hash, uint32_t* perm, uint32_t* cs, uint32_t* ce) {
        if (scene >= cfg.nscenes) return SPH_ERR_INVALID_ARGUMENT;
        if (!stepped) return SPH_ERR_NOT_INITIALISED;
        SPH_TRY_ST(check_order());
        DeviceGuard guard(device);
        const uint32_t a = scene_off[scene];
        const size_t n = scenes[scene].num_particles, ncell = scenes[scene].num_cells;
        const uint32_t base = scenes[scene].cell_base;
        if (sorted_hash)
            SPH_TRY(hipMemcpyAsync(sorted_hash, last_skey + a, n * 4, hipMemcpyDeviceToHost, stream));
        if (perm) SPH_TRY(hipMemcpyAsync(perm, pid[cur_pid] + a, n * 4, hipMemcpyDeviceToHost, stream));
        std::vector<uint32_t> tmp;
        uint32_t* csp = cs;
        if (ce && !cs) {
            tmp.resize(ncell);
            csp = tmp.data();
        }
        if (csp)
            SPH_TRY(hipMemcpyAsync(csp, cell_start[cur_cell] + base, ncell * 4, hipMemcpyDeviceToHost, stream));
        if (ce) SPH_TRY(hipMemcpyAsync(ce, cell_end[cur_cell] + base, ncell * 4, hipMemcpyDeviceToHost, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        if (sorted_hash && base)
            for (size_t k = 0; k < n; ++k) sorted_hash[k] -= base;
        if (perm && a)
            for (size_t k = 0; k < n; ++k) perm[k] -= a;
        // cellEnd is only maintained for occupied cells; report empty ones as
        // the reference marks an empty cellStart.
        for (size_t k = 0; k < ncell && csp; ++k) {
            if (csp[k] == kEmptyCell) {
                if (ce) ce[k] = kEmptyCell;
            } else {
                if (ce) ce[k] -= a;
                csp[k] -= a;
            }
        }
        return SPH_OK;
    }

    int set_tune(const uint32_t t[4]) {
        for (int k = 0; k < 4; ++k) cfg.tune[k] = t[k];
        for (auto& s : scenes)
            for (int k = 0; k < 4; ++k) s.tune[k] = t[k];
        cfg.one = scenes[0];
        drop_graph();
        return SPH_OK;
    }

    int set_params(uint32_t scene, const sph_params* p) {
        int st = validate(p);
        if (st != SPH_OK) return st;
        sph_params& cur = scenes[scene];
        if (p->num_particles != cur.num_particles || p->num_cells != cur.num_cells ||
            p->grid_size[0] != cur.grid_size[0] || p->grid_size[1] != cur.grid_size[1] ||
            p->grid_size[2] != cur.grid_size[2] || (cfg.nscenes > 1 && p->precision != cfg.precision))
            return SPH_ERR_INVALID_PARAMS;
        if (!batch && ((1ull << p->key_bits) < cfg.num_cells || p->key_bits > 31))
            return SPH_ERR_INVALID_PARAMS;
        if (batch && changes_per_step(*p)) return SPH_ERR_UNSUPPORTED;
        DeviceGuard guard(device);
        if ((p->cfl > 0.0f || p->t_end > 0.0f) && !device_clock) {
            // from here on the clock and time step advance on the device
            SPH_TRY(hipMemcpyAsync(stats + kSClock, &host_clock, sizeof(double), hipMemcpyHostToDevice,
                                   stream));
            device_clock = true;
        }
        const uint32_t base = cur.cell_base, kb = cur.key_bits;
        cur = *p;
        derive_shape(cur);
        cur.cell_base = base;
        if (batch) cur.key_bits = kb;
        if (!batch) cfg.key_bits = cur.key_bits;
        cfg.precision = cur.precision;
        cfg.max_cap = 0;
        for (auto& s : scenes) cfg.max_cap = std::max(cfg.max_cap, s.max_par_in_cell);
        for (int k = 0; k < 4; ++k) cfg.tune[k] = scenes[0].tune[k];
        cfg.one = scenes[0];
        cfg.physics = 0;
        for (const auto& sc : scenes) cfg.physics |= sc.physics;
        SPH_TRY(ensure_colour());
        SPH_TRY(ensure_fast());
        if (batch)
            SPH_TRY(hipMemcpyAsync(d_tab + scene, &cur, sizeof(sph_params), hipMemcpyHostToDevice, stream));
        SPH_TRY(hipMemcpyAsync(dt_dev + scene, &cur.time_step, sizeof(float), hipMemcpyHostToDevice, stream));
        SPH_TRY(hipStreamSynchronize(stream));
        drop_graph();
        return SPH_OK;
    }

    // ---- autotuner --------------------------------------------------------
    std::string tune_key() const {
        hipDeviceProp_t prop{};
        (void)hipGetDeviceProperties(&prop, device);
        char buf[512];
        std::snprintf(buf, sizeof(buf), "v2|%s|cu%d|n%u|s%u|c%u|cap%u|p%u|nb%u", prop.gcnArchName,
                      prop.multiProcessorCount, cfg.n, cfg.nscenes, cfg.num_cells, cfg.max_cap,
                      cfg.precision, cfg.max_scene_n);
        return buf;
    }

    static bool cache_lookup(const char* path, const std::string& key, uint32_t t[4], float* ms) {
        if (!path) return false;
        FILE* f = std::fopen(path, "r");
        if (!f) return false;
        char line[1024];
        bool found = false;
        while (std::fgets(line, sizeof(line), f)) {
            char k[512];
            unsigned a = 0, b = 0, d = 0;
            float m = 0;
            if (std::sscanf(line, "{\"key\": \"%511[^\"]\", \"tune0\": %u, \"tune1\": %u, \"tune3\": %u, \"ms\": %f}",
                            k, &a, &b, &d, &m) == 5 &&
                key == k) {
                t[0] = a;
                t[1] = b;
                t[2] = 0;
                t[3] = d;
                if (ms) *ms = m;
                found = true;  // the last matching line wins
            }
        }
        std::fclose(f);
        return found;
    }

    static void cache_store(const char* path, const std::string& key, const uint32_t t[4], float ms) {
        if (!path) return;
        FILE* f = std::fopen(path, "a");
        if (!f) return;
        std::fprintf(f, "{\"key\": \"%s\", \"tune0\": %u, \"tune1\": %u, \"tune3\": %u, \"ms\": %.6f}\n",
                     key.c_str(), t[0], t[1], t[3], ms);
        std::fclose(f);
    }

    // Time each candidate launch configuration over a few steps of the live
    // state, restore the state, keep the fastest.  Every candidate computes
    // identical results, so the choice only moves time.
    int autotune(const char* cache_path, uint32_t out[4], float* best_ms_out) {
        if (!ready()) return SPH_ERR_NOT_INITIALISED;
        if (!buffers_ok()) return SPH_ERR_OUT_OF_MEMORY;
        const std::string key = tune_key();
        uint32_t t[4] = {0, 0, 0, 0};
        float ms = 0.0f;
        if (cache_lookup(cache_path, key, t, &ms)) {
            set_tune(t);
            for (int k = 0; k < 4; ++k) out[k] = t[k];
            if (best_ms_out) *best_ms_out = ms;
            return SPH_OK;
        }
        DeviceGuard guard(device);
        hipDeviceProp_t prop{};
        SPH_TRY(hipGetDeviceProperties(&prop, device));
        const uint32_t cus = (uint32_t)std::max(1, prop.multiProcessorCount);
        // snapshot of everything a step reads or writes that outlives it
        const size_t N = cfg.n;
        float4 *sp = nullptr, *sv = nullptr;
        uint32_t *spid = nullptr, *ssid = nullptr;
        double* sst = nullptr;
        float* sdt = nullptr;
        hipError_t e = dalloc(&sp, N);
        if (e == hipSuccess) e = dalloc(&sv, N);
        if (e == hipSuccess) e = dalloc(&spid, N);
        if (e == hipSuccess && batch) e = dalloc(&ssid, N);
        if (e == hipSuccess) e = dalloc(&sst, (size_t)kNumStats * cfg.nscenes);
        if (e == hipSuccess) e = dalloc(&sdt, (size_t)cfg.nscenes);
        auto cleanup = [&]() {
            (void)hipFree(sp);
            (void)hipFree(sv);
            (void)hipFree(spid);
            (void)hipFree(ssid);
            (void)hipFree(sst);
            (void)hipFree(sdt);
        };
        if (e != hipSuccess) {
            cleanup();
            return to_status(e);
        }
        const int p0 = cur_pos, v0 = cur_vel, q0 = cur_pid;
        const uint64_t steps0 = steps;
        const double clock0 = host_clock;
        const bool stepped0 = stepped;
        auto snap = [&](bool save) -> hipError_t {
            hipError_t r = hipSuccess;
            const hipMemcpyKind k = hipMemcpyDeviceToDevice;
            float4* P = pos[cur_pos];
            float4* V = vel[cur_vel];
            uint32_t* Q = pid[cur_pid];
            uint32_t* S = sid[cur_pid];
            if (r == hipSuccess) r = hipMemcpyAsync(save ? sp : P, save ? P : sp, N * 16, k, stream);
            if (r == hipSuccess) r = hipMemcpyAsync(save ? sv : V, save ? V : sv, N * 16, k, stream);
            if (r == hipSuccess) r = hipMemcpyAsync(save ? spid : Q, save ? Q : spid, N * 4, k, stream);
            if (r == hipSuccess && batch) r = hipMemcpyAsync(save ? ssid : S, save ? S : ssid, N * 4, k, stream);
            if (r == hipSuccess)
                r = hipMemcpyAsync(save ? sst : stats, save ? stats : sst,
                                   (size_t)kNumStats * cfg.nscenes * 8, k, stream);
            if (r == hipSuccess)
                r = hipMemcpyAsync(save ? sdt : dt_dev, save ? dt_dev : sdt, (size_t)cfg.nscenes * 4, k,
                                   stream);
            return r;
        };
        e = snap(true);
        // 3 = the pair-entry path (the force walks the density kernel's
        // pair entries), 2 = the pair-mask path (force lists filled from the
        // density kernel's masks), else the list size of the
        // candidate-testing force
        const uint32_t lists[] = {3, 2, 16, 24, 32};
        const bool fast = cfg.precision == SPH_PREC_FAST;  // one launch shape: time it once
        // gather grid and variant: per-lane loads in XCD order, cell tables in
        // LDS, or a persistent grid
        const uint32_t shapes[][2] = {{0, 0}, {0, 2}, {2 * cus, 0}, {4 * cus, 0}, {8 * cus, 0}};
        float best = 1e30f;
        uint32_t best_t[4] = {cfg.tune[0], cfg.tune[1], 0, cfg.tune[3]};
        const uint32_t saved_tune[4] = {cfg.tune[0], cfg.tune[1], cfg.tune[2], cfg.tune[3]};
        int st = SPH_OK;
        for (uint32_t li : lists) {
            if (fast && li != lists[0]) break;
            if (!fast && li == 3 && (cfg.max_cap > kFastMaxCap || !fast_nbr)) continue;
            if (!fast && li == 2 && (cfg.max_cap > kMaskCap || !masks)) continue;
            if (!fast && li > 3 && cfg.max_cap > li) continue;
            for (const auto& sh : shapes) {
                if (e != hipSuccess || st != SPH_OK) break;
                if ((li <= 3 || fast) && (sh[0] != 0 || sh[1] != 0)) continue;  // one launch shape
                const uint32_t g = sh[0];
                const uint32_t cand[4] = {li, g, 0, sh[1]};
                set_tune(cand);
                cur_pos = p0;
                cur_vel = v0;
                cur_pid = q0;
                e = snap(false);
                // the slots are back in the saved order, the cell ranges are not
                if (e == hipSuccess) e = resort_invalidate(resort, stream);
                if (e != hipSuccess) break;
                st = enqueue_step(false);  // warm-up
                if (st != SPH_OK) break;
                e = hipEventRecord(ev[0], stream);
                const int reps = 4;
                for (int r = 0; r < reps && st == SPH_OK; ++r) st = enqueue_step(false);
                if (e == hipSuccess) e = hipEventRecord(ev[5], stream);
                if (e == hipSuccess) e = hipEventSynchronize(ev[5]);
                float m = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&m, ev[0], ev[5]);
                m /= reps;
                if (e == hipSuccess && st == SPH_OK && m < best) {
                    best = m;
                    best_t[0] = li;
                    best_t[1] = g;
                    best_t[3] = sh[1];
                }
            }
        }
        cur_pos = p0;
        cur_vel = v0;
        cur_pid = q0;
        if (e == hipSuccess) e = snap(false);
        if (e == hipSuccess) e = resort_invalidate(resort, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        cleanup();
        steps = steps0;
        host_clock = clock0;
        stepped = stepped0;
        last_kind = 0;  // densities and cell ranges are those of a timing step
        if (st != SPH_OK) {
            set_tune(saved_tune);
            return st;
        }
        if (e != hipSuccess) {
            set_tune(saved_tune);
            return to_status(e);
        }
        set_tune(best_t);
        cache_store(cache_path, key, best_t, best);
        for (int k = 0; k < 4; ++k) out[k] = best_t[k];
        if (best_ms_out) *best_ms_out = best;
        return SPH_OK;
    }
};

}  // namespace sph

struct sph_ctx {
    sph::Engine eng;
};

struct sph_batch {
    sph::Engine eng;
};
