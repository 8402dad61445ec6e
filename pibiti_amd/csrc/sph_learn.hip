// sph_learn.hip - learned pair-force surrogate behind the C ABI: models,
// training windows kept on a device context, and learned stepping at a
// multiple of the time step with a plain-solver fallback.
//
// Training data come from the plain solver itself: a window starts after a
// plain step (its state as the force kernel saw it is kept on the device)
// and the pair-force velocity increments of that step and the following
// ones are summed per original particle index.  Gradients and Adam updates
// run on the device against that window; nothing leaves HBM except the loss.
#include <cmath>
#include <initializer_list>

#include "sph_engine.hpp"
#include "sph_mlp.hpp"

struct sph_model {
    int device = 0;
    sph_mlp_norm norm{};
    float* params = nullptr;  // FP32 master weights
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    sph::MlpPacked pk{};
    uint64_t adam_t = 0;
};

namespace sph {

struct LearnState {
    uint32_t n = 0;
    // kept training window (sorted order of its step)
    bool have = false;
    uint32_t window = 0;
    float4* pos = nullptr;
    float4* vel = nullptr;  // before the force update
    float2* rhop = nullptr;
    uint32_t* pid = nullptr;
    uint32_t* sid = nullptr;
    uint32_t* nbr = nullptr;
    uint32_t* nnb = nullptr;
    float4* acc = nullptr;  // summed force increments, original order
    // training scratch
    bool train_ready = false;
    MlpTrainBuffers tb{};
    // learned stepping
    bool step_ready = false;
    uint32_t* lnbr = nullptr;
    uint32_t* lnnb = nullptr;
    float4* spos = nullptr;
    float4* svel = nullptr;
    uint32_t* spid = nullptr;
    double* dev_part = nullptr;
    double* dev = nullptr;       // D of each surrogate step of a chunk
    uint32_t* ovf = nullptr;     // per step: list overflow, listed pairs, skin moves
    float* dt_coarse = nullptr;
    float4* bpos = nullptr;      // positions at the last sort (skin > 0)
    double* host = nullptr;      // pinned copies of dev / ovf, one region per chunk slot
    uint32_t* host_cnt = nullptr;
    hipEvent_t done[2] = {};     // end of the chunk in each slot
};

// Surrogate steps between host checks (sph_params::tune[2], 0 = default).
constexpr uint32_t kDefaultCheck = 4;
constexpr uint32_t kMaxCheck = 64;
// Counters kept per surrogate step: list overflow, listed pairs, skin moves
// (kCntNbrOverflow .. kCntSkinMoved, consecutive).
constexpr uint32_t kStepCounters = 3;
static_assert(kCntPairs == kCntNbrOverflow + 1 && kCntSkinMoved == kCntNbrOverflow + 2,
              "step counters are copied as one range");

void learn_release(Engine& e) {
    LearnState* L = e.learn;
    if (!L) return;
    void* ptrs[] = {L->pos,      L->vel,         L->rhop,         L->pid,          L->sid,
                    L->nbr,      L->nnb,         L->acc,          L->tb.h,         L->tb.out,
                    L->tb.dh,    L->tb.part_pair, L->tb.part_node, L->tb.loss_part, L->tb.grad,
                    L->tb.loss,  L->tb.chunks,   L->lnbr,         L->lnnb,         L->spos,         L->svel,
                    L->spid,     L->dev_part,    L->dev,          L->dt_coarse,   L->ovf,
                    L->bpos};
    for (void* q : ptrs) (void)hipFree(q);
    if (L->host) (void)hipHostFree(L->host);
    if (L->host_cnt) (void)hipHostFree(L->host_cnt);
    for (hipEvent_t ev : L->done)
        if (ev) (void)hipEventDestroy(ev);
    delete L;
    e.learn = nullptr;
}

int learn_neighbour_counts(Engine& e) {
    const LearnState* L = e.learn;
    if (!L || !L->lnnb) return SPH_ERR_UNSUPPORTED;
    SPH_TRY(launch_scatter_to_original(e.cfg.n, e.pid[e.cur_pid], L->lnnb, e.staging, 1, 1, e.stream));
    return SPH_OK;
}

namespace {

LearnState* learn_state(Engine& e) {
    if (!e.learn) {
        e.learn = new (std::nothrow) LearnState();
        if (e.learn) e.learn->n = e.cfg.n;
    }
    return e.learn;
}

// Frees a group of buffers whose allocation failed part way, so that a
// later call starts the group again from nothing.
void drop_all(std::initializer_list<void**> ptrs) {
    for (void** q : ptrs) {
        (void)hipFree(*q);
        *q = nullptr;
    }
}

hipError_t ensure_window(LearnState& L, bool batch) {
    if (L.pos) return hipSuccess;
    const size_t N = L.n;
    hipError_t e = dalloc(&L.pos, N);
    if (e == hipSuccess) e = dalloc(&L.vel, N);
    if (e == hipSuccess) e = dalloc(&L.rhop, N);
    if (e == hipSuccess) e = dalloc(&L.pid, N);
    if (e == hipSuccess && batch) e = dalloc(&L.sid, N);
    if (e == hipSuccess) e = dalloc(&L.nbr, N * kNbrCap);
    if (e == hipSuccess) e = dalloc(&L.nnb, N);
    if (e == hipSuccess) e = dalloc(&L.acc, N);
    if (e != hipSuccess)
        drop_all({(void**)&L.pos, (void**)&L.vel, (void**)&L.rhop, (void**)&L.pid, (void**)&L.sid, (void**)&L.nbr,
                  (void**)&L.nnb, (void**)&L.acc});
    return e;
}

hipError_t ensure_train(LearnState& L) {
    if (L.train_ready) return hipSuccess;
    const size_t N = L.n;
    MlpTrainBuffers& t = L.tb;
    hipError_t e = dalloc(&t.h, N * mlp::kHid);
    if (e == hipSuccess) e = dalloc(&t.out, N);
    if (e == hipSuccess) e = dalloc(&t.dh, N * mlp::kHid);
    if (e == hipSuccess) e = dalloc(&t.part_pair, (size_t)mlp_pair_grad_waves(L.n) * mlp::kPairParams);
    if (e == hipSuccess) e = dalloc(&t.part_node, (size_t)mlp_node_blocks(L.n) * mlp::kNodeParams);
    if (e == hipSuccess) e = dalloc(&t.loss_part, (size_t)mlp_node_blocks(L.n));
    if (e == hipSuccess) e = dalloc(&t.chunks, (size_t)mlp::kRedChunks * (mlp::kNumParams + 1));
    if (e == hipSuccess) e = dalloc(&t.grad, (size_t)mlp::kNumParams);
    if (e == hipSuccess) e = dalloc(&t.loss, (size_t)1);
    if (e == hipSuccess)
        L.train_ready = true;
    else
        drop_all({(void**)&t.h, (void**)&t.out, (void**)&t.dh, (void**)&t.part_pair, (void**)&t.part_node,
                  (void**)&t.loss_part, (void**)&t.chunks, (void**)&t.grad, (void**)&t.loss});
    return e;
}

hipError_t ensure_step(LearnState& L, uint32_t nscenes) {
    if (L.step_ready) return hipSuccess;
    const size_t N = L.n;
    hipError_t e = dalloc(&L.lnbr, N * kNbrCap);
    if (e == hipSuccess) e = dalloc(&L.lnnb, N);
    if (e == hipSuccess) e = dalloc(&L.spos, 2 * N);
    if (e == hipSuccess) e = dalloc(&L.svel, 2 * N);
    if (e == hipSuccess) e = dalloc(&L.spid, 2 * N);
    if (e == hipSuccess) e = dalloc(&L.dev_part, (size_t)256);
    if (e == hipSuccess) e = dalloc(&L.dev, (size_t)kMaxCheck);
    if (e == hipSuccess) e = dalloc(&L.ovf, (size_t)kStepCounters * kMaxCheck);
    if (e == hipSuccess) e = dalloc(&L.dt_coarse, (size_t)nscenes);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L.host), 2 * kMaxCheck * sizeof(double));
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&L.host_cnt),
                          2 * kStepCounters * kMaxCheck * sizeof(uint32_t));
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&L.done[k], hipEventDisableTiming);
    if (e == hipSuccess) {
        L.step_ready = true;
        return e;
    }
    drop_all({(void**)&L.lnbr, (void**)&L.lnnb, (void**)&L.spos, (void**)&L.svel, (void**)&L.spid,
              (void**)&L.dev_part, (void**)&L.dev, (void**)&L.ovf, (void**)&L.dt_coarse});
    if (L.host) (void)hipHostFree(L.host);
    if (L.host_cnt) (void)hipHostFree(L.host_cnt);
    L.host = nullptr;
    L.host_cnt = nullptr;
    for (hipEvent_t& ev : L.done) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    return e;
}

hipError_t ensure_skin(LearnState& L) {
    if (L.bpos) return hipSuccess;
    return dalloc(&L.bpos, (size_t)L.n);
}

MlpState window_state(const LearnState& L) {
    MlpState st;
    st.n = L.n;
    st.pos = L.pos;
    st.vel = L.vel;
    st.rhop = L.rhop;
    st.sid = L.sid;
    st.nbr = L.nbr;
    st.nnb = L.nnb;
    return st;
}

int check_model(const Engine& e, const sph_model* m) {
    if (!m) return SPH_ERR_INVALID_ARGUMENT;
    if (m->device != e.device) return SPH_ERR_INVALID_ARGUMENT;
    return SPH_OK;
}

// Keep the state of the step just taken as a training window.
int learn_begin(Engine& e, const sph_model* m) {
    SPH_TRY_ST(check_model(e, m));
    if (!(e.flags & SPH_FLAG_KEEP_DVEL)) return SPH_ERR_UNSUPPORTED;
    // the targets are pair-force increments: no collider or accelerator terms
    for (const auto& sc : e.scenes)
        if (sc.physics != 0) return SPH_ERR_UNSUPPORTED;
    if (e.last_kind != 1) return SPH_ERR_NOT_INITIALISED;
    LearnState* L = learn_state(e);
    if (!L) return SPH_ERR_OUT_OF_MEMORY;
    DeviceGuard guard(e.device);
    SPH_TRY(ensure_window(*L, e.batch));
    const size_t N = e.cfg.n;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    SPH_TRY(hipMemcpyAsync(L->pos, e.pos[e.cur_pos], N * 16, k, e.stream));
    SPH_TRY(hipMemcpyAsync(L->vel, e.vel[e.cur_vel ^ 1], N * 16, k, e.stream));
    SPH_TRY(hipMemcpyAsync(L->pid, e.pid[e.cur_pid], N * 4, k, e.stream));
    if (e.batch) SPH_TRY(hipMemcpyAsync(L->sid, e.sid[e.cur_pid], N * 4, k, e.stream));
    // density and neighbour lists of the window state (the cell ranges are
    // still those of the step just taken): the same kernel arithmetic as the
    // step's density, so the densities are the ones its force kernel saw
    StepBuffers b = e.buffers();
    b.pos_cur = L->pos;
    b.sid_cur = L->sid;
    b.rhop = L->rhop;
    SPH_TRY(launch_fill_u32(e.counters + kCntNbrOverflow, 0u, 1, e.stream));
    SPH_TRY(launch_density_lists(e.launch_cfg(), b, L->nbr, L->nnb, e.stream));
    SPH_TRY(mlp_accumulate_original(e.cfg.n, e.pid[e.cur_pid], e.dvel, L->acc, true, e.stream));
    uint32_t overflow = 0;
    SPH_TRY(hipMemcpyAsync(&overflow, e.counters + kCntNbrOverflow, 4, hipMemcpyDeviceToHost, e.stream));
    SPH_TRY(hipStreamSynchronize(e.stream));
    if (overflow) {
        L->have = false;
        return SPH_ERR_CAPACITY;
    }
    L->have = true;
    L->window = 1;
    return SPH_OK;
}

int learn_add(Engine& e, const sph_model* m) {
    SPH_TRY_ST(check_model(e, m));
    LearnState* L = e.learn;
    if (!L || !L->have || e.last_kind != 1) return SPH_ERR_NOT_INITIALISED;
    DeviceGuard guard(e.device);
    SPH_TRY(mlp_accumulate_original(e.cfg.n, e.pid[e.cur_pid], e.dvel, L->acc, false, e.stream));
    L->window += 1;
    return SPH_OK;
}

int learn_gradient(Engine& e, sph_model* m, double* loss_out, float* grad_out) {
    SPH_TRY_ST(check_model(e, m));
    LearnState* L = e.learn;
    if (!L || !L->have) return SPH_ERR_NOT_INITIALISED;
    DeviceGuard guard(e.device);
    SPH_TRY(ensure_train(*L));
    SPH_TRY(mlp_gradient(e.launch_cfg(), m->norm, m->params, m->pk, window_state(*L), L->pid, L->acc,
                         L->window, L->tb, e.stream));
    double loss = 0.0;
    SPH_TRY(hipMemcpyAsync(&loss, L->tb.loss, sizeof(double), hipMemcpyDeviceToHost, e.stream));
    if (grad_out)
        SPH_TRY(hipMemcpyAsync(grad_out, L->tb.grad, mlp::kNumParams * sizeof(float), hipMemcpyDeviceToHost,
                               e.stream));
    SPH_TRY(hipStreamSynchronize(e.stream));
    if (loss_out) *loss_out = loss;
    return SPH_OK;
}

int learn_train(Engine& e, sph_model* m, uint32_t iters, float lr, double* loss_out) {
    SPH_TRY_ST(check_model(e, m));
    if (!(lr > 0.0f)) return SPH_ERR_INVALID_ARGUMENT;
    LearnState* L = e.learn;
    if (!L || !L->have) return SPH_ERR_NOT_INITIALISED;
    DeviceGuard guard(e.device);
    SPH_TRY(ensure_train(*L));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t it = 0; it < iters; ++it) {
        SPH_TRY(mlp_gradient(e.launch_cfg(), m->norm, m->params, m->pk, window_state(*L), L->pid, L->acc,
                             L->window, L->tb, e.stream));
        m->adam_t += 1;
        const double t = (double)m->adam_t;
        const float bc1 = (float)(1.0 / (1.0 - std::pow((double)b1, t)));
        const float bc2 = (float)(1.0 / (1.0 - std::pow((double)b2, t)));
        SPH_TRY(mlp_adam(m->params, m->adam_m, m->adam_v, L->tb.grad, lr, b1, b2, eps, bc1, bc2, m->pk,
                         e.stream));
    }
    if (loss_out) {
        double loss = 0.0;
        if (iters) {
            SPH_TRY(hipMemcpyAsync(&loss, L->tb.loss, sizeof(double), hipMemcpyDeviceToHost, e.stream));
            SPH_TRY(hipStreamSynchronize(e.stream));
        }
        *loss_out = loss;
    }
    return SPH_OK;
}

int learn_predict(Engine& e, sph_model* m, float* out4) {
    SPH_TRY_ST(check_model(e, m));
    if (!out4) return SPH_ERR_INVALID_ARGUMENT;
    LearnState* L = e.learn;
    if (!L || !L->have) return SPH_ERR_NOT_INITIALISED;
    DeviceGuard guard(e.device);
    SPH_TRY(ensure_train(*L));
    SPH_TRY(mlp_infer(e.launch_cfg(), m->norm, m->params, m->pk, window_state(*L), nullptr, L->tb.out,
                      nullptr, e.stream));
    SPH_TRY(launch_scatter_to_original(e.cfg.n, L->pid, L->tb.out, e.staging, 4, 4, e.stream));
    SPH_TRY(hipMemcpyAsync(out4, e.staging, (size_t)e.cfg.n * 16, hipMemcpyDeviceToHost, e.stream));
    SPH_TRY(hipStreamSynchronize(e.stream));
    return SPH_OK;
}

int learn_get_window(Engine& e, uint32_t scene, int which, float* out) {
    if (!out || scene >= e.cfg.nscenes) return SPH_ERR_INVALID_ARGUMENT;
    LearnState* L = e.learn;
    if (!L || !L->have) return SPH_ERR_NOT_INITIALISED;
    DeviceGuard guard(e.device);
    const uint32_t n = e.cfg.n;
    int rec = 4;
    const float* src = nullptr;
    switch (which) {
        case SPH_ARRAY_POSITION:
            SPH_TRY(launch_scatter_to_original(n, L->pid, L->pos, e.staging, 4, 4, e.stream));
            break;
        case SPH_ARRAY_VELOCITY:
            SPH_TRY(launch_scatter_to_original(n, L->pid, L->vel, e.staging, 4, 4, e.stream));
            break;
        case SPH_ARRAY_DENSITY:
            SPH_TRY(launch_scatter_to_original(n, L->pid, L->rhop, e.staging, 1, 2, e.stream));
            rec = 1;
            break;
        case SPH_ARRAY_PRESSURE:
            SPH_TRY(launch_scatter_to_original(n, L->pid, reinterpret_cast<float*>(L->rhop) + 1, e.staging, 1,
                                               2, e.stream));
            rec = 1;
            break;
        case SPH_ARRAY_TARGET:
            src = reinterpret_cast<const float*>(L->acc);
            break;
        default:
            return SPH_ERR_INVALID_ARGUMENT;
    }
    if (!src) src = reinterpret_cast<const float*>(e.staging);
    const size_t first = (size_t)e.scene_off[scene] * rec;
    const size_t count = (size_t)e.scenes[scene].num_particles * rec;
    SPH_TRY(hipMemcpyAsync(out, src + first, count * sizeof(float), hipMemcpyDeviceToHost, e.stream));
    SPH_TRY(hipStreamSynchronize(e.stream));
    if (which == SPH_ARRAY_TARGET) {
        const double scale = 1.0 / ((double)L->window * (double)e.scenes[scene].time_step);
        for (size_t k = 0; k < count; ++k) out[k] = (float)((double)out[k] * scale);
    }
    return SPH_OK;
}

// One surrogate step: plain integrate (coarse dt); with skin = 0 or on a
// rebuild step, sort and reorder (with skin > 0 keeping the sorted
// positions); density with neighbour lists (the 27-cell walk after a sort,
// else the walk of launch_density_reach); the density deviation into
// dev_out and the step's counters into cnt_out (both device); the surrogate
// velocity update.  Events: ev[1]/ev[2] bracket the density kernel,
// ev[3]/ev[4] the surrogate, ev[5]/ev[6] a rebuild's integrate, order and
// reorder.
int enqueue_coarse(Engine& e, LearnState& L, const sph_model& m, float skin, bool rebuild, double* dev_out,
                   uint32_t* cnt_out, bool timed) {
    const LaunchCfg c = e.launch_cfg();
    SPH_TRY(launch_fill_u32(e.counters, 0u, kCntNum, e.stream));
    StepBuffers b = e.buffers();
    b.dt_dev = L.dt_coarse;
    const bool sorted = skin == 0.0f || rebuild;
    if (sorted) {
        // the plain step's order: only the particles that changed cell are
        // placed again (the full sort when a cell crowds)
        if (timed && skin > 0.0f) SPH_TRY(hipEventRecord(e.ev[5], e.stream));
        SPH_TRY(launch_integrate_moves(c, b, e.resort, e.cur_cell, e.key_tmp, e.stream));
        SPH_TRY(resort_order(c, b, e.resort, e.cur_cell, e.stream));
        SPH_TRY(resort_apply(c, b, e.resort, e.cur_cell, e.sort, e.stream));
        e.cur_pos ^= 1;
        e.cur_vel ^= 1;
        e.cur_pid ^= 1;
        e.cur_cell ^= 1;
        e.cur_key ^= 1;
        e.last_skey = e.skey_buf[e.cur_key];
        if (skin > 0.0f) {
            SPH_TRY(hipMemcpyAsync(L.bpos, e.pos[e.cur_pos], (size_t)c.n * sizeof(float4),
                                   hipMemcpyDeviceToDevice, e.stream));
            if (timed) SPH_TRY(hipEventRecord(e.ev[6], e.stream));
        }
    } else {
        // a step between sorts keeps the slot order and the cell ranges
        // (and so the order stays valid for the next sort)
        SPH_TRY(launch_integrate(c, b, e.stream));
    }
    b = e.buffers();
    if (timed) SPH_TRY(hipEventRecord(e.ev[1], e.stream));
    if (sorted)
        SPH_TRY(launch_density_lists(c, b, L.lnbr, L.lnnb, e.stream));
    else
        SPH_TRY(launch_density_reach(c, b, skin, L.bpos, L.lnbr, L.lnnb, e.stream));
    if (timed) SPH_TRY(hipEventRecord(e.ev[2], e.stream));
    SPH_TRY(mlp_density_deviation(c.n, b.rhop, 1.0f / e.scenes[0].rest_density, L.dev_part, dev_out, e.stream,
                                  e.counters, kCntNbrOverflow, kStepCounters, cnt_out));
    MlpState st;
    st.n = c.n;
    st.pos = b.pos_cur;
    st.vel = b.vel_cur;
    st.rhop = b.rhop;
    st.sid = b.sid_cur;
    st.nbr = L.lnbr;
    st.nnb = L.lnnb;
    if (timed) SPH_TRY(hipEventRecord(e.ev[3], e.stream));
    SPH_TRY(mlp_infer(c, m.norm, m.params, m.pk, st, b.vel_alt, nullptr, L.dt_coarse, e.stream));
    if (timed) SPH_TRY(hipEventRecord(e.ev[4], e.stream));
    e.cur_vel ^= 1;
    return SPH_OK;
}

// D of the current state (after a plain step), synchronously.
int plain_deviation(Engine& e, LearnState& L, double* out) {
    SPH_TRY(mlp_density_deviation(e.cfg.n, e.rhop, 1.0f / e.scenes[0].rest_density, L.dev_part, L.dev,
                                  e.stream));
    SPH_TRY(hipMemcpyAsync(L.host, L.dev, sizeof(double), hipMemcpyDeviceToHost, e.stream));
    SPH_TRY(hipStreamSynchronize(e.stream));
    *out = L.host[0];
    return SPH_OK;
}

// Surrogate steps run in chunks of `check` steps between host checks.  A
// chunk starts from a snapshot of the state (and, with a skin, its first
// step sorts while the others keep that sort's order); its steps'
// deviations D and counters are copied to the host at its end.  Two chunks
// are in flight: the next one is queued before the host reads this one's
// results, so the device never waits for the host.  The first step whose
// drift |D - anchor| exceeds tol (or whose neighbour lists overflowed) is
// replaced: the queue is drained, the chunk is undone from its snapshot,
// the steps before the bad one are taken again (the surrogate step is
// deterministic, so they reproduce exactly) and `factor` plain steps stand
// in for it; the chunk queued behind it is discarded.  A step in which a
// particle moved farther than the skin since the sort ends the chunk the
// same way, but is then taken again as the first step of the next chunk,
// which sorts.  With timing on, one chunk is in flight at a time and HIP
// events bracket the chunk, its last step's kernels and its first step's
// sort.
namespace {

struct ChunkRun {
    uint32_t len = 0;
    int slot = 0;
    int p0 = 0, v0 = 0, q0 = 0;  // state buffers at its start
};

}  // namespace

int step_learned(Engine& e, sph_model* m, uint32_t ncoarse, uint32_t factor, float tol, float skin,
                 sph_learned_report* rep) {
    SPH_TRY_ST(check_model(e, m));
    if (factor == 0 || !(tol > 0.0f) || !(skin >= 0.0f) || !(skin <= 1.0f)) return SPH_ERR_INVALID_ARGUMENT;
    // the surrogate replaces the pair force only: no collider, accelerators,
    // moving wave or colours
    if (e.batch || e.device_clock || e.scenes[0].physics != 0 || changes_per_step(e.scenes[0]) ||
        e.scenes[0].clr_type != SPH_CLR_NONE)
        return SPH_ERR_UNSUPPORTED;
    if (!e.ready() || e.last_kind == 0) return SPH_ERR_NOT_INITIALISED;
    if (!e.buffers_ok()) return SPH_ERR_OUT_OF_MEMORY;
    LearnState* L = learn_state(e);
    if (!L) return SPH_ERR_OUT_OF_MEMORY;
    DeviceGuard guard(e.device);
    SPH_TRY(ensure_step(*L, e.cfg.nscenes));
    if (skin > 0.0f) SPH_TRY(ensure_skin(*L));
    e.drop_graph();
    const uint32_t check = std::min<uint32_t>(e.cfg.tune[2] ? e.cfg.tune[2] : kDefaultCheck, kMaxCheck);
    const size_t N = e.cfg.n;
    const double dt = (double)e.scenes[0].time_step;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    const size_t depth = e.timing ? 1 : 2;
    SPH_TRY(mlp_scale_dt(e.cfg.nscenes, e.dt_dev, (float)factor, L->dt_coarse, e.stream));
    double anchor = 0.0;
    SPH_TRY_ST(plain_deviation(e, *L, &anchor));
    sph_learned_report r{};
    double t_dens = 0.0, t_mlp = 0.0, t_chunks = 0.0, t_build = 0.0, pairs = 0.0;
    uint32_t timed_steps = 0, chunk_steps = 0, timed_builds = 0;

    // Snapshot, steps and result copies of one chunk, then its event.
    auto enqueue_chunk = [&](uint32_t len, int slot, ChunkRun& c) -> int {
        c.len = len;
        c.slot = slot;
        c.p0 = e.cur_pos;
        c.v0 = e.cur_vel;
        c.q0 = e.cur_pid;
        if (e.timing) SPH_TRY(hipEventRecord(e.ev[0], e.stream));
        SPH_TRY(hipMemcpyAsync(L->spos + slot * N, e.pos[c.p0], N * 16, k, e.stream));
        SPH_TRY(hipMemcpyAsync(L->svel + slot * N, e.vel[c.v0], N * 16, k, e.stream));
        SPH_TRY(hipMemcpyAsync(L->spid + slot * N, e.pid[c.q0], N * 4, k, e.stream));
        for (uint32_t i = 0; i < len; ++i) {
            const bool timed = e.timing && (i + 1 == len || (i == 0 && skin > 0.0f));
            SPH_TRY_ST(enqueue_coarse(e, *L, *m, skin, i == 0, L->dev + i, L->ovf + kStepCounters * i, timed));
            if (timed && i == 0 && skin > 0.0f) {
                // the events of the first step are read before the last step reuses them
                SPH_TRY(hipEventSynchronize(e.ev[6]));
                float bt = 0.0f;
                SPH_TRY(hipEventElapsedTime(&bt, e.ev[5], e.ev[6]));
                t_build += bt;
                timed_builds += 1;
            }
        }
        if (e.timing) SPH_TRY(hipEventRecord(e.ev[7], e.stream));
        SPH_TRY(hipMemcpyAsync(L->host + slot * kMaxCheck, L->dev, len * sizeof(double), hipMemcpyDeviceToHost,
                               e.stream));
        SPH_TRY(hipMemcpyAsync(L->host_cnt + slot * kStepCounters * kMaxCheck, L->ovf,
                               kStepCounters * len * sizeof(uint32_t), hipMemcpyDeviceToHost, e.stream));
        SPH_TRY(hipEventRecord(L->done[slot], e.stream));
        return SPH_OK;
    };

    ChunkRun pend[2];
    size_t npend = 0;
    int next_slot = 0;
    uint32_t done = 0, queued = 0;
    while (done < ncoarse) {
        while (npend < depth && done + queued < ncoarse) {
            const uint32_t len = std::min(check, ncoarse - done - queued);
            SPH_TRY_ST(enqueue_chunk(len, next_slot, pend[npend]));
            next_slot ^= 1;
            queued += len;
            npend += 1;
        }
        const ChunkRun c = pend[0];
        pend[0] = pend[1];
        npend -= 1;
        queued -= c.len;
        SPH_TRY(hipEventSynchronize(L->done[c.slot]));
        const double* hd = L->host + c.slot * kMaxCheck;
        const uint32_t* hc = L->host_cnt + c.slot * kStepCounters * kMaxCheck;
        if (skin > 0.0f) r.rebuilds += 1;
        uint32_t bad = c.len;
        bool moved = false;
        for (uint32_t i = 0; i < c.len; ++i) {
            const uint32_t* cnt = hc + kStepCounters * i;
            if (cnt[2] != 0) {
                bad = i;
                moved = true;
                break;
            }
            const double drift = std::fabs(hd[i] - anchor);
            if (drift > (double)tol || cnt[0] != 0) {
                bad = i;
                r.drift_last = drift;
                break;
            }
            r.drift_max = std::max(r.drift_max, drift);
            r.drift_last = drift;
            pairs += cnt[1];
        }
        if (e.timing && bad == c.len) {
            float a = 0.0f, b = 0.0f, t = 0.0f;
            SPH_TRY(hipEventElapsedTime(&a, e.ev[1], e.ev[2]));
            SPH_TRY(hipEventElapsedTime(&b, e.ev[3], e.ev[4]));
            SPH_TRY(hipEventElapsedTime(&t, e.ev[0], e.ev[7]));
            t_dens += a;
            t_mlp += b;
            t_chunks += t;
            timed_steps += 1;
            chunk_steps += c.len;
        }
        if (bad < c.len) {
            // drop the chunk queued behind this one, undo this one
            SPH_TRY(hipStreamSynchronize(e.stream));
            npend = 0;
            queued = 0;
            e.cur_pos = c.p0;
            e.cur_vel = c.v0;
            e.cur_pid = c.q0;
            SPH_TRY(hipMemcpyAsync(e.pos[c.p0], L->spos + c.slot * N, N * 16, k, e.stream));
            SPH_TRY(hipMemcpyAsync(e.vel[c.v0], L->svel + c.slot * N, N * 16, k, e.stream));
            SPH_TRY(hipMemcpyAsync(e.pid[c.q0], L->spid + c.slot * N, N * 4, k, e.stream));
            SPH_TRY(resort_invalidate(e.resort, e.stream));  // slots back in the snapshot's order
            for (uint32_t i = 0; i < bad; ++i)
                SPH_TRY_ST(enqueue_coarse(e, *L, *m, skin, i == 0, L->dev + i, L->ovf + kStepCounters * i, false));
            if (skin > 0.0f && bad > 0) r.rebuilds += 1;
            if (moved) {
                r.early_rebuilds += 1;
            } else {
                for (uint32_t f = 0; f < factor; ++f) {
                    SPH_TRY_ST(e.enqueue_step(false));
                    e.after_steps(1);
                }
                SPH_TRY_ST(plain_deviation(e, *L, &anchor));
                r.fallbacks += 1;
                r.plain_steps += factor;
            }
        }
        const uint32_t kept = bad;
        r.coarse_steps += kept;
        e.steps += kept;
        e.host_clock += dt * factor * kept;
        e.stepped = true;
        if (bad == c.len) e.last_kind = 2;
        done += kept + (bad < c.len && !moved ? 1 : 0);
    }
    r.anchor = anchor;
    r.sim_time = e.host_clock;
    if (r.coarse_steps) r.pairs = pairs / r.coarse_steps;
    if (timed_steps) {
        r.density_ms = t_dens / timed_steps;
        r.mlp_ms = t_mlp / timed_steps;
        r.coarse_ms = t_chunks / chunk_steps;
    }
    if (timed_builds) r.rebuild_ms = t_build / timed_builds;
    if (rep) *rep = r;
    return SPH_OK;
}

int model_create(int device, const sph_mlp_norm* norm, uint64_t seed, sph_model** out) {
    if (!out || !norm) return SPH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPH_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return SPH_ERR_NO_DEVICE;
    sph_model* m = new (std::nothrow) sph_model();
    if (!m) return SPH_ERR_OUT_OF_MEMORY;
    m->device = device;
    m->norm = *norm;
    if (m->norm.rho_scale == 0.0f) m->norm.rho_scale = 1.0f;
    // weights from the seed: SplitMix64 stream, Box-Muller normals
    std::vector<float> w(mlp::kNumParams, 0.0f);
    uint64_t state = seed;
    auto next = [&]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        const double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        const double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    const double s1 = std::sqrt(2.0 / 9.0), s2 = std::sqrt(2.0 / mlp::kHid), s3 = 0.01;
    for (int u = 0; u < mlp::kHid; ++u)
        for (int f = 0; f < mlp::kFeat; ++f)
            if (mlp::is_feature_column(f)) w[mlp::kOffW1 + u * mlp::kFeat + f] = (float)(s1 * normal());
    for (int k = 0; k < mlp::kHid * mlp::kHid; ++k) w[mlp::kOffW2 + k] = (float)(s2 * normal());
    for (int k = 0; k < mlp::kOut * mlp::kW3Cols; ++k) w[mlp::kOffW3 + k] = (float)(s3 * normal());
    DeviceGuard guard(device);
    const size_t bytes = mlp::kNumParams * sizeof(float);
    hipError_t e = device_malloc((void**)&m->params, bytes);
    if (e == hipSuccess) e = device_malloc((void**)&m->adam_m, bytes);
    if (e == hipSuccess) e = device_malloc((void**)&m->adam_v, bytes);
    if (e == hipSuccess) e = device_malloc((void**)&m->pk.w1f, mlp_w1f_bytes());
    if (e == hipSuccess) e = device_malloc((void**)&m->pk.w2f, mlp_w2f_bytes());
    if (e == hipSuccess) e = device_malloc((void**)&m->pk.w2tf, mlp_w2f_bytes());
    if (e == hipSuccess) e = hipMemcpy(m->params, w.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->adam_m, 0, bytes);
    if (e == hipSuccess) e = hipMemset(m->adam_v, 0, bytes);
    if (e == hipSuccess) e = mlp_pack(m->params, m->pk, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        void* ptrs[] = {m->params, m->adam_m, m->adam_v, m->pk.w1f, m->pk.w2f, m->pk.w2tf};
        for (void* q : ptrs) (void)hipFree(q);
        delete m;
        return to_status(e);
    }
    *out = m;
    return SPH_OK;
}

}  // namespace
}  // namespace sph

using namespace sph;

extern "C" {

int sph_model_create(int device, const sph_mlp_norm* norm, uint64_t seed, sph_model** out) {
    return model_create(device, norm, seed, out);
}

int sph_model_destroy(sph_model* m) {
    if (!m) return SPH_OK;
    {
        DeviceGuard guard(m->device);
        (void)hipDeviceSynchronize();
        void* ptrs[] = {m->params, m->adam_m, m->adam_v, m->pk.w1f, m->pk.w2f, m->pk.w2tf};
        for (void* q : ptrs) (void)hipFree(q);
    }
    delete m;
    return SPH_OK;
}

int sph_model_get_norm(sph_model* m, sph_mlp_norm* out) {
    if (!m || !out) return SPH_ERR_INVALID_ARGUMENT;
    *out = m->norm;
    return SPH_OK;
}

int sph_model_get_weights(sph_model* m, float* out) {
    if (!m || !out) return SPH_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(m->device);
    SPH_TRY(hipDeviceSynchronize());
    SPH_TRY(hipMemcpy(out, m->params, mlp::kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_model_set_weights(sph_model* m, const float* in) {
    if (!m || !in) return SPH_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < mlp::kNumParams; ++k)
        if (!std::isfinite(in[k])) return SPH_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(m->device);
    const size_t bytes = mlp::kNumParams * sizeof(float);
    SPH_TRY(hipDeviceSynchronize());
    SPH_TRY(hipMemcpy(m->params, in, bytes, hipMemcpyHostToDevice));
    SPH_TRY(hipMemset(m->adam_m, 0, bytes));
    SPH_TRY(hipMemset(m->adam_v, 0, bytes));
    SPH_TRY(mlp_pack(m->params, m->pk, nullptr));
    SPH_TRY(hipDeviceSynchronize());
    m->adam_t = 0;
    return SPH_OK;
}

int sph_learn_begin(sph_ctx* c, sph_model* m) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return learn_begin(c->eng, m);
}

int sph_learn_add(sph_ctx* c, sph_model* m) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return learn_add(c->eng, m);
}

int sph_learn_train(sph_ctx* c, sph_model* m, uint32_t iters, float lr, double* loss_out) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return learn_train(c->eng, m, iters, lr, loss_out);
}

int sph_learn_gradient(sph_ctx* c, sph_model* m, double* loss_out, float* grad_out) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return learn_gradient(c->eng, m, loss_out, grad_out);
}

int sph_learn_predict(sph_ctx* c, sph_model* m, float* accel4_out) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return learn_predict(c->eng, m, accel4_out);
}

int sph_learn_get_window(sph_ctx* c, int which, float* out) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    return learn_get_window(c->eng, 0, which, out);
}

int sph_batch_learn_begin(sph_batch* b, sph_model* m) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return learn_begin(b->eng, m);
}

int sph_batch_learn_add(sph_batch* b, sph_model* m) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return learn_add(b->eng, m);
}

int sph_batch_learn_train(sph_batch* b, sph_model* m, uint32_t iters, float lr, double* loss_out) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return learn_train(b->eng, m, iters, lr, loss_out);
}

int sph_batch_learn_gradient(sph_batch* b, sph_model* m, double* loss_out, float* grad_out) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return learn_gradient(b->eng, m, loss_out, grad_out);
}

int sph_batch_learn_get_window(sph_batch* b, uint32_t scene, int which, float* out) {
    if (!b) return SPH_ERR_INVALID_ARGUMENT;
    return learn_get_window(b->eng, scene, which, out);
}

int sph_step_learned(sph_ctx* c, sph_model* m, uint32_t ncoarse, uint32_t factor, float tol, float skin,
                     sph_learned_report* out) {
    if (!c) return SPH_ERR_INVALID_ARGUMENT;
    const int st = step_learned(c->eng, m, ncoarse, factor, tol, skin, out);
    if (st != SPH_OK) return st;
    return c->eng.check_order();
}

}  // extern "C"
