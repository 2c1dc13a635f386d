// ensemble.hip -- many independent boxes advanced together on one GPU.
//
// Each box is a full context.  One ensemble timestep is ONE captured graph.
// By default (batched) the boxes are grouped by grid size, one graph branch
// per size; within a branch every pass is ONE launch over all boxes of that
// size (their grids stacked along z, each workgroup reading its box's
// arguments from a table), so small boxes fill the 256 CUs together and a
// timestep is a few dozen launches instead of ten per box.  Unbatched, the
// branches are the boxes' own timesteps (the capture forks from the
// ensemble stream into every box's stream and joins back).  The tracer
// re-ordering (a radix sort, every sort_every steps) runs eagerly on
// the ensemble stream before the replay.  Graphs are keyed by every box's
// current buffers (the ping-pong parities, tracer columns).  Boxes share
// nothing; a box stopped by the CFL check stays frozen (its kernels return
// at once) while the others continue.  Both forms compute every box with
// the same kernels and arithmetic as a box stepped alone.
#include "internal.hpp"

#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace mhit {
int step_async(mhit_ctx* c, int nsteps);     // api.hip: enqueue without waiting
int finish_steps(mhit_ctx* c);               // api.hip: wait, read the scalars back
int finish_synced(mhit_ctx* c);              // api.hip: the same once S_host is read
int step_eager_part(mhit_ctx* c);            // api.hip: tracer re-ordering if due
int step_kernels(mhit_ctx* c);               // api.hip: the kernels of one step
void clear_graphs(mhit_ctx* c);              // api.hip: drop a box's captured steps
}

using namespace mhit;

struct EnsembleGraph {
    hipGraphExec_t exec = nullptr;
    char* arena = nullptr;              // argument tables of the batched launches
    size_t arena_cap = 0;
    struct Box {
        double* in[mhit_ctx::kPingPong];    // roles of the ping-pong buffers before the step
        double* out[mhit_ctx::kPingPong];   // ... and after it
        double* tag;                        // first tracer column
        unsigned long long gen;             // the box's graph generation at capture
    };
    std::vector<Box> box;
};

struct mhit_ensemble {
    std::vector<mhit_ctx*> boxes;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    EnsembleGraph graphs[4];
    bool use_graph = true;
    bool batched = true;                    // one launch per pass and grid size
    std::vector<hipStream_t> size_stream;   // graph branch of each grid size
    std::vector<hipEvent_t> size_join;
    // the boxes' scalars read back in one copy: a device table of their
    // DevScalars pointers, a device gather buffer and its pinned host copy
    const DevScalars** s_src = nullptr;
    DevScalars* s_dev = nullptr;
    DevScalars* s_host = nullptr;
};

__global__ void k_gather_scalars(const DevScalars* const* __restrict__ src, DevScalars* __restrict__ dst,
                                 int nb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) dst[b] = *src[b];
}

// After the replays on the ensemble stream: every box's scalars to the host
// with one gather kernel and one copy (one round trip instead of one per
// box), then each box's own end-of-call bookkeeping.
static int finish_ensemble(mhit_ensemble* e, int* status) {
    const int nb = (int)e->boxes.size();
    k_gather_scalars<<<div_up(nb, 64), 64, 0, e->stream>>>(e->s_src, e->s_dev, nb);
    MHIT_HIP_TRY(hipGetLastError());
    MHIT_HIP_TRY(hipMemcpyAsync(e->s_host, e->s_dev, nb * sizeof(DevScalars), hipMemcpyDeviceToHost, e->stream));
    MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
    *status = MHIT_OK;
    for (int b = 0; b < nb; b++) {
        mhit_ctx* c = e->boxes[b];
        c->S_host = e->s_host[b];
        const int rc = finish_synced(c);
        if (rc == MHIT_ERR_CFL) *status = MHIT_ERR_CFL;
        else if (rc) return rc;
    }
    return MHIT_OK;
}

// The caller has synchronised the ensemble stream (no replay in flight).
static void clear_ensemble_graphs(mhit_ensemble* e) {
    for (auto& gr : e->graphs) {
        if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
        if (gr.arena) (void)hipFree(gr.arena);
        gr = EnsembleGraph{};
    }
}

static bool graph_matches(const EnsembleGraph& gr, const mhit_ensemble* e) {
    if (!gr.exec) return false;
    for (size_t b = 0; b < e->boxes.size(); b++) {
        mhit_ctx* c = e->boxes[b];
        double* cur[mhit_ctx::kPingPong];
        pingpong_current(c, cur);
        const EnsembleGraph::Box& g = gr.box[b];
        if (g.tag != c->ptag || g.gen != c->graph_gen) return false;
        for (int p = 0; p < mhit_ctx::kPingPong; p++)
            if (g.in[p] != cur[p]) return false;
    }
    return true;
}

static int capture_ensemble_step(mhit_ensemble* e, EnsembleGraph* slot) {
    const size_t nb = e->boxes.size();
    slot->box.resize(nb);
    for (size_t b = 0; b < nb; b++) {
        pingpong_current(e->boxes[b], slot->box[b].in);
        slot->box[b].tag = e->boxes[b]->ptag;
        slot->box[b].gen = e->boxes[b]->graph_gen;
    }
    hipGraph_t graph = nullptr;
    MHIT_HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipEventRecord(e->fork, e->stream);
    int rc = MHIT_OK;
    for (size_t b = 0; b < nb && err == hipSuccess && rc == MHIT_OK; b++) {
        mhit_ctx* c = e->boxes[b];
        err = hipStreamWaitEvent(c->stream, e->fork, 0);   // the box stream joins the capture
        if (err == hipSuccess) rc = step_kernels(c);        // performs the role swaps on the host
        if (err == hipSuccess && rc == MHIT_OK) err = hipEventRecord(e->join[b], c->stream);
        if (err == hipSuccess && rc == MHIT_OK) err = hipStreamWaitEvent(e->stream, e->join[b], 0);
    }
    hipError_t end = hipStreamEndCapture(e->stream, &graph);
    // the capture advanced each box's host-side buffer roles; undo it so the
    // launch applies the step once
    for (size_t b = 0; b < nb; b++) pingpong_undo(e->boxes[b], slot->box[b].in, slot->box[b].out);
    if (rc != MHIT_OK || err != hipSuccess || end != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != MHIT_OK) return rc;
        MHIT_HIP_TRY(err != hipSuccess ? err : end);
    }
    err = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MHIT_HIP_TRY(err);
    return MHIT_OK;
}

// Boxes in order of first appearance of key(box), grouped.
template <typename K>
static std::vector<std::vector<mhit_ctx*>> group_by(const std::vector<mhit_ctx*>& boxes, K key) {
    std::vector<std::vector<mhit_ctx*>> groups;
    for (mhit_ctx* c : boxes) {
        bool placed = false;
        for (auto& gp : groups)
            if (key(gp[0]) == key(c)) { gp.push_back(c); placed = true; break; }
        if (!placed) groups.push_back({c});
    }
    return groups;
}

// One timestep of boxes of one grid size on stream st, each pass batched
// over the boxes that share its kernel and launch shape; the same order of
// passes as a single box (api.hip, enqueue_step_kernels).
static int enqueue_size_class(const std::vector<mhit_ctx*>& boxes, ArgArena& ar, hipStream_t st) {
    int rc;
    for (mhit_ctx* c : boxes)
        if (c->P.partflag && c->np > 0 && (rc = launch_tracers(c, st))) return rc;
    std::vector<mhit_ctx*> phase;
    for (mhit_ctx* c : boxes)
        if (c->P.phiflag) phase.push_back(c);
    auto ac_key = [](mhit_ctx* c) { return std::make_tuple(c->tune.ac_mz, c->tune.ac_my, c->tune.xcd); };
    for (auto& gp : group_by(phase, ac_key))
        if ((rc = launch_acdi_march_batch(gp.data(), (int)gp.size(), ar, st))) return rc;
    auto ns_key = [](mhit_ctx* c) {
        return std::make_tuple(c->P.phiflag != 0, c->fp32, c->tune.ns_mz, c->tune.ns_my, c->tune.xcd);
    };
    for (auto& gp : group_by(boxes, ns_key))
        if ((rc = launch_ns_march_batch(gp.data(), (int)gp.size(), ar, st))) return rc;
    auto prec_key = [](mhit_ctx* c) {
        return std::make_tuple(c->fp32, c->tune.fft_wave, fft_width(c, 0), fft_width(c, 1), fft_width(c, 2),
                               c->tune.fft_xy, c->tune.fft_xplan);
    };
    const auto by_prec = group_by(boxes, prec_key);
    for (auto& gp : by_prec)
        if ((rc = poisson_project_batch(gp.data(), (int)gp.size(), ar, st))) return rc;
    for (auto& gp : by_prec)
        if ((rc = launch_correct_batch(gp.data(), (int)gp.size(), ar, st))) return rc;
    return launch_finalize_batch(boxes.data(), (int)boxes.size(), ar, st);
}

// Launch shapes for boxes batched in an ensemble: 64 x 16 tiles (96 x 10
// where n is not a multiple of 64) in XCD order and z
// chunks of min(n, 128) planes, then chunks halved size class by size
// class (the class with the longest chunks first, ties to the larger grid;
// never below 8 planes) until the NS launches of one step hold at least
// four workgroups per CU in total.  Measured on 16 boxes of one size: 64^3
// 0.88 -> 0.30 ms per ensemble step (8-plane chunks), 96^3 1.57 -> 0.96 ms
// (12), 128^3 2.52 -> 1.76 ms (32); the 16 x 64^3 + 16 x 96^3 + 16 x 128^3
// mix, which fills the GPU already at full chunks through its concurrent
// size classes, keeps 2.9 ms within 2%.  MHIT_TUNE overrides apply last.
static void batched_shapes(const std::vector<mhit_ctx*>& boxes) {
    for (mhit_ctx* c : boxes) default_tuning(c, true, false);
    int cus = 0;
    if (boxes.empty() ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, boxes[0]->device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    const long long target = 4LL * cus;
    std::vector<std::pair<int, int>> chunk;   // (n, z chunk) per size class
    for (mhit_ctx* c : boxes) {
        bool seen = false;
        for (auto& sc : chunk) seen = seen || sc.first == c->n;
        if (!seen) chunk.push_back({c->n, c->tune.ns_mz});
    }
    auto zc_of = [&](int n) {
        for (auto& sc : chunk)
            if (sc.first == n) return sc.second;
        return n;
    };
    auto workgroups = [&]() {
        long long t = 0;
        for (mhit_ctx* c : boxes)
            t += (long long)div_up(c->n, tile_cols(c->tune.ns_my)) * div_up(c->n, c->tune.ns_my) *
                 div_up(c->n, zc_of(c->n));
        return t;
    };
    while (workgroups() < target) {
        std::pair<int, int>* pick = nullptr;
        for (auto& sc : chunk)
            if (!pick || sc.second > pick->second || (sc.second == pick->second && sc.first > pick->first))
                pick = &sc;
        if (!pick || pick->second / 2 < 8) break;
        pick->second /= 2;
    }
    for (mhit_ctx* c : boxes) {
        c->tune.ns_mz = c->tune.ac_mz = zc_of(c->n);
        tuning_env(c);
    }
}

static bool can_batch(const mhit_ensemble* e) {
    if (!e->batched) return false;
    for (mhit_ctx* c : e->boxes)
        if (!c->march) return false;
    return true;
}

static int capture_ensemble_batched(mhit_ensemble* e, EnsembleGraph* slot) {
    const size_t nb = e->boxes.size();
    const auto classes = group_by(e->boxes, [](mhit_ctx* c) { return c->n; });
    while (e->size_stream.size() < classes.size()) {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        MHIT_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(s);
            set_error("ensemble: event creation failed");
            return MHIT_ERR_HIP;
        }
        e->size_stream.push_back(s);   // the pair is kept together
        e->size_join.push_back(ev);
    }
    // at most ~20 tables per size class and a few hundred bytes per box and pass
    const size_t cap = 8192 * nb + 16384 * classes.size();
    if (slot->arena_cap < cap) {
        if (slot->arena) (void)hipFree(slot->arena);
        slot->arena = nullptr;
        slot->arena_cap = 0;
        MHIT_HIP_TRY(hipMalloc((void**)&slot->arena, cap));
        slot->arena_cap = cap;
    }
    ArgArena ar;
    ar.dev = slot->arena;
    ar.cap = slot->arena_cap;
    slot->box.resize(nb);
    for (size_t b = 0; b < nb; b++) {
        pingpong_current(e->boxes[b], slot->box[b].in);
        slot->box[b].tag = e->boxes[b]->ptag;
        slot->box[b].gen = e->boxes[b]->graph_gen;
    }
    hipGraph_t graph = nullptr;
    MHIT_HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipEventRecord(e->fork, e->stream);
    int rc = MHIT_OK;
    for (size_t k = 0; k < classes.size() && err == hipSuccess && rc == MHIT_OK; k++) {
        hipStream_t st = e->size_stream[k];
        err = hipStreamWaitEvent(st, e->fork, 0);
        if (err == hipSuccess) rc = enqueue_size_class(classes[k], ar, st);   // role swaps on the host
        if (err == hipSuccess && rc == MHIT_OK) err = hipEventRecord(e->size_join[k], st);
        if (err == hipSuccess && rc == MHIT_OK) err = hipStreamWaitEvent(e->stream, e->size_join[k], 0);
    }
    hipError_t end = hipStreamEndCapture(e->stream, &graph);
    for (size_t b = 0; b < nb; b++) pingpong_undo(e->boxes[b], slot->box[b].in, slot->box[b].out);
    if (rc != MHIT_OK || err != hipSuccess || end != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != MHIT_OK) return rc;
        MHIT_HIP_TRY(err != hipSuccess ? err : end);
    }
    err = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MHIT_HIP_TRY(err);
    // the tables, before the first replay (in order with it on the ensemble
    // stream, and complete before the host staging goes away)
    MHIT_HIP_TRY(hipMemcpyAsync(slot->arena, ar.host.data(), ar.host.size(), hipMemcpyHostToDevice,
                                e->stream));
    MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
    return MHIT_OK;
}

static int ensemble_graph_step(mhit_ensemble* e) {
    int rc;
    // eager part on the ensemble stream, in order with the replays
    for (mhit_ctx* c : e->boxes) {
        if (c->S_host.halted) continue;
        hipStream_t own = c->stream;
        c->stream = e->stream;
        rc = step_eager_part(c);
        c->stream = own;
        if (rc) return rc;
    }
    EnsembleGraph* slot = nullptr;
    for (auto& gr : e->graphs)
        if (graph_matches(gr, e)) { slot = &gr; break; }
    if (!slot) {
        for (auto& gr : e->graphs)
            if (!gr.exec) { slot = &gr; break; }
        if (!slot) {
            MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
            clear_ensemble_graphs(e);
            slot = &e->graphs[0];
        }
        if ((rc = can_batch(e) ? capture_ensemble_batched(e, slot) : capture_ensemble_step(e, slot)))
            return rc;
    }
    MHIT_HIP_TRY(hipGraphLaunch(slot->exec, e->stream));
    for (size_t b = 0; b < e->boxes.size(); b++)
        pingpong_apply(e->boxes[b], slot->box[b].in, slot->box[b].out);
    return MHIT_OK;
}

extern "C" {

int mhit_ensemble_create(int32_t nbox, const mhit_params* params, mhit_ensemble** out) {
    if (!out || nbox <= 0 || !params) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    *out = nullptr;
    mhit_ensemble* e = new mhit_ensemble();
    for (int32_t b = 0; b < nbox; b++) {
        mhit_ctx* c = nullptr;
        int rc = mhit_create(&params[b], &c);
        if (rc) {
            set_error("box " + std::to_string(b) + ": " + mhit_last_error());
            mhit_ensemble_destroy(e);
            return rc;
        }
        e->boxes.push_back(c);
        e->use_graph = e->use_graph && c->use_graph;
    }
    batched_shapes(e->boxes);   // the shapes for batched launches
    e->join.assign(nbox, nullptr);
    bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
    for (auto& ev : e->join)
        ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    std::vector<const DevScalars*> src(nbox);
    for (int32_t b = 0; b < nbox; b++) src[b] = e->boxes[b]->S;
    ok = ok && hipMalloc((void**)&e->s_src, nbox * sizeof(DevScalars*)) == hipSuccess &&
         hipMalloc((void**)&e->s_dev, nbox * sizeof(DevScalars)) == hipSuccess &&
         hipHostMalloc((void**)&e->s_host, nbox * sizeof(DevScalars), hipHostMallocDefault) == hipSuccess &&
         hipMemcpy(e->s_src, src.data(), nbox * sizeof(DevScalars*), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        set_error("ensemble stream/event creation failed");
        mhit_ensemble_destroy(e);
        return MHIT_ERR_HIP;
    }
    *out = e;
    return MHIT_OK;
}

int mhit_ensemble_box(mhit_ensemble* e, int32_t i, mhit_ctx** box) {
    if (!e || !box || i < 0 || i >= (int32_t)e->boxes.size()) { set_error("bad box index"); return MHIT_ERR_ARG; }
    *box = e->boxes[i];
    return MHIT_OK;
}

int mhit_ensemble_step(mhit_ensemble* e, int nsteps) {
    if (!e || nsteps < 0) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    int rc;
    bool graph = e->use_graph;
    for (mhit_ctx* c : e->boxes) graph = graph && !c->prof;
    if (graph) {
        // work queued on the boxes' own streams (uploads, initial
        // conditions) comes first
        for (size_t b = 0; b < e->boxes.size(); b++) {
            MHIT_HIP_TRY(hipEventRecord(e->join[b], e->boxes[b]->stream));
            MHIT_HIP_TRY(hipStreamWaitEvent(e->stream, e->join[b], 0));
        }
        for (int s = 0; s < nsteps; s++) {
            Range r("Iteration");
            if ((rc = ensemble_graph_step(e))) return rc;
        }
        int status = MHIT_OK;
        if ((rc = finish_ensemble(e, &status))) return rc;
        if (status == MHIT_ERR_CFL) set_error("at least one box stopped on cou > 7");
        return status;
    } else {
        // one stream per box, interleaved step by step so that no stream
        // runs far ahead (a box already stopped on the CFL check is skipped)
        for (int s = 0; s < nsteps; s++)
            for (mhit_ctx* c : e->boxes)
                if (!c->S_host.halted && (rc = step_async(c, 1))) return rc;
    }
    int status = MHIT_OK;
    for (mhit_ctx* c : e->boxes) {
        rc = finish_steps(c);
        if (rc == MHIT_ERR_CFL) status = MHIT_ERR_CFL;
        else if (rc) return rc;
    }
    if (status == MHIT_ERR_CFL) set_error("at least one box stopped on cou > 7");
    return status;
}

int mhit_ensemble_set_batched(mhit_ensemble* e, int32_t on) {
    if (!e) { set_error("null ensemble"); return MHIT_ERR_ARG; }
    MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
    clear_ensemble_graphs(e);
    e->batched = on != 0;
    // the launch shapes suited to the mode: a box on its own branch needs
    // the many-workgroup shapes of a box alone
    if (e->batched) batched_shapes(e->boxes);
    for (mhit_ctx* c : e->boxes) {
        if (!e->batched) default_tuning(c, false);
        clear_graphs(c);
    }
    return MHIT_OK;
}

// ms per ensemble step over `steps` replays after one untimed step (the
// graph capture); the caller restores the boxes' states afterwards.
static int time_ensemble(mhit_ensemble* e, int steps, double* ms) {
    MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
    clear_ensemble_graphs(e);
    for (mhit_ctx* c : e->boxes) clear_graphs(c);   // launch shapes are baked into graphs
    int rc = mhit_ensemble_step(e, 1);
    if (rc) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    MHIT_HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        set_error("ensemble autotune: event creation failed");
        return MHIT_ERR_HIP;
    }
    hipError_t err = hipEventRecord(a, e->stream);
    for (int s = 0; s < steps && err == hipSuccess && rc == MHIT_OK; s++) rc = ensemble_graph_step(e);
    if (err == hipSuccess && rc == MHIT_OK) err = hipEventRecord(b, e->stream);
    if (err == hipSuccess && rc == MHIT_OK) err = hipEventSynchronize(b);
    float f = 0.0f;
    if (err == hipSuccess && rc == MHIT_OK) err = hipEventElapsedTime(&f, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    MHIT_HIP_TRY(err);
    int status = MHIT_OK;
    if ((rc = finish_ensemble(e, &status))) return rc;
    if (status != MHIT_OK) {
        set_error("ensemble autotune: a trial stopped on cou > 7");
        return status;
    }
    *ms = (double)f / steps;
    return MHIT_OK;
}

int mhit_ensemble_autotune(mhit_ensemble* e, int steps, double* ms_per_step) {
    if (!e || steps <= 0) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    if (!e->batched || !e->use_graph || !can_batch(e)) {
        set_error("ensemble autotune: batched graph mode with the fused passes only");
        return MHIT_ERR_STATE;
    }
    for (mhit_ctx* c : e->boxes)
        if (c->prof) { set_error("ensemble autotune: disable profiling first"); return MHIT_ERR_STATE; }
    MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
    for (mhit_ctx* c : e->boxes)
        if (c->S_host.halted) { set_error("stopped: cou > 7"); return MHIT_ERR_CFL; }
    const size_t nb = e->boxes.size();
    std::vector<StateCopy*> snaps(nb, nullptr);
    int rc = MHIT_OK;
    for (size_t b = 0; b < nb && rc == MHIT_OK; b++) rc = state_copy_take(e->boxes[b], &snaps[b]);
    auto restore_all = [&]() {
        int r = MHIT_OK;
        for (size_t b = 0; b < nb; b++) {
            const int q = state_copy_restore(e->boxes[b], snaps[b]);
            if (r == MHIT_OK) r = q;
        }
        return r;
    };
    auto trial = [&](double* ms) {
        int r = time_ensemble(e, steps, ms);
        const int q = restore_all();
        return r != MHIT_OK ? r : q;
    };
    double best_ms = 0.0;
    if (rc == MHIT_OK) rc = trial(&best_ms);
    std::vector<Tuning> start(nb);
    for (size_t b = 0; b < nb; b++) start[b] = e->boxes[b]->tune;
    // one size and precision at a time, one setting at a time (the same
    // value for every box of the class), keeping the fastest
    const auto classes = group_by(e->boxes, [](mhit_ctx* c) { return std::make_pair(c->n, c->fp32); });
    for (const auto& cls : classes) {
        if (rc != MHIT_OK) break;
        mhit_ctx* c0 = cls[0];
        const int n = c0->n;
        std::vector<int> zc = {c0->tune.ns_mz};
        for (int z = n; z >= 8; z /= 2)
            if (z != c0->tune.ns_mz) zc.push_back(z);
        auto widths = [&](int pass) {
            int w[3];
            const int k = fft_widths(n, c0->fp32, pass, w);
            return std::vector<int>(w, w + k);
        };
        struct Param {
            int Tuning::*field;
            int Tuning::*also;   // set together (the two fused passes' z chunk)
            std::vector<int> cand;
        };
        const std::vector<Param> params = {
            {&Tuning::ns_mz, &Tuning::ac_mz, zc},
            {&Tuning::ns_my, nullptr, {16, 8, 32, 10}},
            {&Tuning::ac_my, nullptr, {16, 8, 32, 10}},
            {&Tuning::xcd, nullptr, {0, 1}},
            {&Tuning::fft_wave, nullptr, {1, 0}},
            {&Tuning::fft_ycols, nullptr, widths(0)},
            {&Tuning::fft_zcols, nullptr, widths(1)},
            {&Tuning::fft_xrows, nullptr, widths(2)},
            {&Tuning::fft_xy, nullptr, fft_xy_supported(n, c0->fp32) ? std::vector<int>{1, 0} : std::vector<int>{}},
            {&Tuning::fft_xplan, nullptr, fft_xplan_differs(n) ? std::vector<int>{0, 1} : std::vector<int>{}},
        };
        for (const Param& p : params) {
            if (rc != MHIT_OK) break;
            if (p.cand.size() < 2) continue;
            const int keep = c0->tune.*(p.field);
            int best = keep;
            double bms = best_ms;   // the current setting's time
            for (int v : p.cand) {
                if (v == keep) continue;
                for (mhit_ctx* c : cls) {
                    c->tune.*(p.field) = v;
                    if (p.also) c->tune.*(p.also) = v;
                }
                double ms = 0.0;
                if ((rc = trial(&ms)) != MHIT_OK) break;
                if (ms < bms) { bms = ms; best = v; }
            }
            for (mhit_ctx* c : cls) {
                c->tune.*(p.field) = best;
                if (p.also) c->tune.*(p.also) = best;
            }
            best_ms = bms;
        }
    }
    if (rc == MHIT_OK) {
        // trials of a few steps are noisy: the tuned settings and the
        // starting ones, timed alternately twice each, and the faster kept
        std::vector<Tuning> tuned(nb);
        for (size_t b = 0; b < nb; b++) tuned[b] = e->boxes[b]->tune;
        double t_tuned = 1e300, t_start = 1e300;
        for (int rep = 0; rep < 2 && rc == MHIT_OK; rep++) {
            for (int which = 0; which < 2 && rc == MHIT_OK; which++) {
                for (size_t b = 0; b < nb; b++) e->boxes[b]->tune = which ? start[b] : tuned[b];
                double ms = 0.0;
                rc = trial(&ms);
                double& t = which ? t_start : t_tuned;
                if (rc == MHIT_OK && ms < t) t = ms;
            }
        }
        const bool keep = t_tuned <= t_start;
        for (size_t b = 0; b < nb; b++) e->boxes[b]->tune = keep ? tuned[b] : start[b];
        best_ms = keep ? t_tuned : t_start;
    }
    for (StateCopy* s : snaps) state_copy_free(s);
    MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
    clear_ensemble_graphs(e);
    for (mhit_ctx* c : e->boxes) clear_graphs(c);
    if (rc != MHIT_OK) return rc;
    if (ms_per_step) *ms_per_step = best_ms;
    return MHIT_OK;
}

int mhit_ensemble_destroy(mhit_ensemble* e) {
    if (!e) return MHIT_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    clear_ensemble_graphs(e);
    for (hipEvent_t ev : e->size_join)
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t st : e->size_stream)
        if (st) (void)hipStreamDestroy(st);
    for (mhit_ctx* c : e->boxes) mhit_destroy(c);
    for (hipEvent_t ev : e->join)
        if (ev) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->s_src) (void)hipFree(e->s_src);
    if (e->s_dev) (void)hipFree(e->s_dev);
    if (e->s_host) (void)hipHostFree(e->s_host);
    delete e;
    return MHIT_OK;
}

}  // extern "C"
