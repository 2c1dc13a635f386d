// ensemble.hip -- many independent boxes advanced together on one GPU.
//
// Each box is a full context.  One ensemble timestep is ONE captured graph
// whose branches are the boxes' timesteps: the capture forks from the
// ensemble stream into every box's stream and joins back, so the replay
// lets the boxes' kernels overlap on the GPU (small boxes cannot fill 256
// CUs alone) at the cost of a single launch.  The tracer re-ordering
// (hipCUB radix sort, every sort_every steps) runs eagerly on the ensemble
// stream before the replay.  Graphs are keyed by every box's current
// buffers (the ping-pong parities, tracer columns).  Boxes share nothing; a
// box stopped by the CFL check stays frozen (its kernels return at once)
// while the others continue.
#include "internal.hpp"

#include <string>
#include <vector>

namespace mhit {
int step_async(mhit_ctx* c, int nsteps);     // api.hip: enqueue without waiting
int finish_steps(mhit_ctx* c);               // api.hip: wait, read the scalars back
int step_eager_part(mhit_ctx* c);            // api.hip: tracer re-ordering if due
int step_kernels(mhit_ctx* c);               // api.hip: the kernels of one step
}

using namespace mhit;

struct EnsembleGraph {
    hipGraphExec_t exec = nullptr;
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
};

static void clear_ensemble_graphs(mhit_ensemble* e) {
    for (auto& gr : e->graphs) {
        if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
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
            clear_ensemble_graphs(e);
            slot = &e->graphs[0];
        }
        if ((rc = capture_ensemble_step(e, slot))) return rc;
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
    e->join.assign(nbox, nullptr);
    bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
    for (auto& ev : e->join)
        ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
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
        for (int s = 0; s < nsteps; s++)
            if ((rc = ensemble_graph_step(e))) return rc;
        MHIT_HIP_TRY(hipStreamSynchronize(e->stream));
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

int mhit_ensemble_destroy(mhit_ensemble* e) {
    if (!e) return MHIT_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    clear_ensemble_graphs(e);
    for (mhit_ctx* c : e->boxes) mhit_destroy(c);
    for (hipEvent_t ev : e->join)
        if (ev) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return MHIT_OK;
}

}  // extern "C"
