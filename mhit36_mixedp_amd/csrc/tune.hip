// tune.hip -- launch-shape autotuner for one context.
//
// The reference tunes its decomposition at start-up with cuDecomp's
// autotuner (multi/main.f90:80-136: process grid, transpose and halo
// backends).  On one GPU the free choices are launch shapes instead: the
// tile shape (64x16, 64x8 or 32x32) and the z planes per workgroup of the two
// fused marching passes, and the XCD-aware tile order.  None of them
// changes a result bit (every cell is
// computed by the same arithmetic in any shape; the correction's reduction
// grid is fixed), so the tuner may pick freely.  The tracer re-ordering
// period (sort_every) is settable but not tuned: its pay-off depends on how
// far tracers move, which a short trial does not see.
//
// The FFT passes are tuned too: the column transforms' form (fft_wave)
// and the launch widths (columns per workgroup of the y and z passes, row
// pairs per workgroup of the x passes; each width a compiled kernel, so
// the choice also sets the launch bounds: 64 threads per unit of width and
// 1/8 of the transform length, and the LDS per workgroup).  The wavefront
// form and the register-direct form compute the same DFT in a different
// operation order, so that choice changes results at the rounding level;
// a width changes the form only where it moves a pass to or from eight
// columns (the wavefront form's width), and otherwise no bit.
//
// mhit_autotune times candidate values one parameter at a time on the
// context's own state, with the production path (captured graphs), and
// restores the state after every trial: the run continues as if the tuner
// had not been called.
#include "internal.hpp"

#include <algorithm>
#include <vector>

namespace mhit {
int step_async(mhit_ctx* c, int nsteps);
int finish_steps(mhit_ctx* c);
int sync_scalars_public(mhit_ctx* c);
int push_scalars_public(mhit_ctx* c);
void clear_graphs(mhit_ctx* c);
}

using namespace mhit;

namespace {

// Device copy of everything a timestep reads, by role (the ping-pong and
// re-ordering swap buffers, so restore goes to the current owner of a role),
// and of the pressure, which a download or dump returns between steps.
struct Snapshot {
    std::vector<double*> copies;
    std::vector<size_t> counts;
    void* pres = nullptr;   // the last projection's pressure, in its own precision
    size_t pres_bytes = 0;
    DevScalars S{};
    long long steps_since_sort = 0;
};

size_t pres_bytes(const mhit_ctx* c) { return c->N3 * (c->fp32 ? sizeof(float) : sizeof(double)); }

std::vector<double**> roles(mhit_ctx* c) {
    std::vector<double**> r = {&c->u, &c->v, &c->w, &c->rhsu_o, &c->rhsv_o, &c->rhsw_o};
    if (c->P.phiflag) { r.push_back(&c->phi); r.push_back(&c->rhsphi_o); }
    if (c->np > 0)
        for (double** col : {&c->ptag, &c->px, &c->py, &c->pz, &c->pu1, &c->pv1, &c->pw1})
            r.push_back(col);
    return r;
}

size_t role_count(mhit_ctx* c, size_t k) { return k < (c->P.phiflag ? 8u : 6u) ? c->N3 : (size_t)c->np; }

void release(Snapshot& s) {
    for (double* p : s.copies) (void)hipFree(p);
    s.copies.clear();
    if (s.pres) (void)hipFree(s.pres);
    s.pres = nullptr;
}

int take(mhit_ctx* c, Snapshot& s) {
    int rc;
    if ((rc = sync_scalars_public(c))) return rc;
    s.S = c->S_host;
    s.steps_since_sort = c->steps_since_sort;
    auto r = roles(c);
    for (size_t k = 0; k < r.size(); k++) {
        const size_t cnt = role_count(c, k);
        double* d = nullptr;
        if (hipMalloc((void**)&d, cnt * sizeof(double)) != hipSuccess) {
            release(s);
            set_error("autotune: no device memory for the state snapshot");
            return MHIT_ERR_NOMEM;
        }
        s.copies.push_back(d);
        s.counts.push_back(cnt);
        if (hipMemcpyAsync(d, *r[k], cnt * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
            release(s);
            set_error("autotune: state snapshot copy failed");
            return MHIT_ERR_HIP;
        }
    }
    s.pres_bytes = pres_bytes(c);
    if (hipMalloc(&s.pres, s.pres_bytes) != hipSuccess) {
        s.pres = nullptr;
        release(s);
        set_error("autotune: no device memory for the state snapshot");
        return MHIT_ERR_NOMEM;
    }
    if (hipMemcpyAsync(s.pres, c->pres, s.pres_bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
        release(s);
        set_error("autotune: state snapshot copy failed");
        return MHIT_ERR_HIP;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        release(s);
        set_error("autotune: state snapshot copy failed");
        return MHIT_ERR_HIP;
    }
    return MHIT_OK;
}

int restore(mhit_ctx* c, const Snapshot& s) {
    auto r = roles(c);
    for (size_t k = 0; k < r.size(); k++)
        MHIT_HIP_TRY(hipMemcpyAsync(*r[k], s.copies[k], s.counts[k] * sizeof(double),
                                    hipMemcpyDeviceToDevice, c->stream));
    MHIT_HIP_TRY(hipMemcpyAsync(c->pres, s.pres, s.pres_bytes, hipMemcpyDeviceToDevice, c->stream));
    c->S_host = s.S;
    c->steps_since_sort = s.steps_since_sort;
    int rc;
    if ((rc = push_scalars_public(c))) return rc;
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    return MHIT_OK;
}

// ms per step of `steps` steps after one untimed step (graph capture).
int time_steps(mhit_ctx* c, int steps, double* ms) {
    int rc;
    clear_graphs(c);   // launch shapes are baked into captured graphs
    if ((rc = step_async(c, 1))) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    MHIT_HIP_TRY(hipEventCreate(&a));
    MHIT_HIP_TRY(hipEventCreate(&b));
    hipError_t e = hipEventRecord(a, c->stream);
    if (e == hipSuccess) rc = step_async(c, steps);
    if (e == hipSuccess && rc == MHIT_OK) e = hipEventRecord(b, c->stream);
    if (e == hipSuccess && rc == MHIT_OK) e = hipEventSynchronize(b);
    float f = 0.0f;
    if (e == hipSuccess && rc == MHIT_OK) e = hipEventElapsedTime(&f, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    MHIT_HIP_TRY(e);
    if ((rc = finish_steps(c))) return rc;   // MHIT_ERR_CFL if a trial stopped the run
    *ms = (double)f / steps;
    return MHIT_OK;
}

void to_public(const mhit_ctx* c, mhit_tuning* out, double ms) {
    const Tuning& t = c->tune;
    out->ns_mz = t.ns_mz;
    out->ac_mz = t.ac_mz;
    out->xcd = t.xcd;
    out->sort_every = t.sort_every;
    out->ns_my = t.ns_my;
    out->ac_my = t.ac_my;
    out->fft_wave = t.fft_wave;
    out->fft_ycols = fft_width(c, 0);
    out->fft_zcols = fft_width(c, 1);
    out->fft_xrows = fft_width(c, 2);
    out->fft_xy = (t.fft_xy && fft_xy_supported(c->n, c->fp32)) ? 1 : 0;
    out->fft_xplan = (t.fft_xplan && fft_xplan_differs(c->n)) ? 1 : 0;
    out->ms_per_step = ms;
}

// 0 (the default) or a width compiled for the context's size and precision
bool valid_width(const mhit_ctx* c, int pass, int w) {
    if (w == 0) return true;
    int ws[3];
    const int k = fft_widths(c->n, c->fp32, pass, ws);
    for (int i = 0; i < k; i++)
        if (ws[i] == w) return true;
    return false;
}

bool valid(const mhit_ctx* c, const mhit_tuning* t) {
    auto rows = [](int r) { return r == 8 || r == 16 || r == 32 || r == 10; };
    return t->ns_mz > 0 && t->ac_mz > 0 && (t->xcd == 0 || t->xcd == 1) && t->sort_every >= 0 &&
           rows(t->ns_my) && rows(t->ac_my) && (t->fft_wave == 0 || t->fft_wave == 1) &&
           valid_width(c, 0, t->fft_ycols) && valid_width(c, 1, t->fft_zcols) && valid_width(c, 2, t->fft_xrows) &&
           (t->fft_xy == 0 || t->fft_xy == 1) && (t->fft_xplan == 0 || t->fft_xplan == 1);
}

}  // namespace

namespace mhit {
// A context's state copy for other tuners (the ensemble's).
struct StateCopy {
    Snapshot snap;
};
int state_copy_take(mhit_ctx* c, StateCopy** out) {
    StateCopy* s = new StateCopy();
    const int rc = take(c, s->snap);
    if (rc) { delete s; return rc; }
    *out = s;
    return MHIT_OK;
}
int state_copy_restore(mhit_ctx* c, const StateCopy* s) { return restore(c, s->snap); }
void state_copy_free(StateCopy* s) {
    if (!s) return;
    release(s->snap);
    delete s;
}
}  // namespace mhit

extern "C" {

int mhit_get_tuning(mhit_ctx* c, mhit_tuning* out) {
    if (!c || !out) { set_error("null argument"); return MHIT_ERR_ARG; }
    to_public(c, out, 0.0);
    return MHIT_OK;
}

int mhit_set_tuning(mhit_ctx* c, const mhit_tuning* t) {
    if (!c || !t || !valid(c, t)) { set_error("bad tuning"); return MHIT_ERR_ARG; }
    c->tune.ns_mz = t->ns_mz;
    c->tune.ac_mz = t->ac_mz;
    c->tune.xcd = t->xcd;
    c->tune.sort_every = t->sort_every;
    c->tune.ns_my = t->ns_my;
    c->tune.ac_my = t->ac_my;
    c->tune.fft_wave = t->fft_wave;
    c->tune.fft_ycols = t->fft_ycols;
    c->tune.fft_zcols = t->fft_zcols;
    c->tune.fft_xrows = t->fft_xrows;
    c->tune.fft_xy = t->fft_xy;
    c->tune.fft_xplan = t->fft_xplan;
    clear_graphs(c);
    return MHIT_OK;
}

int mhit_autotune(mhit_ctx* c, int steps, mhit_tuning* out) {
    if (!c || steps <= 0) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    int rc;
    if ((rc = sync_scalars_public(c))) return rc;
    if (c->S_host.halted) { set_error("stopped: cou > 7"); return MHIT_ERR_CFL; }
    if (c->prof) { set_error("autotune: disable profiling first"); return MHIT_ERR_STATE; }
    Snapshot snap;
    if ((rc = take(c, snap))) return rc;

    const int n = c->n;
    auto zlist = [&](std::initializer_list<int> vals) {
        std::vector<int> v;
        for (int z : vals)
            if (z <= n) v.push_back(z);
        if (v.empty()) v.push_back(n);
        return v;
    };
    auto widths = [&](int pass) {
        int w[3];
        const int k = fft_widths(n, c->fp32, pass, w);
        return std::vector<int>(w, w + k);
    };
    struct Param { int* field; std::vector<int> cand; bool active; };
    std::vector<Param> params = {
        {&c->tune.ns_my, {16, 8, 32, 10}, true},
        {&c->tune.ns_mz, zlist({4, 8, 16, 32, 64, 128}), true},
        {&c->tune.ac_my, {16, 8, 32, 10}, c->P.phiflag != 0},
        {&c->tune.ac_mz, zlist({4, 8, 16, 32, 64, 128}), c->P.phiflag != 0},
        {&c->tune.xcd, {0, 1}, true},
        {&c->tune.fft_wave, {1, 0}, true},
        {&c->tune.fft_ycols, widths(0), widths(0).size() > 1},
        {&c->tune.fft_zcols, widths(1), widths(1).size() > 1},
        {&c->tune.fft_xrows, widths(2), widths(2).size() > 1},
        {&c->tune.fft_xy, {1, 0}, fft_xy_supported(n, c->fp32)},
        {&c->tune.fft_xplan, {0, 1}, fft_xplan_differs(n)},
    };
    double best_ms = 0.0;
    rc = MHIT_OK;
    const Tuning start = c->tune;
    for (Param& p : params) {
        if (!p.active || rc) continue;
        int best = *p.field;
        double bms = 1e300;
        for (int v : p.cand) {
            *p.field = v;
            double ms = 0.0;
            rc = time_steps(c, steps, &ms);
            const int rr = restore(c, snap);
            if (!rc) rc = rr;
            if (rc) break;
            if (ms < bms) { bms = ms; best = v; }
        }
        *p.field = best;
        best_ms = bms;
    }
    if (!rc) {
        // trials of a few steps are noisy: the tuned settings and the
        // starting ones, timed alternately twice each, and the faster kept
        const Tuning tuned = c->tune;
        double t_tuned = 1e300, t_start = 1e300;
        for (int rep = 0; rep < 2 && !rc; rep++) {
            for (int which = 0; which < 2 && !rc; which++) {
                c->tune = which ? start : tuned;
                double ms = 0.0;
                rc = time_steps(c, steps, &ms);
                const int rr = restore(c, snap);
                if (!rc) rc = rr;
                double& t = which ? t_start : t_tuned;
                if (!rc && ms < t) t = ms;
            }
        }
        const bool keep = t_tuned <= t_start;
        c->tune = keep ? tuned : start;
        best_ms = keep ? t_tuned : t_start;
    }
    release(snap);
    clear_graphs(c);
    if (rc) return rc;
    if (out) to_public(c, out, best_ms);
    return MHIT_OK;
}

}  // extern "C"
