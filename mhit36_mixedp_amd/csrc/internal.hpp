// internal.hpp -- context layout and launcher declarations shared by the
// translation units of libmhit36.so.
#pragma once

#include "common.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstring>
#include <utility>
#include <vector>

struct mhit_dump_state;   // dump.hip

struct mhit_ctx {
    mhit_params P;
    int n = 0;
    size_t N3 = 0;
    mhit::Grid g{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool fp32 = true;

    // velocity state (stored with the deferred mean, see DevScalars::mean)
    double *u = nullptr, *v = nullptr, *w = nullptr;
    // provisional velocity u* (main.f90:783-796)
    double *us = nullptr, *vs = nullptr, *ws = nullptr;
    double *rhsu_o = nullptr, *rhsv_o = nullptr, *rhsw_o = nullptr;
    // the other half of the AB2 history: the fused NS pass reads its
    // neighbours' previous rhs, so it writes the new one here and swaps
    double *rhsu_n = nullptr, *rhsv_n = nullptr, *rhsw_n = nullptr;
    // phase field (module.f90:66-71)
    double *phi = nullptr, *rhsphi = nullptr, *rhsphi_o = nullptr;
    double *phi_next = nullptr;   // ping-pong target of the fused ACDI pass
    bool march = true;            // fused z-marching stencil passes (default)
    mhit::Tuning tune{};          // launch shapes (MHIT_TUNE)
    long long steps_since_sort = 0;
    double *psi = nullptr, *tanh_psi = nullptr;
    double *normx = nullptr, *normy = nullptr, *normz = nullptr;
    double *fxst = nullptr, *fyst = nullptr, *fzst = nullptr;
    // Poisson: pressure (FP32 or FP64), half spectrum [z][y][kx < n/2] and
    // the kx = n/2 (Nyquist) plane [z][y], complex FP32 or FP64
    void *pres = nullptr, *spec = nullptr, *nyq = nullptr;
    void *tw = nullptr;           // W_n^j, j < n, complex in the Poisson precision
    double *kk = nullptr;         // per-axis k^2 by wavenumber index (n)
    // forcing tables and extended grid, one allocation of 3(n+1):
    // mysin[0..n], mycos[0..n], xext[0..n] (main.f90:210-250)
    double *mysin = nullptr, *mycos = nullptr, *xext = nullptr;
    // reductions
    double *partials = nullptr;   // correction partial sums/maxima
    int npart_blocks = 0;
    double *dpartials = nullptr;  // diagnostics partials
    int ndiag_blocks = 0;
    mhit::DevScalars *S = nullptr;
    // tracers (particles.f90:11-22, SoA)
    int64_t np = 0, np_cap = 0;
    double *ptag = nullptr, *px = nullptr, *py = nullptr, *pz = nullptr;
    double *pu = nullptr, *pv = nullptr, *pw = nullptr;
    double *pu1 = nullptr, *pv1 = nullptr, *pw1 = nullptr;
    // cell-order sort of the tracers: keys, permutation, the other half of
    // the double-buffered columns and the radix-sort workspace
    unsigned* skey[2] = {nullptr, nullptr};
    int* sidx[2] = {nullptr, nullptr};
    double* alt[7] = {};
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int64_t sort_cap = 0;
    // host copy of scalars after the last sync
    mhit::DevScalars S_host{};
    // captured timesteps, keyed by the buffers they use (phi ping-pong
    // parity, tracer columns after a re-ordering)
    // buffer pairs whose roles alternate from step to step
    static constexpr int kPingPong = 4;   // phi, rhsu_o, rhsv_o, rhsw_o
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        double* in[kPingPong] = {};    // current buffer of each pair before the step
        double* out[kPingPong] = {};   // ... and after it
        double* tag = nullptr;         // first tracer column
    };
    bool use_graph = true;
    StepGraph graphs[4];
    unsigned long long graph_gen = 0;   // bumped whenever captured graphs become invalid
    // per-stage event profiling (mhit_profile_enable)
    bool prof = false;
    hipEvent_t prof_ev[40] = {};
    int prof_stage[40] = {};
    int prof_n = 0;
    double prof_ms[MHIT_PROF_MAX] = {};
    long long prof_calls[MHIT_PROF_MAX] = {};
    mhit_dump_state* dump = nullptr;   // asynchronous field dumps (mhit_dump_async)
};

namespace mhit {

// Stages of one timestep, in launch order.
enum Stage {
    ST_TRACER_SORT, ST_TRACERS, ST_ACDI_PSI, ST_ACDI_RHS, ST_ACDI_NORM, ST_ACDI_UPDATE, ST_ACDI,
    ST_CSF, ST_NS,
    ST_FFT_X_FWD, ST_FFT_Y_FWD, ST_FFT_Z, ST_FFT_Y_INV, ST_FFT_X_INV, ST_CORRECT, ST_FINALIZE,
    ST_COUNT
};
const char* stage_name(int s);
double stage_bytes(const mhit_ctx* c, int s);

// The buffer pairs whose roles a timestep swaps on the host (the fused
// ACDI pass's phi ping-pong, the fused NS pass's AB2-history ping-pong).
// A captured step replays with the roles it was captured with, so graphs
// are keyed by the current buffer of each pair.
inline void pingpong_pairs(mhit_ctx* c, double** a[mhit_ctx::kPingPong],
                           double** b[mhit_ctx::kPingPong]) {
    a[0] = &c->phi;    b[0] = &c->phi_next;
    a[1] = &c->rhsu_o; b[1] = &c->rhsu_n;
    a[2] = &c->rhsv_o; b[2] = &c->rhsv_n;
    a[3] = &c->rhsw_o; b[3] = &c->rhsw_n;
}
// After a capture (which performed the swaps on the host): record the
// roles after the step in out[], and put the roles back to in[].
inline void pingpong_undo(mhit_ctx* c, double* const in[], double* out[]) {
    double** a[mhit_ctx::kPingPong];
    double** b[mhit_ctx::kPingPong];
    pingpong_pairs(c, a, b);
    for (int p = 0; p < mhit_ctx::kPingPong; p++) {
        out[p] = *a[p];
        *a[p] = in[p];
        if (*b[p] == in[p]) *b[p] = out[p];
    }
}
// After a replay: apply the swaps the captured step performs.
inline void pingpong_apply(mhit_ctx* c, double* const in[], double* const out[]) {
    double** a[mhit_ctx::kPingPong];
    double** b[mhit_ctx::kPingPong];
    pingpong_pairs(c, a, b);
    for (int p = 0; p < mhit_ctx::kPingPong; p++)
        if (out[p] != in[p]) std::swap(*a[p], *b[p]);
}
inline void pingpong_current(mhit_ctx* c, double* in[]) {
    double** a[mhit_ctx::kPingPong];
    double** b[mhit_ctx::kPingPong];
    pingpong_pairs(c, a, b);
    for (int p = 0; p < mhit_ctx::kPingPong; p++) in[p] = *a[p];
}

// Argument tables of batched launches (several boxes of one size in one
// launch).  The device block is allocated before a step is captured, so
// the tables' device addresses are known to the captured launches; the
// tables are filled on the host during the capture and copied to the
// device (upload) before the first replay.
struct ArgArena {
    char* dev = nullptr;
    size_t cap = 0;
    std::vector<unsigned char> host;
    template <typename A>
    const A* put(const A* t, int n) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        const size_t bytes = sizeof(A) * (size_t)n;
        if (!dev || off + bytes > cap) return nullptr;
        host.resize(off + bytes);
        std::memcpy(host.data() + off, t, bytes);
        return reinterpret_cast<const A*>(dev + off);
    }
};

// A blocking copy in order with the context's stream.  The stream is
// non-blocking, so a plain hipMemcpy (legacy stream) would not wait for the
// kernels queued on it, and an upload from pageable memory may return
// before its DMA has landed; here both are ordered and complete at return.
inline hipError_t copy_on_stream(mhit_ctx* c, void* dst, const void* src, size_t bytes,
                                 hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}
inline hipError_t fill_on_stream(mhit_ctx* c, void* dst, int value, size_t bytes) {
    hipError_t e = hipMemsetAsync(dst, value, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

// roctx ranges around the host-side enqueue of a timestep and of its
// stages, named as the reference's NVTX ranges (commented out there:
// "Iteration" main.f90:460, "Phase-field" 517, "Projection" 655, "Poisson"
// 831, "Correction" 940; "Particles" has no counterpart).  Under rocprofv3
// --marker-trace they group the launches; with no tool attached a push/pop
// pair costs tens of nanoseconds.  A captured step replays inside its
// "Iteration" range only (its stages were enqueued once, at capture).
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

// Record the end of stage s on the context stream (profiling only).
inline void prof_mark(mhit_ctx* c, int s) {
    if (!c->prof || c->prof_n >= 40) return;
    (void)hipEventRecord(c->prof_ev[c->prof_n], c->stream);
    c->prof_stage[c->prof_n] = s;
    c->prof_n++;
}

// api.hip: the device array of field id fieldn (1..9 but p), or nullptr
double* state_field(mhit_ctx* c, int fieldn);
// api.hip: launch shapes by grid size (for a box alone, or batched in an
// ensemble), then the MHIT_TUNE overrides
void default_tuning(mhit_ctx* c, bool batched, bool env = true);
void tuning_env(mhit_ctx* c);   // the MHIT_TUNE overrides alone
// dump.hip: wait for a pending dump and free its buffers
void dump_release(mhit_ctx* c);

// Supported grid sizes (FFT lengths).
bool fft_size_supported(int n);
// poisson.hip: whether the x and y transforms of a z plane run as one launch
// for this size and precision (Tuning::fft_xy)
bool fft_xy_supported(int n, bool fp32);
// poisson.hip: whether the x passes have a second radix plan for this size
// (Tuning::fft_xplan: n = 128, 256, 384, 512)
bool fft_xplan_differs(int n);
// Launch widths compiled for an FFT pass (0: y passes, columns per
// workgroup; 1: z pass, columns; 2: x passes, row pairs) of size n and
// precision: out[0] is the default, the count is returned (at most 3).
int fft_widths(int n, bool fp32, int pass, int out[3]);
// The width a context's pass uses: its Tuning value if compiled, else the
// default.
int fft_width(const mhit_ctx* c, int pass);

// tune.hip: a device copy of everything a timestep reads (restored into
// the current owners of the buffer roles) and the scalars
struct StateCopy;
int state_copy_take(mhit_ctx* c, StateCopy** out);
int state_copy_restore(mhit_ctx* c, const StateCopy* s);
void state_copy_free(StateCopy* s);

// poisson.hip
enum class XSrc { Divergence, Real };
int poisson_project(mhit_ctx* c);                       // u*,v*,w* -> pres
// the projection of the fused path (div u* in pres -> p in pres) for
// boxes[0..nb) of one size and precision, batched
int poisson_project_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st);
int poisson_solve_field(mhit_ctx* c, const double* rhs_dev);  // rhs -> pres
int fft_r2c_dev(mhit_ctx* c, const double* in_dev);     // -> spec, nyq
int fft_c2r_dev(mhit_ctx* c);                           // spec, nyq -> pres

// stencil.hip
int launch_acdi(mhit_ctx* c);
int launch_ns(mhit_ctx* c);

int launch_correct(mhit_ctx* c);
int launch_finalize(mhit_ctx* c);
int launch_correct_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st);
int launch_finalize_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st);
int launch_diag(mhit_ctx* c, double* host_out /* 8 values */);
int correct_block_count(int n);
int diag_block_count(int n);
int launch_copy_probe(mhit_ctx* c, const void* src, void* dst, size_t bytes);
int launch_stream_probe(mhit_ctx* c, const double* const* in, int R, double* const* out, int W, size_t n);

// march.hip
int launch_acdi_march(mhit_ctx* c);
// Swap the roles of the two AB2 history buffers of the NS pass.
void ns_swap(mhit_ctx* c);
int launch_ns_march(mhit_ctx* c);
// Batched forms: boxes[0..nb) share the size, flags and launch shapes; one
// launch on stream st, argument table in the arena; the role swaps of every
// box are done on the host as in the single-box form.
int launch_acdi_march_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st);
int launch_ns_march_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st);

// tracers.hip
int launch_tracers(mhit_ctx* c, hipStream_t st);
int sort_tracers(mhit_ctx* c);          // stable cell-order permutation
void free_sort_buffers(mhit_ctx* c);

// init.hip
int launch_init_tgv(mhit_ctx* c);
int launch_init_drop(mhit_ctx* c);
int launch_init_modes(mhit_ctx* c, const double* modes_dev, int nmodes, const double2* E);
int launch_init_drops(mhit_ctx* c, const double* centers_dev, int ndrops, double radius);
int launch_init_tracers(mhit_ctx* c, int64_t n, uint64_t seed);

}  // namespace mhit
