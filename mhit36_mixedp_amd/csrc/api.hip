// api.hip -- the C-ABI of include/mhit36.h: context lifetime, transfers,
// the timestep driver and the initial-condition generators.
#include "fft.hpp"
#include "internal.hpp"

#include <algorithm>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace mhit {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

template <typename P>
static int dmalloc(P** p, size_t count) {
    *p = nullptr;
    if (count == 0) return MHIT_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(P));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(P)) + " bytes: " +
                  hipGetErrorString(e));
        *p = nullptr;
        return MHIT_ERR_NOMEM;
    }
    return MHIT_OK;
}

static void dfree(void* p) {
    if (p) (void)hipFree(p);
}

static int check_params(const mhit_params* p) {
    if (!p) { set_error("null params"); return MHIT_ERR_ARG; }
    if (!fft_size_supported(p->nx)) {
        set_error("nx=" + std::to_string(p->nx) + " unsupported (2^a 3^b, 8..512)");
        return MHIT_ERR_ARG;
    }
    if (!(p->dt > 0) || !(p->rho > 0) || !(p->epsr > 0) || p->mu < 0) {
        set_error("dt, rho, epsr must be > 0 and mu >= 0");
        return MHIT_ERR_ARG;
    }
    if (p->poisson_precision != MHIT_POISSON_FP32 && p->poisson_precision != MHIT_POISSON_FP64) {
        set_error("poisson_precision must be MHIT_POISSON_FP32 or MHIT_POISSON_FP64");
        return MHIT_ERR_ARG;
    }
    return MHIT_OK;
}

static int sync_scalars(mhit_ctx* c) {
    MHIT_HIP_TRY(hipMemcpyAsync(&c->S_host, c->S, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    return MHIT_OK;
}

static int push_scalars(mhit_ctx* c) {
    MHIT_HIP_TRY(hipMemcpyAsync(c->S, &c->S_host, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    return MHIT_OK;
}

// "key=value,key=value" with keys ns_mz, ac_mz, ns_my, ac_my, xcd, sort;
// unknown keys, non-positive chunk lengths and tile rows other than 8, 16,
// 32 or 10
// are ignored.
static void parse_tuning(const char* s, Tuning& t) {
    std::string str(s);
    size_t pos = 0;
    while (pos < str.size()) {
        size_t end = str.find(',', pos);
        if (end == std::string::npos) end = str.size();
        const std::string kv = str.substr(pos, end - pos);
        const size_t eq = kv.find('=');
        if (eq != std::string::npos) {
            const std::string key = kv.substr(0, eq);
            const int val = std::atoi(kv.c_str() + eq + 1);
            if (key == "ns_mz" && val > 0) t.ns_mz = val;
            else if (key == "ac_mz" && val > 0) t.ac_mz = val;
            else if (key == "ns_my" && (val == 8 || val == 16 || val == 32 || val == 10)) t.ns_my = val;
            else if (key == "ac_my" && (val == 8 || val == 16 || val == 32 || val == 10)) t.ac_my = val;
            else if (key == "xcd") t.xcd = val != 0;
            else if (key == "sort" && val >= 0) t.sort_every = val;
            else if (key == "fft") t.fft_wave = val != 0;
            // widths that are not compiled for the size fall back to the default
            else if (key == "ycols" && val > 0) t.fft_ycols = val;
            else if (key == "zcols" && val > 0) t.fft_zcols = val;
            else if (key == "xrows" && val > 0) t.fft_xrows = val;
            else if (key == "xy") t.fft_xy = val != 0;
            else if (key == "xplan") t.fft_xplan = val != 0;
        }
        pos = end + 1;
    }
}

// Launch shapes by grid size, then MHIT_TUNE; mhit_autotune refines.
// A box alone: 64x8 tiles up to 64^3, 64x16 above (less halo); z chunks
// of min(128, n/4) planes, halved (never below 4) until the NS pass has a
// workgroup per CU.  Measured, ms per step, FP32 Poisson, single phase /
// multiphase: 64^3 0.075 / 0.113 at 8 planes -> 0.066 / 0.091 at 4;
// 128^3 0.137 / 0.206 at 16 planes (64x8 ACDI tiles) -> 0.124 / 0.175 at
// 8; 256^3 0.700 / 1.021 at 32 -> 0.674 / 0.983 at 64; 512^3 keeps 128
// (fewer warm-up planes than 64: 8.17 -> 8.07 ms).  Boxes batched in an
// ensemble: 64x16 tiles and min(n, 128) planes, which the ensemble then
// shortens until its launches fill the GPU.
void default_tuning(mhit_ctx* c, bool batched, bool env) {
    const int n = c->n;
    // 96 x 10 tiles where 64-column tiles would leave part of each row of
    // tiles empty (96^3: a third of the threads); against 32 x 32 tiles they
    // read fewer partial lines per cell (0.108 against 0.137 128-byte lines
    // of each field): the 64^3 / 96^3 / 128^3 ensemble 18.96-19.05k ->
    // 19.13-19.19k box-steps/s, 16 x 96^3 alone 19.0k -> 19.6k
    const int rows = (n > 64 && n % 64 != 0 && n % 96 == 0) ? 10
                   : (n > 64 && n % 64 != 0 && n % 32 == 0) ? 32 : ((!batched && n <= 64) ? 8 : 16);
    int zc = std::min(128, n);
    if (!batched) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
            cus <= 0)
            cus = 256;
        zc = std::min(128, std::max(4, n / 4));
        const long long tiles = (long long)div_up(n, tile_cols(rows)) * div_up(n, rows);
        while (tiles * div_up(n, zc) < cus && zc / 2 >= 4) zc /= 2;
    }
    c->tune.ns_mz = zc;
    c->tune.ac_mz = zc;
    c->tune.ns_my = rows;
    c->tune.ac_my = rows;
    // XCD tile order: neighbouring tiles of a box then share an XCD's L2.
    // Batched (the 64^3 / 96^3 / 128^3 ensemble 18.78-18.84k -> 18.87-
    // 18.96k box-steps/s) and a box alone up to 256^3 (ms per step: 64^3
    // 0.049 -> 0.046, 128^3 multiphase FP64 0.190 -> 0.183, 256^3 0.764 ->
    // 0.751); at 512^3 it measured slower (7.56 -> 7.68 ms) and stays off
    c->tune.xcd = (batched || n <= 256) ? 1 : 0;
    // the wavefront form of the column transforms pays in FP32 (the z pass);
    // with the FP64 Poisson the register-direct form measured faster
    // (512^3: 8.84-8.89 -> 8.77 ms per step)
    c->tune.fft_wave = c->fp32 ? 1 : 0;
    // one launch for the x and y transforms where compiled: 64^3 FP64
    // 0.051 -> 0.049 ms per step, 96^3 FP64 0.099 -> 0.095; batched 16 x
    // 64^3 / 96^3 / 128^3 (FP32) 56.0k / 18.2k / 9.79k -> 58.2k / 18.7k /
    // 9.91k box-steps/s
    c->tune.fft_xy = fft_xy_supported(n, c->fp32) ? 1 : 0;
    if (env) tuning_env(c);
}

void tuning_env(mhit_ctx* c) {
    if (const char* t = std::getenv("MHIT_TUNE")) parse_tuning(t, c->tune);
}

double* state_field(mhit_ctx* c, int fieldn) {
    switch (fieldn) {
        case MHIT_FIELD_U: return c->u;
        case MHIT_FIELD_V: return c->v;
        case MHIT_FIELD_W: return c->w;
        case MHIT_FIELD_PHI: return c->phi;
        case MHIT_FIELD_RHSU_O: return c->rhsu_o;
        case MHIT_FIELD_RHSV_O: return c->rhsv_o;
        case MHIT_FIELD_RHSW_O: return c->rhsw_o;
        case MHIT_FIELD_RHSPHI_O: return c->rhsphi_o;
        default: return nullptr;
    }
}

const char* stage_name(int s) {
    static const char* names[ST_COUNT] = {
        "tracer_sort", "tracers", "acdi_psi", "acdi_rhs", "acdi_normals", "acdi_update", "acdi",
        "csf", "ns_ab2",
        "fft_x_fwd_div", "fft_y_fwd", "fft_z_solve", "fft_y_inv", "fft_x_inv", "correct",
        "finalize"};
    return (s >= 0 && s < ST_COUNT) ? names[s] : "?";
}

// Algorithmic HBM bytes of one call of stage s: every array the stage
// reads or writes counted once (stencil neighbours are free).
double stage_bytes(const mhit_ctx* c, int s) {
    const double N3 = (double)c->N3, n = c->n;
    const double cs = c->fp32 ? 8.0 : 16.0, rs = c->fp32 ? 4.0 : 8.0;
    const double spec = (n / 2) * n * n * cs + n * n * cs;   // half spectrum + Nyquist plane
    switch (s) {
        case ST_TRACER_SORT: return 120.0 * (double)c->np;   // keys, sort, 7 columns r/w
        case ST_TRACERS: return 144.0 * (double)c->np;       // pos r/w, u_old r/w, 6 gathers
        case ST_ACDI_PSI: return 24.0 * N3;                  // r phi; w psi, tanh
        case ST_ACDI_RHS: return 72.0 * N3;                  // r u,v,w,phi,psi; w rphi, n x3
        case ST_ACDI_NORM: return 48.0 * N3;                 // r/w n x3
        case ST_ACDI_UPDATE: return 72.0 * N3;               // r phi,rphi,rphi_o,T,n x3; w phi,rphi_o
        case ST_ACDI: return 56.0 * N3;                      // r phi,u,v,w,rphi_o; w phi',rphi_o
        case ST_CSF: return 32.0 * N3;                       // r phi; w f x3
        case ST_NS:   // r u,v,w,rhs_o[,phi or f x3]; w u*,rhs_o[, div u* (fused pass)]
            return ((c->P.phiflag ? (c->march ? 104.0 : 120.0) : 96.0) + (c->march ? rs : 0.0)) * N3;
        case ST_FFT_X_FWD:   // r div u* (fused NS) or u*,v*,w*; w spectrum
            return (c->march ? rs : 24.0) * N3 + spec;
        case ST_FFT_Y_FWD: case ST_FFT_Z: case ST_FFT_Y_INV: return 2.0 * spec;
        case ST_FFT_X_INV: return spec + rs * N3;            // r spectrum; w p
        case ST_CORRECT: return (48.0 + rs) * N3;            // r u*, p; w u
        case ST_FINALIZE: return 48.0 * c->npart_blocks;
        default: return 0.0;
    }
}

// the kernels of one timestep, main.f90:456-1007 (the tracer re-ordering
// is not part of it)
static int enqueue_step_kernels(mhit_ctx* c) {
    int rc;
    if (c->P.partflag && c->np > 0) {
        Range r("Particles");
        if ((rc = launch_tracers(c, c->stream))) return rc;                  // 469-510
    }
    if (c->P.phiflag) {
        Range r("Phase-field");
        if ((rc = c->march ? launch_acdi_march(c) : launch_acdi(c))) return rc; // 521-637
    }
    {
        Range r("Projection");
        if ((rc = c->march ? launch_ns_march(c) : launch_ns(c))) return rc;  // 658-796
    }
    {
        Range r("Poisson");
        if ((rc = poisson_project(c))) return rc;                            // 833-923
    }
    {
        Range r("Correction");
        if ((rc = launch_correct(c))) return rc;                             // 942-962
        if ((rc = launch_finalize(c))) return rc;                            // 964-1007
    }
    return MHIT_OK;
}

void clear_graphs(mhit_ctx* c) {
    c->graph_gen++;   // also invalidates the ensemble graphs that hold this box
    for (auto& gr : c->graphs) {
        if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
        gr = mhit_ctx::StepGraph{};
    }
}

// One timestep as a replay of a captured graph.  The graph is captured on
// first use for the current buffers (the fused ACDI pass alternates phi
// and phi_next; a tracer re-ordering swaps the tracer columns).  All
// per-step scalars live in device memory, so replays need no host update.
static int graph_step(mhit_ctx* c) {
    double* cur[mhit_ctx::kPingPong];
    pingpong_current(c, cur);
    mhit_ctx::StepGraph* slot = nullptr;
    for (auto& gr : c->graphs) {
        if (!gr.exec || gr.tag != c->ptag) continue;
        bool same = true;
        for (int p = 0; p < mhit_ctx::kPingPong; p++) same = same && gr.in[p] == cur[p];
        if (same) { slot = &gr; break; }
    }
    if (!slot) {
        for (auto& gr : c->graphs)
            if (!gr.exec) { slot = &gr; break; }
        if (!slot) {            // all four in use: start over
            clear_graphs(c);
            slot = &c->graphs[0];
        }
        double* tag = c->ptag;
        hipGraph_t graph = nullptr;
        MHIT_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step_kernels(c);    // performs the role swaps on the host
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        // the capture advanced the host-side buffer roles; undo it so the
        // launch below applies the step once
        for (int p = 0; p < mhit_ctx::kPingPong; p++) slot->in[p] = cur[p];
        pingpong_undo(c, slot->in, slot->out);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        MHIT_HIP_TRY(e);
        e = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MHIT_HIP_TRY(e);
        slot->tag = tag;
    }
    MHIT_HIP_TRY(hipGraphLaunch(slot->exec, c->stream));
    pingpong_apply(c, slot->in, slot->out);
    return MHIT_OK;
}

int sync_scalars_public(mhit_ctx* c) { return sync_scalars(c); }
int push_scalars_public(mhit_ctx* c) { return push_scalars(c); }

// The part of a timestep that stays outside a captured graph: the tracer
// re-ordering, every sort_every steps.
int step_eager_part(mhit_ctx* c) {
    int rc;
    if (c->P.partflag && c->np > 1 && c->tune.sort_every > 0 &&
        c->steps_since_sort >= c->tune.sort_every) {
        if ((rc = sort_tracers(c))) return rc;
        c->steps_since_sort = 0;
    }
    c->steps_since_sort++;
    return MHIT_OK;
}

// The kernels of one timestep on the context's stream (for capture).
int step_kernels(mhit_ctx* c) { return enqueue_step_kernels(c); }

// one timestep, main.f90:456-1007
static int enqueue_step(mhit_ctx* c) {
    int rc;
    Range r("Iteration");
    if (c->prof) {
        c->prof_n = 0;
        prof_mark(c, -1);
    }
    if ((rc = step_eager_part(c))) return rc;
    if (c->use_graph && !c->prof) return graph_step(c);
    if ((rc = enqueue_step_kernels(c))) return rc;
    if (c->prof) {
        MHIT_HIP_TRY(hipEventSynchronize(c->prof_ev[c->prof_n - 1]));
        for (int e = 1; e < c->prof_n; e++) {
            float ms = 0.0f;
            MHIT_HIP_TRY(hipEventElapsedTime(&ms, c->prof_ev[e - 1], c->prof_ev[e]));
            const int s = c->prof_stage[e];
            c->prof_ms[s] += ms;
            c->prof_calls[s] += 1;
        }
    }
    return MHIT_OK;
}

static int ensure_tracers(mhit_ctx* c, int64_t n) {
    c->steps_since_sort = c->tune.sort_every;   // re-order on the next step
    clear_graphs(c);                            // tracer count and columns change
    if (n <= c->np_cap) { c->np = n; return MHIT_OK; }
    double** cols[] = {&c->ptag, &c->px, &c->py, &c->pz, &c->pu1, &c->pv1, &c->pw1};
    for (double** col : cols) { dfree(*col); *col = nullptr; }
    int rc;
    for (double** col : cols)
        if ((rc = dmalloc(col, (size_t)n))) return rc;
    c->np_cap = n;
    c->np = n;
    return MHIT_OK;
}

// Enqueue nsteps timesteps on the context's stream without waiting.
int step_async(mhit_ctx* c, int nsteps) {
    int rc;
    if (c->S_host.halted) { set_error("stopped: cou > 7"); return MHIT_ERR_CFL; }
    for (int s = 0; s < nsteps; s++)
        if ((rc = enqueue_step(c))) return rc;
    return MHIT_OK;
}

int finish_synced(mhit_ctx* c);

// Wait for the queued steps and read the scalars back; MHIT_ERR_CFL if the
// CFL check stopped the run (main.f90:1006).
int finish_steps(mhit_ctx* c) {
    int rc;
    if ((rc = sync_scalars(c))) return rc;
    return finish_synced(c);
}

// finish_steps once S_host holds the device scalars
int finish_synced(mhit_ctx* c) {
    int rc;
    if (c->S_host.skipped) {
        // steps queued behind a CFL stop wrote no AB2 history but the host
        // still swapped the history buffers after each: an odd count leaves
        // the last history in the "new" buffer, so swap once more
        if (c->march && (c->S_host.skipped & 1)) ns_swap(c);
        c->S_host.skipped = 0;
        if ((rc = push_scalars(c))) return rc;
    }
    if (c->S_host.halted) {
        set_error("cou = " + std::to_string(c->S_host.cou) + " > 7 at step " +
                  std::to_string(c->S_host.step) + " (main.f90:1006)");
        return MHIT_ERR_CFL;
    }
    return MHIT_OK;
}

}  // namespace mhit

using namespace mhit;

extern "C" {

const char* mhit_last_error(void) { return g_last_error.c_str(); }

const char* mhit_version(void) { return "mhit36-mi355x 0.1 (gfx950)"; }

int mhit_params_default(mhit_params* p) {
    if (!p) return MHIT_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    // multi/input.inp:1-18
    p->restart = 0; p->tstart = 0; p->tfin = 10000; p->dump = 100000;
    p->inflow = 0; p->inphi = 0;
    p->dt = 0.0002; p->mu = 0.006; p->rho = 1.0;
    p->f1 = 1.0; p->f2 = 1.0; p->f3 = 1.0; p->k0 = 2.0;
    p->radius = 0.5; p->sigma = 1.0; p->epsr = 1.0;
    p->npart = 10000; p->inpart = 1;
    p->nx = 256;                     // module.f90:2
    p->phiflag = 0; p->partflag = 0; // main.f90:55-57
    p->poisson_precision = MHIT_POISSON_FP32;
    p->k2_mode = MHIT_K2_CONTINUOUS;
    p->interp_mode = MHIT_INTERP_REFERENCE;
    p->cfl_abort = 1;
    return MHIT_OK;
}

int mhit_create(const mhit_params* P, mhit_ctx** out) {
    if (!out) { set_error("null output"); return MHIT_ERR_ARG; }
    *out = nullptr;
    int rc = check_params(P);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return MHIT_ERR_NOGPU;
    }
    mhit_ctx* c = new mhit_ctx();
    c->P = *P;
    const int n = c->n = P->nx;
    c->N3 = (size_t)n * n * n;
    c->fp32 = P->poisson_precision == MHIT_POISSON_FP32;
    // MHIT_STENCIL_PATH=cell selects the one-cell-per-thread kernels that
    // follow the reference's loop split (a cross-check of the fused passes)
    if (const char* sp = std::getenv("MHIT_STENCIL_PATH")) c->march = std::strcmp(sp, "cell") != 0;
    // MHIT_GRAPH=0 launches the kernels of every step one by one
    if (const char* gr = std::getenv("MHIT_GRAPH")) c->use_graph = std::strcmp(gr, "0") != 0;
    // readinput.f90:39-49
    const double twopi = 8.0 * std::atan(1.0);
    Grid& g = c->g;
    g.n = n;
    g.lx = twopi;
    g.dx = g.lx / (double)n;
    g.dxi = 1.0 / g.dx;
    g.ddxi = 1.0 / g.dx / g.dx;
    g.rho = P->rho;
    g.rhoi = 1.0 / P->rho;
    g.mu = P->mu;
    g.dt = P->dt;
    g.eps = P->epsr * g.dx;
    g.epsi = 1.0 / g.eps;
    g.enm = (double)1.e-16f;   // `enum=1.e-16`: a default-real literal
    g.sigma = P->sigma;
    g.f1 = P->f1; g.f2 = P->f2; g.f3 = P->f3;

    auto fail = [&](int code) { mhit_destroy(c); return code; };
    if (hipGetDevice(&c->device) != hipSuccess) { set_error("hipGetDevice failed"); return fail(MHIT_ERR_HIP); }
    default_tuning(c, false);   // launch shapes for this device's CU count
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(MHIT_ERR_HIP);
    }
    const size_t N3 = c->N3;
    // state and AB2 history; the other half of the AB2 ping-pong (fused
    // path) or the intermediates of the split ACDI/CSF kernels (cell path)
    // only where they are used: at 512^3 multiphase, 17 instead of 27 GB
    std::vector<double**> fields = {&c->u, &c->v, &c->w, &c->us, &c->vs, &c->ws,
                                    &c->rhsu_o, &c->rhsv_o, &c->rhsw_o};
    if (c->march) fields.insert(fields.end(), {&c->rhsu_n, &c->rhsv_n, &c->rhsw_n});
    if (P->phiflag) {
        fields.insert(fields.end(), {&c->phi, &c->rhsphi_o, &c->phi_next});
        if (!c->march)
            fields.insert(fields.end(), {&c->rhsphi, &c->psi, &c->tanh_psi, &c->normx, &c->normy,
                                         &c->normz, &c->fxst, &c->fyst, &c->fzst});
    }
    for (double** f : fields)
        if ((rc = dmalloc(f, N3))) return fail(rc);
    const size_t H = n / 2;
    const size_t csz = c->fp32 ? sizeof(float2) : sizeof(double2);
    const size_t rsz = c->fp32 ? sizeof(float) : sizeof(double);
    if ((rc = dmalloc((char**)&c->pres, N3 * rsz))) return fail(rc);
    if ((rc = dmalloc((char**)&c->spec, (size_t)n * n * H * csz))) return fail(rc);
    if ((rc = dmalloc((char**)&c->nyq, (size_t)n * n * csz))) return fail(rc);
    if ((rc = dmalloc((char**)&c->tw, (size_t)n * csz))) return fail(rc);
    if ((rc = dmalloc(&c->kk, (size_t)n))) return fail(rc);
    if ((rc = dmalloc(&c->mysin, 3 * (size_t)(n + 1)))) return fail(rc);
    c->mycos = c->mysin + (n + 1);
    c->xext = c->mysin + 2 * (n + 1);
    c->npart_blocks = correct_block_count(n);
    if ((rc = dmalloc(&c->partials, (size_t)c->npart_blocks * 6))) return fail(rc);
    c->ndiag_blocks = diag_block_count(n);
    if ((rc = dmalloc(&c->dpartials, (size_t)c->ndiag_blocks * 8))) return fail(rc);
    if ((rc = dmalloc(&c->S, 1))) return fail(rc);

    // host tables: grid by cumulative sum and x_ext(n+1) = lx (main.f90:210-217),
    // wavenumbers (234-240), forcing tables (244-250) on the extended grid
    std::vector<double> tab(3 * (size_t)(n + 1));
    double* xe = tab.data() + 2 * (n + 1);
    xe[0] = 0.0;
    for (int i = 1; i < n; i++) xe[i] = xe[i - 1] + g.dx;
    xe[n] = g.lx;
    for (int t = 0; t <= n; t++) {
        tab[t] = std::sin(P->k0 * xe[t] + g.dx / 2);
        tab[(n + 1) + t] = std::cos(P->k0 * xe[t] + g.dx / 2);
    }
    std::vector<double> kx(n), kk(n);
    for (int i = 1; i <= n / 2; i++) kx[i - 1] = (double)(i - 1) * (twopi / g.lx);
    for (int i = n / 2 + 1; i <= n; i++) kx[i - 1] = (double)(i - 1 - n) * (twopi / g.lx);
    for (int i = 0; i < n; i++) {
        if (P->k2_mode == MHIT_K2_MODIFIED) {
            const double s = 2.0 * std::sin(kx[i] * g.dx * 0.5) * g.dxi;
            kk[i] = s * s;
        } else {
            kk[i] = kx[i] * kx[i];
        }
    }
    std::vector<double2> twd(n);
    std::vector<float2> twf(n);
    for (int j = 0; j < n; j++) {
        const double a = twopi * (double)j / (double)n;
        twd[j].x = std::cos(a); twd[j].y = -std::sin(a);
        twf[j].x = (float)twd[j].x; twf[j].y = (float)twd[j].y;
    }
    if (copy_on_stream(c, c->mysin, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        copy_on_stream(c, c->kk, kk.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        copy_on_stream(c, c->tw, c->fp32 ? (const void*)twf.data() : (const void*)twd.data(), n * csz,
                  hipMemcpyHostToDevice) != hipSuccess) {
        set_error("table upload failed");
        return fail(MHIT_ERR_HIP);
    }
    // zero state; AB2 start of main.f90:440-444
    for (double** f : fields)
        if (fill_on_stream(c, *f, 0, N3 * sizeof(double)) != hipSuccess) return fail(MHIT_ERR_HIP);
    if (fill_on_stream(c, c->pres, 0, N3 * rsz) != hipSuccess) return fail(MHIT_ERR_HIP);
    c->S_host = DevScalars{};
    c->S_host.alpha = 1.0;
    c->S_host.beta = 0.0;
    c->S_host.gumax = 1.0;
    if ((rc = push_scalars(c))) return fail(rc);
    *out = c;
    return MHIT_OK;
}

int mhit_destroy(mhit_ctx* c) {
    if (!c) return MHIT_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->u, c->v, c->w, c->us, c->vs, c->ws, c->rhsu_o, c->rhsv_o, c->rhsw_o,
                    c->rhsu_n, c->rhsv_n, c->rhsw_n,
                    c->phi, c->phi_next, c->rhsphi, c->rhsphi_o, c->psi, c->tanh_psi, c->normx, c->normy,
                    c->normz, c->fxst, c->fyst, c->fzst, c->pres, c->spec, c->nyq, c->tw, c->kk,
                    c->mysin, c->partials, c->dpartials, c->S, c->ptag, c->px, c->py, c->pz,
                    c->pu1, c->pv1, c->pw1};
    dump_release(c);   // waits for a dump still being written
    clear_graphs(c);
    for (void* p : ptrs) dfree(p);
    free_sort_buffers(c);
    for (hipEvent_t& e : c->prof_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MHIT_OK;
}

int mhit_upload_field(mhit_ctx* c, int fieldn, const double* host) {
    if (!c || !host) { set_error("null argument"); return MHIT_ERR_ARG; }
    int rc;
    const size_t N3 = c->N3;
    if (fieldn == MHIT_FIELD_P) {
        if (c->fp32) {
            std::vector<float> tmp(N3);
            for (size_t q = 0; q < N3; q++) tmp[q] = (float)host[q];
            MHIT_HIP_TRY(copy_on_stream(c, c->pres, tmp.data(), N3 * sizeof(float), hipMemcpyHostToDevice));
        } else {
            MHIT_HIP_TRY(copy_on_stream(c, c->pres, host, N3 * sizeof(double), hipMemcpyHostToDevice));
        }
        return MHIT_OK;
    }
    double* f = state_field(c, fieldn);
    if (!f) { set_error("field not allocated (phiflag=0?) or bad fieldn"); return MHIT_ERR_ARG; }
    if ((rc = sync_scalars(c))) return rc;
    MHIT_HIP_TRY(copy_on_stream(c, f, host, N3 * sizeof(double), hipMemcpyHostToDevice));
    if (fieldn <= MHIT_FIELD_W) {
        c->S_host.mean[fieldn - 1] = 0.0;
        if ((rc = push_scalars(c))) return rc;
    }
    return MHIT_OK;
}

int mhit_download_field(mhit_ctx* c, int fieldn, double* host) {
    if (!c || !host) { set_error("null argument"); return MHIT_ERR_ARG; }
    int rc;
    const size_t N3 = c->N3;
    if ((rc = sync_scalars(c))) return rc;
    if (fieldn == MHIT_FIELD_P) {
        if (c->fp32) {
            std::vector<float> tmp(N3);
            MHIT_HIP_TRY(copy_on_stream(c, tmp.data(), c->pres, N3 * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t q = 0; q < N3; q++) host[q] = (double)tmp[q];
        } else {
            MHIT_HIP_TRY(copy_on_stream(c, host, c->pres, N3 * sizeof(double), hipMemcpyDeviceToHost));
        }
        return MHIT_OK;
    }
    double* f = state_field(c, fieldn);
    if (!f) { set_error("field not allocated (phiflag=0?) or bad fieldn"); return MHIT_ERR_ARG; }
    MHIT_HIP_TRY(copy_on_stream(c, host, f, N3 * sizeof(double), hipMemcpyDeviceToHost));
    if (fieldn <= MHIT_FIELD_W) {
        // main.f90:977-981 applied on the way out (the device keeps u + mean)
        const double m = c->S_host.mean[fieldn - 1];
        if (m != 0.0)
            for (size_t q = 0; q < N3; q++) host[q] = host[q] - m;
    }
    return MHIT_OK;
}

int mhit_upload_tracers(mhit_ctx* c, int64_t n, const double* x, const double* y, const double* z) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !z))) { set_error("bad tracer arguments"); return MHIT_ERR_ARG; }
    int rc;
    if ((rc = ensure_tracers(c, n))) return rc;
    if (n == 0) return MHIT_OK;
    std::vector<double> tag(n);
    for (int64_t i = 0; i < n; i++) tag[i] = (double)(i + 1);   // particles.f90:46-49
    const size_t b = (size_t)n * sizeof(double);
    MHIT_HIP_TRY(copy_on_stream(c, c->ptag, tag.data(), b, hipMemcpyHostToDevice));
    MHIT_HIP_TRY(copy_on_stream(c, c->px, x, b, hipMemcpyHostToDevice));
    MHIT_HIP_TRY(copy_on_stream(c, c->py, y, b, hipMemcpyHostToDevice));
    MHIT_HIP_TRY(copy_on_stream(c, c->pz, z, b, hipMemcpyHostToDevice));
    MHIT_HIP_TRY(fill_on_stream(c, c->pu1, 0, b));
    MHIT_HIP_TRY(fill_on_stream(c, c->pv1, 0, b));
    MHIT_HIP_TRY(fill_on_stream(c, c->pw1, 0, b));
    return MHIT_OK;
}

int mhit_upload_tracer_state(mhit_ctx* c, int64_t n, const double* tag, const double* x,
                             const double* y, const double* z, const double* up, const double* vp,
                             const double* wp) {
    if (!c || n < 0 || (n > 0 && (!tag || !x || !y || !z || !up || !vp || !wp))) {
        set_error("bad tracer arguments");
        return MHIT_ERR_ARG;
    }
    int rc;
    if ((rc = ensure_tracers(c, n))) return rc;
    if (n == 0) return MHIT_OK;
    const size_t b = (size_t)n * sizeof(double);
    struct { double* d; const double* h; } cols[] = {{c->ptag, tag}, {c->px, x}, {c->py, y}, {c->pz, z},
                                                     {c->pu1, up}, {c->pv1, vp}, {c->pw1, wp}};
    for (auto& col : cols) MHIT_HIP_TRY(copy_on_stream(c, col.d, col.h, b, hipMemcpyHostToDevice));
    return MHIT_OK;
}

int mhit_download_tracers(mhit_ctx* c, double* tag, double* x, double* y, double* z, double* up,
                          double* vp, double* wp) {
    if (!c) { set_error("null context"); return MHIT_ERR_ARG; }
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    const size_t b = (size_t)c->np * sizeof(double);
    if (b == 0) return MHIT_OK;
    struct { double* h; double* d; } cols[] = {{tag, c->ptag}, {x, c->px}, {y, c->py}, {z, c->pz},
                                               {up, c->pu1}, {vp, c->pv1}, {wp, c->pw1}};
    for (auto& col : cols)
        if (col.h) MHIT_HIP_TRY(copy_on_stream(c, col.h, col.d, b, hipMemcpyDeviceToHost));
    return MHIT_OK;
}

int64_t mhit_tracer_count(const mhit_ctx* c) { return c ? c->np : -1; }

int mhit_step(mhit_ctx* c, int nsteps) {
    if (!c || nsteps < 0) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    int rc;
    if ((rc = step_async(c, nsteps))) return rc;
    return finish_steps(c);
}

int mhit_set_ab2_state(mhit_ctx* c, double alpha, double beta, double gumax) {
    if (!c) return MHIT_ERR_ARG;
    int rc;
    if ((rc = sync_scalars(c))) return rc;
    c->S_host.alpha = alpha;
    c->S_host.beta = beta;
    c->S_host.gumax = gumax;
    c->S_host.halted = 0;
    return push_scalars(c);
}

int mhit_sync(mhit_ctx* c) {
    if (!c) return MHIT_ERR_ARG;
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    return MHIT_OK;
}

int mhit_profile_enable(mhit_ctx* c, int on) {
    if (!c) return MHIT_ERR_ARG;
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    if (on && !c->prof_ev[0])
        for (hipEvent_t& e : c->prof_ev) MHIT_HIP_TRY(hipEventCreate(&e));
    c->prof = on != 0;
    for (int s = 0; s < MHIT_PROF_MAX; s++) { c->prof_ms[s] = 0.0; c->prof_calls[s] = 0; }
    return MHIT_OK;
}

int mhit_copy_probe(mhit_ctx* c, int64_t bytes, int iters, double* gbps) {
    if (!c || bytes < 16 || iters < 1 || !gbps) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    bytes -= bytes % 16;
    void *src = nullptr, *dst = nullptr;
    int rc;
    if ((rc = dmalloc((char**)&src, (size_t)bytes))) return rc;
    if ((rc = dmalloc((char**)&dst, (size_t)bytes))) { dfree(src); return rc; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = MHIT_ERR_HIP;
    if (hipMemsetAsync(src, 0, (size_t)bytes, c->stream) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
        launch_copy_probe(c, src, dst, (size_t)bytes) == MHIT_OK &&   // warm-up
        hipEventRecord(e0, c->stream) == hipSuccess) {
        bool ok = true;
        for (int t = 0; t < iters && ok; t++) ok = launch_copy_probe(c, src, dst, (size_t)bytes) == MHIT_OK;
        float ms = 0.0f;
        if (ok && hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
            *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
            rc = MHIT_OK;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c->stream);
    dfree(src);
    dfree(dst);
    if (rc) set_error("copy probe failed");
    return rc;
}

int mhit_stream_probe(mhit_ctx* c, int64_t bytes_per_stream, int reads, int writes, int iters, double* gbps) {
    if (!c || reads < 1 || reads > 8 || writes < 1 || writes > 8 || iters < 1 || !gbps ||
        bytes_per_stream < 8) {
        set_error("bad arguments");
        return MHIT_ERR_ARG;
    }
    const size_t n = (size_t)bytes_per_stream / sizeof(double);
    std::vector<double*> in(reads, nullptr), out(writes, nullptr);
    int rc = MHIT_OK;
    for (auto* v : {&in, &out})
        for (double*& p : *v)
            if (rc == MHIT_OK) rc = dmalloc(&p, n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == MHIT_OK) {
        rc = MHIT_ERR_HIP;
        bool ok = true;
        for (double* p : in) ok = ok && hipMemsetAsync(p, 0, n * sizeof(double), c->stream) == hipSuccess;
        if (ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
            launch_stream_probe(c, in.data(), reads, out.data(), writes, n) == MHIT_OK &&   // warm-up
            hipEventRecord(e0, c->stream) == hipSuccess) {
            for (int t = 0; t < iters && ok; t++)
                ok = launch_stream_probe(c, in.data(), reads, out.data(), writes, n) == MHIT_OK;
            float ms = 0.0f;
            if (ok && hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
                *gbps = (double)(reads + writes) * n * sizeof(double) * iters / (ms * 1e-3) / 1e9;
                rc = MHIT_OK;
            }
        }
        if (rc) set_error("stream probe failed");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c->stream);
    for (double* p : in) dfree(p);
    for (double* p : out) dfree(p);
    return rc;
}

int mhit_profile_read(mhit_ctx* c, mhit_profile* out) {
    if (!c || !out) return MHIT_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->nstages = ST_COUNT;
    for (int s = 0; s < ST_COUNT; s++) {
        std::strncpy(out->name[s], stage_name(s), 31);
        out->calls[s] = c->prof_calls[s];
        out->ms[s] = c->prof_ms[s];
        out->bytes_per_call[s] = stage_bytes(c, s);
    }
    return MHIT_OK;
}

static void fill_scalars(const mhit_ctx* c, mhit_diag* d) {
    const DevScalars& S = c->S_host;
    std::memset(d, 0, sizeof(*d));
    d->step = S.step;
    d->alpha = S.alpha; d->beta = S.beta; d->gumax = S.gumax; d->cou = S.cou;
    d->umean = S.mean[0]; d->vmean = S.mean[1]; d->wmean = S.mean[2];
}

int mhit_scalars(mhit_ctx* c, mhit_diag* d) {
    if (!c || !d) { set_error("null argument"); return MHIT_ERR_ARG; }
    int rc;
    if ((rc = sync_scalars(c))) return rc;
    fill_scalars(c, d);
    return MHIT_OK;
}

int mhit_diag_compute(mhit_ctx* c, mhit_diag* d) {
    if (!c || !d) { set_error("null argument"); return MHIT_ERR_ARG; }
    int rc;
    if ((rc = sync_scalars(c))) return rc;
    double r[8];
    if ((rc = launch_diag(c, r))) return rc;
    fill_scalars(c, d);
    const double N3 = (double)c->N3;
    d->ke = 0.5 * r[0] / N3;
    d->dissipation = 2.0 * (c->P.mu / c->P.rho) * r[1] / N3;
    d->maxdiv = r[2];
    d->phi_mass = c->P.phiflag ? r[3] * c->g.dx * c->g.dx * c->g.dx : 0.0;
    d->phi_min = c->P.phiflag ? r[5] : 0.0;
    d->phi_max = c->P.phiflag ? r[6] : 0.0;
    d->nan_count = (int64_t)r[7];
    return MHIT_OK;
}

// ------------------------------------------------------------ initial conditions
int mhit_init_taylor_green(mhit_ctx* c) {
    if (!c) return MHIT_ERR_ARG;
    int rc;
    if ((rc = launch_init_tgv(c))) return rc;
    if ((rc = sync_scalars(c))) return rc;
    c->S_host.mean[0] = c->S_host.mean[1] = c->S_host.mean[2] = 0.0;
    return push_scalars(c);
}

int mhit_init_drop(mhit_ctx* c) {
    if (!c) return MHIT_ERR_ARG;
    if (!c->P.phiflag) { set_error("phiflag = 0: no phase field"); return MHIT_ERR_STATE; }
    int rc;
    if ((rc = launch_init_drop(c))) return rc;
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    return MHIT_OK;
}

int mhit_init_random_velocity(mhit_ctx* c, uint64_t seed, double kmax, double urms) {
    if (!c || !(kmax >= 1.0)) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    const int n = c->n;
    const double kme = std::fmin(kmax, (double)(n / 2 - 1));
    const int K = (int)std::floor(kme);
    const double twopi = 8.0 * std::atan(1.0);
    std::vector<double> modes;
    double ms = 0.0;
    const uint64_t base = splitmix64(seed ^ 0x5EEDull);
    for (int mz = -K; mz <= K; mz++)
        for (int my = -K; my <= K; my++)
            for (int mx = -K; mx <= K; mx++) {
                const int m2 = mx * mx + my * my + mz * mz;
                if (m2 == 0 || (double)m2 > kme * kme) continue;
                if (!(mz > 0 || (mz == 0 && my > 0) || (mz == 0 && my == 0 && mx > 0))) continue;
                // six normal deviates from a counter keyed by the mode
                const uint64_t key = splitmix64(base + (uint64_t)((mx + 64) + 128 * ((my + 64) + 128 * (mz + 64))));
                double gsn[6];
                for (int t = 0; t < 3; t++) {
                    double a = u01(splitmix64(key + 2 * t)), b = u01(splitmix64(key + 2 * t + 1));
                    if (a < 1e-300) a = 1e-300;
                    const double r = std::sqrt(-2.0 * std::log(a));
                    gsn[2 * t] = r * std::cos(twopi * b);
                    gsn[2 * t + 1] = r * std::sin(twopi * b);
                }
                const double amp = 1.0 / (double)m2;   // |A| ~ |m|^-2
                double Ar[3], Ai[3], cr[3], ci[3];
                const int md[3] = {mx, my, mz};
                double cc = 0.0;
                for (int d = 0; d < 3; d++) {
                    Ar[d] = amp * gsn[2 * d];
                    Ai[d] = amp * gsn[2 * d + 1];
                    const double th = twopi * (double)md[d] / (double)n;
                    cr[d] = std::cos(th) - 1.0;
                    ci[d] = std::sin(th);
                    cc += cr[d] * cr[d] + ci[d] * ci[d];
                }
                // remove the component along conj(c): sum_d A_d c_d = 0 makes the
                // forward-difference divergence of the mode vanish
                double dr = 0.0, di = 0.0;
                for (int d = 0; d < 3; d++) {
                    dr += Ar[d] * cr[d] - Ai[d] * ci[d];
                    di += Ar[d] * ci[d] + Ai[d] * cr[d];
                }
                for (int d = 0; d < 3; d++) {
                    // (dr + i di) * conj(c_d) / cc
                    Ar[d] -= (dr * cr[d] + di * ci[d]) / cc;
                    Ai[d] -= (di * cr[d] - dr * ci[d]) / cc;
                }
                modes.push_back(mx); modes.push_back(my); modes.push_back(mz);
                for (int d = 0; d < 3; d++) {
                    modes.push_back(2.0 * Ar[d]);
                    modes.push_back(2.0 * Ai[d]);
                    ms += 2.0 * (Ar[d] * Ar[d] + Ai[d] * Ai[d]);
                }
            }
    const int nmodes = (int)(modes.size() / 9);
    if (nmodes == 0 || ms <= 0.0) { set_error("no modes below kmax"); return MHIT_ERR_ARG; }
    const double scale = urms / std::sqrt(ms / 3.0);
    for (int m = 0; m < nmodes; m++)
        for (int t = 3; t < 9; t++) modes[(size_t)m * 9 + t] *= scale;
    std::vector<double2> E(n);
    for (int t = 0; t < n; t++) {
        E[t].x = std::cos(twopi * t / n);
        E[t].y = std::sin(twopi * t / n);
    }
    double* dmodes = nullptr;
    double2* dE = nullptr;
    int rc;
    if ((rc = dmalloc(&dmodes, modes.size()))) return rc;
    if ((rc = dmalloc(&dE, (size_t)n))) { dfree(dmodes); return rc; }
    hipError_t e1 = copy_on_stream(c, dmodes, modes.data(), modes.size() * sizeof(double), hipMemcpyHostToDevice);
    hipError_t e2 = copy_on_stream(c, dE, E.data(), n * sizeof(double2), hipMemcpyHostToDevice);
    rc = (e1 == hipSuccess && e2 == hipSuccess) ? launch_init_modes(c, dmodes, nmodes, dE) : MHIT_ERR_HIP;
    if (rc == MHIT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = MHIT_ERR_HIP;
    dfree(dmodes);
    dfree(dE);
    if (rc) { set_error("random velocity init failed"); return rc; }
    if ((rc = sync_scalars(c))) return rc;
    c->S_host.mean[0] = c->S_host.mean[1] = c->S_host.mean[2] = 0.0;
    return push_scalars(c);
}

int mhit_drop_centers(int32_t ndrops, double radius, double min_gap, uint64_t seed, double* xyz) {
    if (ndrops < 0 || !(radius > 0) || (ndrops > 0 && !xyz)) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    const double L = 8.0 * std::atan(1.0);
    const double dmin = 2.0 * radius + min_gap;
    uint64_t ctr = splitmix64(seed ^ 0xD50Bull);
    for (int d = 0; d < ndrops; d++) {
        bool ok = false;
        for (int attempt = 0; attempt < 100000 && !ok; attempt++) {
            double c3[3];
            for (int t = 0; t < 3; t++) c3[t] = u01(splitmix64(ctr++)) * L;
            ok = true;
            for (int e = 0; e < d && ok; e++) {
                double r2 = 0.0;
                for (int t = 0; t < 3; t++) {
                    double a = std::fabs(c3[t] - xyz[3 * e + t]);
                    a = std::fmin(a, L - a);
                    r2 += a * a;
                }
                if (r2 < dmin * dmin) ok = false;
            }
            if (ok)
                for (int t = 0; t < 3; t++) xyz[3 * d + t] = c3[t];
        }
        if (!ok) { set_error("could not place " + std::to_string(ndrops) + " drops without overlap"); return MHIT_ERR_ARG; }
    }
    return MHIT_OK;
}

int mhit_init_drops(mhit_ctx* c, int32_t ndrops, double radius, uint64_t seed) {
    if (!c) return MHIT_ERR_ARG;
    if (!c->P.phiflag) { set_error("phiflag = 0: no phase field"); return MHIT_ERR_STATE; }
    std::vector<double> xyz(3 * (size_t)(ndrops > 0 ? ndrops : 1));
    int rc;
    if ((rc = mhit_drop_centers(ndrops, radius, 4.0 * c->g.eps, seed, xyz.data()))) return rc;
    double* dc = nullptr;
    if ((rc = dmalloc(&dc, xyz.size()))) return rc;
    rc = copy_on_stream(c, dc, xyz.data(), xyz.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess
             ? launch_init_drops(c, dc, ndrops, radius) : MHIT_ERR_HIP;
    if (rc == MHIT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = MHIT_ERR_HIP;
    dfree(dc);
    return rc;
}

int mhit_init_tracers_uniform(mhit_ctx* c, int64_t n, uint64_t seed) {
    if (!c || n < 0) return MHIT_ERR_ARG;
    int rc;
    if ((rc = ensure_tracers(c, n))) return rc;
    if (n == 0) return MHIT_OK;
    if ((rc = launch_init_tracers(c, n, seed))) return rc;
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    return MHIT_OK;
}

// ------------------------------------------------------------ building blocks
int mhit_poisson_solve(mhit_ctx* c, const double* rhs, double* p) {
    if (!c || !rhs || !p) { set_error("null argument"); return MHIT_ERR_ARG; }
    double* d = nullptr;
    int rc;
    if ((rc = dmalloc(&d, c->N3))) return rc;
    if (copy_on_stream(c, d, rhs, c->N3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = MHIT_ERR_HIP;
    if (!rc) rc = poisson_solve_field(c, d);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = MHIT_ERR_HIP;
    dfree(d);
    if (rc) return rc;
    return mhit_download_field(c, MHIT_FIELD_P, p);
}

}  // extern "C"

namespace mhit {

template <typename V>
static int download_spectrum(mhit_ctx* c, double* out) {
    const int n = c->n, H = n / 2;
    std::vector<V> spec((size_t)n * n * H), nyq((size_t)n * n);
    MHIT_HIP_TRY(copy_on_stream(c, spec.data(), c->spec, spec.size() * sizeof(V), hipMemcpyDeviceToHost));
    MHIT_HIP_TRY(copy_on_stream(c, nyq.data(), c->nyq, nyq.size() * sizeof(V), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < (size_t)n * n; r++) {
        for (int k = 0; k < H; k++) {
            out[2 * (r * (H + 1) + k)] = (double)spec[r * H + k].x;
            out[2 * (r * (H + 1) + k) + 1] = (double)spec[r * H + k].y;
        }
        out[2 * (r * (H + 1) + H)] = (double)nyq[r].x;
        out[2 * (r * (H + 1) + H) + 1] = (double)nyq[r].y;
    }
    return MHIT_OK;
}

template <typename V>
static int upload_spectrum(mhit_ctx* c, const double* in) {
    const int n = c->n, H = n / 2;
    std::vector<V> spec((size_t)n * n * H), nyq((size_t)n * n);
    for (size_t r = 0; r < (size_t)n * n; r++) {
        for (int k = 0; k < H; k++) {
            spec[r * H + k].x = in[2 * (r * (H + 1) + k)];
            spec[r * H + k].y = in[2 * (r * (H + 1) + k) + 1];
        }
        nyq[r].x = in[2 * (r * (H + 1) + H)];
        nyq[r].y = in[2 * (r * (H + 1) + H) + 1];
    }
    MHIT_HIP_TRY(copy_on_stream(c, c->spec, spec.data(), spec.size() * sizeof(V), hipMemcpyHostToDevice));
    MHIT_HIP_TRY(copy_on_stream(c, c->nyq, nyq.data(), nyq.size() * sizeof(V), hipMemcpyHostToDevice));
    return MHIT_OK;
}

}  // namespace mhit

extern "C" {

int mhit_fft_r2c(mhit_ctx* c, const double* in, double* out) {
    if (!c || !in || !out) { set_error("null argument"); return MHIT_ERR_ARG; }
    double* d = nullptr;
    int rc;
    if ((rc = dmalloc(&d, c->N3))) return rc;
    if (copy_on_stream(c, d, in, c->N3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = MHIT_ERR_HIP;
    if (!rc) rc = fft_r2c_dev(c, d);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = MHIT_ERR_HIP;
    dfree(d);
    if (rc) return rc;
    return c->fp32 ? download_spectrum<float2>(c, out) : download_spectrum<double2>(c, out);
}

int mhit_fft_c2r(mhit_ctx* c, const double* in, double* out) {
    if (!c || !in || !out) { set_error("null argument"); return MHIT_ERR_ARG; }
    int rc = c->fp32 ? upload_spectrum<float2>(c, in) : upload_spectrum<double2>(c, in);
    if (rc) return rc;
    if ((rc = fft_c2r_dev(c))) return rc;
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    return mhit_download_field(c, MHIT_FIELD_P, out);
}

}  // extern "C"
