/*
 * mhit36.h -- C-ABI of the MI355X-native MHIT36 single-GPU time-stepper.
 *
 * The reference (aroccon/MHIT36_mixedp) has no operator API: its whole
 * timestep is inline in `program main` (multi/main.f90:456-1042) and talks to
 * cuFFT, cuDecomp and MPI directly.  This header is the drop-in boundary that
 * replaces those call sites for one GPU.  A Fortran driver binds it through
 * iso_c_binding (driver/mhit_c.f90); Python binds it through ctypes
 * (mhit36_mixedp_amd/api.py).
 *
 * Conventions
 *  - Every call returns an int status, MHIT_OK (0) on success.  This mirrors
 *    CHECK_CUDECOMP_EXIT (multi/main.f90:1); unlike the reference, FFT/HIP
 *    errors are never ignored (cf. main.f90:190-205, 854-913).  The text of
 *    the last error is returned by mhit_last_error().
 *  - Fields are interior n^3 FP64 arrays in Fortran order (x fastest), i.e.
 *    C order [z][y][x], exactly the layout of the raw .dat files written by
 *    writefield (multi/readwrite.f90:39-92).  Host buffers are caller-owned;
 *    uploads and downloads copy.
 *  - One host thread per context, one HIP stream per context.  Calls are
 *    blocking at return, as the reference runs everything synchronously on
 *    the default stream.
 */
#ifndef MHIT36_H
#define MHIT36_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MHIT_OK              0
#define MHIT_ERR_ARG         1  /* bad argument / unsupported size          */
#define MHIT_ERR_HIP         2  /* HIP runtime error                        */
#define MHIT_ERR_CFL         3  /* cou > 7: the reference stops (main.f90:1006) */
#define MHIT_ERR_NOMEM       4  /* device allocation failed                 */
#define MHIT_ERR_STATE       5  /* call not valid in the current state      */
#define MHIT_ERR_NOGPU       6  /* no HIP device visible                    */
#define MHIT_ERR_IO          7  /* a file could not be written              */

/* Poisson precision: the "mixedp" branch solves in FP32 (reference README.md:7-8,
 * main.f90:283-284); the FP64 form is the same algorithm in double.      */
#define MHIT_POISSON_FP64    0
#define MHIT_POISSON_FP32    1

/* Wavenumber used by the spectral divide (main.f90:889).                  */
#define MHIT_K2_CONTINUOUS   0  /* kx^2+ky^2+kz^2, the reference           */
#define MHIT_K2_MODIFIED     1  /* sum (2 sin(k dx/2)/dx)^2, opt-in: exact discrete projection */

/* Tracer interpolation (particles.f90:78-122).                            */
#define MHIT_INTERP_REFERENCE 0 /* 1-D linear per component, the reference */
#define MHIT_INTERP_TRILINEAR 1 /* trilinear at the staggered positions, opt-in */

/* Field numbers, as writefield/readfield (multi/readwrite.f90:21-25).      */
#define MHIT_FIELD_U   1
#define MHIT_FIELD_V   2
#define MHIT_FIELD_W   3
#define MHIT_FIELD_P   4
#define MHIT_FIELD_PHI 5
/* AB2 history (previous-step right-hand sides), for a restart that
 * continues bit for bit.  No counterpart in the reference, which restarts
 * with an Euler step (multi/main.f90:440-441) and keeps these only in
 * module.f90:58-60, 69.                                                   */
#define MHIT_FIELD_RHSU_O   6
#define MHIT_FIELD_RHSV_O   7
#define MHIT_FIELD_RHSW_O   8
#define MHIT_FIELD_RHSPHI_O 9

/*
 * Run parameters.  The first block is input.inp, in its order
 * (multi/input.inp:1-18, multi/readinput.f90:14-35).  The second block is
 * what the reference fixes at compile time (module.f90:2 nx; main.f90:55-57
 * phiflag/partflag) or that this library adds as options.
 */
typedef struct mhit_params {
    /* input.inp */
    int32_t restart;      /* 0 fresh start, 1 restart                    */
    int32_t tstart;       /* first iteration                             */
    int32_t tfin;         /* last iteration                              */
    int32_t dump;         /* output frequency                            */
    int32_t inflow;       /* 0 Taylor-Green, 1 read ./input/{u,v,w}.dat  */
    int32_t inphi;        /* 0 centred drop, 1 read ./input/phi.dat      */
    double  dt;
    double  mu;           /* dynamic viscosity                           */
    double  rho;
    double  f1, f2, f3;   /* ABC forcing amplitudes A, B, C              */
    double  k0;           /* forced wavenumber                           */
    double  radius;       /* initial drop radius                         */
    double  sigma;        /* surface tension                             */
    double  epsr;         /* eps/dx                                      */
    int64_t npart;        /* number of tracers                           */
    int32_t inpart;       /* tracer initial condition                    */
    /* compile-time switches of the reference, runtime here               */
    int32_t nx;           /* grid points per direction (ny = nz = nx)    */
    int32_t phiflag;      /* 1: ACDI phase field + CSF                   */
    int32_t partflag;     /* 1: Lagrangian tracers                       */
    int32_t poisson_precision; /* MHIT_POISSON_FP32 (default) / _FP64    */
    int32_t k2_mode;      /* MHIT_K2_CONTINUOUS (default) / _MODIFIED    */
    int32_t interp_mode;  /* MHIT_INTERP_REFERENCE (default) / _TRILINEAR */
    int32_t cfl_abort;    /* 1 (default): stop when cou > 7              */
} mhit_params;

/* Per-step scalars the reference keeps in module `velocity`
 * (module.f90:59-61) plus diagnostics it never computed (added here).     */
typedef struct mhit_diag {
    int64_t step;         /* completed timesteps since create/upload     */
    double  alpha, beta;  /* AB2 coefficients for the next step          */
    double  gumax;        /* signed max over u,v,w (main.f90:998-1002)   */
    double  cou;          /* gumax*dt/dx (main.f90:1003)                 */
    double  umean, vmean, wmean; /* means removed by the last correction */
    double  ke;           /* 0.5 <u^2+v^2+w^2> over stored face values   */
    double  dissipation;  /* 2 nu <S_ij S_ij>, central differences       */
    double  maxdiv;       /* max |sum_d (u_d(+1)-u_d)/dx|                 */
    double  phi_mass;     /* sum phi dx^3                                */
    double  phi_min, phi_max;
    int64_t nan_count;    /* non-finite values seen in u,v,w,phi         */
} mhit_diag;

typedef struct mhit_ctx mhit_ctx;

/* Reference defaults: multi/input.inp values, nx = 256 (module.f90:2),
 * FP32 Poisson, continuous k^2, reference interpolation.                  */
int mhit_params_default(mhit_params* p);

/* Allocate device state and FFT plans for params p (main.f90:154-310).
 * Supported nx: 2^a 3^b with 8 <= nx <= 512.  AB2 state starts as the
 * reference's (alpha=1, beta=0, gumax=1; main.f90:440-444).               */
int mhit_create(const mhit_params* p, mhit_ctx** out);
int mhit_destroy(mhit_ctx* ctx);

/* Copy an interior n^3 FP64 field in/out (fieldn 1..9, see MHIT_FIELD_*).
 * Stand-in for readfield/readfield_restart/writefield
 * (multi/readwrite.f90:1-318).  Downloading U/V/W returns the velocity
 * after the mean removal of main.f90:977-981.  P is the pressure of the
 * last projection promoted to FP64.                                        */
int mhit_upload_field(mhit_ctx* ctx, int fieldn, const double* host);
int mhit_download_field(mhit_ctx* ctx, int fieldn, double* host);

/* Dump fields to <dir>/<name>_%08d.dat (name u, v, w, p, phi, rhsu_o,
 * rhsv_o, rhsw_o, rhsphi_o), the files of writefield
 * (multi/readwrite.f90:38-98): the interior n^3 array in FP64, Fortran
 * order, with the values mhit_download_field returns.  field_mask has bit
 * (1 << fieldn) set for every field wanted.  The fields are snapshot on
 * the device in order with the timesteps already enqueued; a host thread
 * then writes the files while the caller continues (the reference's time
 * loop waits for its writes).  A call first waits for the previous dump
 * and returns that dump's error, if any, without starting a new one.
 * mhit_dump_wait blocks until the pending dump is on disk and returns its
 * status (MHIT_ERR_IO when a file could not be written).                 */
int mhit_dump_async(mhit_ctx* ctx, const char* dir, int32_t step, uint32_t field_mask);
int mhit_dump_wait(mhit_ctx* ctx);

/* Tracers (particles.f90:11-22): positions in [0, 2pi).  Upload resets the
 * previous-step velocity to zero (the first step is Euler).  Download takes
 * NULL for any column not wanted; `tag` returns the particle tag
 * (part(:,Itag)) so that callers can match particles after reordering.    */
int mhit_upload_tracers(mhit_ctx* ctx, int64_t n, const double* x,
                        const double* y, const double* z);
/* Full tracer state (tag, position, previous-step velocity) as returned by
 * mhit_download_tracers, for a restart that continues bit for bit.  The
 * reference's part(:, Itag..) columns (particles.f90:11-22) are what it
 * would persist; it re-seeds tracers instead (main.f90:399-415).          */
int mhit_upload_tracer_state(mhit_ctx* ctx, int64_t n, const double* tag,
                             const double* x, const double* y, const double* z,
                             const double* up, const double* vp,
                             const double* wp);
int mhit_download_tracers(mhit_ctx* ctx, double* tag, double* x, double* y,
                          double* z, double* up, double* vp, double* wp);
int64_t mhit_tracer_count(const mhit_ctx* ctx);

/* Advance nsteps timesteps (the loop body main.f90:456-1007).  Returns
 * MHIT_ERR_CFL if cou > 7 after some step (when cfl_abort = 1); the state
 * is then the one after that step and further calls return MHIT_ERR_CFL. */
int mhit_step(mhit_ctx* ctx, int nsteps);

/* Scalars of the last step plus diagnostics computed on the current state. */
int mhit_diag_compute(mhit_ctx* ctx, mhit_diag* out);
/* The scalars of the last step only (step, alpha, beta, gumax, cou and the
 * removed means), read back without a pass over the fields: what the
 * reference prints every step (main.f90:1003-1005).  The diagnostic
 * members of *out are zero.                                               */
int mhit_scalars(mhit_ctx* ctx, mhit_diag* out);

/* Set the AB2 / gamma state, e.g. to resume a run (main.f90:440-444).     */
int mhit_set_ab2_state(mhit_ctx* ctx, double alpha, double beta, double gumax);

/* Launch shapes (no counterpart in the reference, whose start-up tunes
 * cuDecomp's process grid and backends instead, multi/main.f90:80-136).
 * None of these changes a result bit, except the FFT transform form
 * (fft_wave, and widths that move a pass to or from eight columns), which
 * changes rounding.                                                       */
typedef struct mhit_tuning {
    int32_t ns_mz;        /* z planes per workgroup, fused NS pass          */
    int32_t ac_mz;        /* z planes per workgroup, fused ACDI pass        */
    int32_t xcd;          /* 1: XCD-aware workgroup order, fused passes (the
                             default up to 256^3 and in ensembles)          */
    int32_t sort_every;   /* tracer re-ordering period in steps (0: never)  */
    int32_t ns_my;        /* tile rows of the fused NS pass: 16 or 8 (64
                             columns), 32 (32 columns) or 10 (96 columns)   */
    int32_t ac_my;        /* tile rows of the fused ACDI pass: the same     */
    int32_t fft_wave;     /* 1: column transforms of eight columns (the z
                             pass; the y pass in FP64) in the wavefront form,
                             lane/register transposes by permlane and DPP
                             (n = 64, 128, 256, 512); 0: the register-direct
                             form with LDS exchanges (the default with the
                             FP64 Poisson, where it measured faster).  The
                             two forms compute the same DFT in a different
                             operation order, so they differ in rounding
                             only.                                          */
    int32_t fft_ycols;    /* columns per workgroup of the y passes          */
    int32_t fft_zcols;    /* columns per workgroup of the z pass            */
    int32_t fft_xrows;    /* row pairs per workgroup of the x passes: the FFT
                             launch widths.  0 in mhit_set_tuning keeps the
                             default for the size and precision;
                             mhit_get_tuning reports the width in use.  Grids
                             of 64 and more have compiled alternatives: y 16
                             or 8 columns (FP32 Poisson), 8 or 4 (FP64); z 8
                             or 16 (FP32), 8 or 4 (FP64); x 8, 4 or 16 row
                             pairs (FP32), 4, 2 or 8 (FP64).  Eight columns
                             take the wavefront form when fft_wave is 1, so a
                             width that moves a pass to or from 8 columns
                             changes rounding; the other fields change no
                             result bit.                                    */
    int32_t fft_xy;       /* 1: the x and y transforms of a z plane in one
                             launch (the plane's half spectrum staged in LDS;
                             the spectrum crosses HBM once instead of twice)
                             where compiled: n = 32, 48, 64, 96, 128 with the
                             FP32 Poisson, n = 32, 48, 64, 96 with FP64 (the
                             default there; mhit_get_tuning reports 0
                             elsewhere).  The x launch widths then go unused.
                             Changes rounding only.                         */
    int32_t fft_xplan;    /* radix plan of the separate x passes: 0 radix 8
                             first (512 = 8.8.8, 256 = 8.4.8, 128 = 8.2.8,
                             the default), 1 radix 4 until 64 points remain
                             (512 = 4.4.4.8, 256 = 4.8.8, 128 = 4.4.8,
                             384 = 3.4.4.8).  Only n = 128, 256, 384, 512
                             have two plans (mhit_get_tuning reports 0
                             elsewhere).  Changes rounding only.            */
    double ms_per_step;   /* mhit_autotune: measured with these settings    */
} mhit_tuning;
int mhit_get_tuning(mhit_ctx* ctx, mhit_tuning* out);
int mhit_set_tuning(mhit_ctx* ctx, const mhit_tuning* t);
/* Time candidate shapes and transform forms on the context's own state
 * (`steps` timed steps per candidate), one setting at a time, keeping the
 * fastest; then time the chosen settings and the starting ones
 * alternately, twice each, and keep the faster set.  The state is
 * restored: the run then continues exactly as if the call had not been
 * made (apart from the rounding of the transform forms it picks).         */
int mhit_autotune(mhit_ctx* ctx, int steps, mhit_tuning* out);

/* Drop statistics from the phase field (post-processing of the reference,
 * run on the live state): cells with phi >= threshold, connected through
 * faces (connectivity 6) or faces, edges and corners (26), across the
 * periodic boundaries (post/unionfind_periodic/main.f90).  Drops are
 * numbered by their smallest cell index (z, y, x order), as the reference's
 * minimum labels; *ndrops returns the total, at most max_drops are written.
 * Centre: circular mean of the cell centres (i - 1/2) dx on each axis;
 * velocity: mean over the drop's cells; volume = cells dx^3;
 * d_eq = (6 volume / pi)^(1/3) (post/DropCountSize/calculate_deq.f90).     */
typedef struct mhit_drop {
    int64_t cells;        /* voxel count                                   */
    int64_t first_cell;   /* smallest linear cell index (z n + y) n + x    */
    double x, y, z;       /* centre in [0, 2 pi)                           */
    double u, v, w;       /* mean velocity                                 */
    double volume, d_eq;
} mhit_drop;
int mhit_drop_stats(mhit_ctx* ctx, double threshold, int connectivity,
                    int64_t max_drops, mhit_drop* out, int64_t* ndrops);

/* Kinetic-energy shell spectrum of the current velocity (the reference
 * computes spectra offline, python_utilities/plot_fields.py:116-151):
 * E[k] = 1/2 sum over modes with |m| in [k - 1/2, k + 1/2) of
 * (|u^|^2 + |v^|^2 + |w^|^2) / n^6, so that sum_k E[k] is mhit_diag's ke.
 * Uses the context's FFT precision.  *nshells returns the number of shells
 * (floor(sqrt(3) n/2 + 1/2) + 1); E gets min(nbins, *nshells) values,
 * zero-filled up to nbins.                                                */
int mhit_energy_spectrum(mhit_ctx* ctx, int32_t nbins, double* E, int32_t* nshells);

/* mhit_energy_spectrum without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The shell sums go to a fixed per-device array of
 * MHIT_SPECTRUM_NOALLOC_MAX_BINS doubles that the library holds until the
 * process ends; min(nbins, shells) above that limit returns MHIT_ERR_ARG.
 * Every call on a device shares its array: calls that overlap on one
 * device (from several threads or contexts) race.  Results as
 * mhit_energy_spectrum's (the shell sums are atomic adds, so the two agree
 * to rounding, not bit for bit).                                          */
#define MHIT_SPECTRUM_NOALLOC_MAX_BINS 4096
int mhit_energy_spectrum_noalloc(mhit_ctx* ctx, int32_t nbins, double* E, int32_t* nshells);

/* Block until all work queued on the context's stream is done.            */
int mhit_sync(mhit_ctx* ctx);

/* Text of the last error on this thread ("" if none).                    */
const char* mhit_last_error(void);

/* ---- initial conditions (computed on the device) -------------------- */
/* Taylor-Green velocity, main.f90:324-337 (in the time-stepper's frame).  */
int mhit_init_taylor_green(mhit_ctx* ctx);
/* One drop of radius p.radius at the box centre, main.f90:370-382.        */
int mhit_init_drop(mhit_ctx* ctx);
/* Seeded random-phase velocity, |A(k)| ~ k^-2 on 0 < |k| <= kmax, exactly
 * divergence-free for the staggered divergence, scaled to u_rms.          */
int mhit_init_random_velocity(mhit_ctx* ctx, uint64_t seed, double kmax,
                              double urms);
/* ndrops seeded non-overlapping drops of the given radius (tanh profile of
 * main.f90:378, periodic minimum-image distance, max over drops).         */
int mhit_init_drops(mhit_ctx* ctx, int32_t ndrops, double radius,
                    uint64_t seed);
/* Drop centres used by mhit_init_drops (host only; xyz interleaved).      */
int mhit_drop_centers(int32_t ndrops, double radius, double min_gap,
                      uint64_t seed, double* xyz);
/* n tracers uniform in the box from a seeded counter-based generator.     */
int mhit_init_tracers_uniform(mhit_ctx* ctx, int64_t n, uint64_t seed);

/* ---- building blocks, exposed for parity tests ---------------------- */
/* Solve lap(p) = rhs with the context's Poisson path: R2C FFT, divide by
 * -k^2 (zero mode 0), C2R (main.f90:853-912).  Host n^3 FP64 in and out;
 * in FP32 mode rhs is rounded to FP32 first, as rhspp (main.f90:844).     */
int mhit_poisson_solve(mhit_ctx* ctx, const double* rhs, double* p);
/* Forward unnormalised 3-D R2C FFT (e^{-i}), output (n/2+1) x n x n
 * complex as interleaved re,im in C order [kz][ky][kx].                   */
int mhit_fft_r2c(mhit_ctx* ctx, const double* in, double* out);
/* Inverse unnormalised 3-D C2R FFT (e^{+i}) of the same layout.           */
int mhit_fft_c2r(mhit_ctx* ctx, const double* in, double* out);

/* ---- ensembles of independent boxes ----------------------------------- */
/* nbox independent MHIT36 boxes (sizes, physics switches and parameters
 * may differ per box) advanced together on one GPU.  The reference runs one
 * box per job; this has no reference counterpart (it serves ensembles of
 * small boxes, e.g. 16 each of 64^3, 96^3 and 128^3).  Each box is a full
 * context: initialise, upload and
 * download through mhit_ensemble_box().  mhit_ensemble_step advances every
 * box nsteps steps as one captured graph; returns MHIT_ERR_CFL if any box
 * stopped on cou > 7 (that box stays frozen, the others keep stepping).
 * mhit_ensemble_set_batched: on (the default), boxes of one grid size share
 * every launch of a pass (one launch per pass and size); off, each box
 * runs its own launches on a graph branch of its own.  Either call resets
 * the boxes' launch shapes to the defaults for that form (batched: z
 * chunks of up to 128 planes, shortened until the batched launches hold
 * about four workgroups per CU; otherwise many workgroups per box; see
 * mhit_tuning).
 * Results are the same bits either way, and the same as each box stepped
 * alone.                                                                  */
typedef struct mhit_ensemble mhit_ensemble;
int mhit_ensemble_create(int32_t nbox, const mhit_params* params, mhit_ensemble** out);
int mhit_ensemble_box(mhit_ensemble* ens, int32_t i, mhit_ctx** box);
int mhit_ensemble_step(mhit_ensemble* ens, int nsteps);
int mhit_ensemble_set_batched(mhit_ensemble* ens, int32_t on);
/* Batched form: tune the launch shapes per grid size and Poisson precision
 * (z chunk of the fused passes, their tile rows and XCD order, the FFT
 * transform forms, the x passes' radix plan and the launch widths; the
 * same value for every box of the class), one setting at a time, each
 * candidate timed over `steps` ensemble steps on the live state; then the
 * chosen settings and the starting ones are timed alternately, twice
 * each, and the faster set is kept.  Every box's state is restored (as mhit_autotune does for one
 * box; the reference's analogue is cuDecomp's start-up autotune,
 * multi/main.f90:80-136).  *ms_per_step returns the ensemble step time
 * with the kept settings, which each box's mhit_get_tuning reports.  The
 * transform forms change rounding only; mhit_ensemble_set_batched resets
 * the shapes.                                                             */
int mhit_ensemble_autotune(mhit_ensemble* ens, int steps, double* ms_per_step);
int mhit_ensemble_destroy(mhit_ensemble* ens);

/* ---- per-stage timing ------------------------------------------------ */
/* With profiling on, mhit_step records a HIP event after every stage of
 * the step on the context's stream and accumulates the elapsed time per
 * stage (this adds one event per stage; leave it off for timing whole
 * steps).  bytes_per_call is the stage's algorithmic HBM traffic: every
 * array it reads or writes counted once.                                 */
#define MHIT_PROF_MAX 24
typedef struct mhit_profile {
    int32_t nstages;
    char    name[MHIT_PROF_MAX][32];
    int64_t calls[MHIT_PROF_MAX];
    double  ms[MHIT_PROF_MAX];              /* accumulated                */
    double  bytes_per_call[MHIT_PROF_MAX];
} mhit_profile;
int mhit_profile_enable(mhit_ctx* ctx, int on);   /* also resets counters */
int mhit_profile_read(mhit_ctx* ctx, mhit_profile* out);

/* Plain device-to-device copy of `bytes` bytes (a read and a write per
 * byte), repeated iters times on the context's stream; *gbps = achieved
 * (read + write) GB/s.  The measured copy ceiling of the GPU, and the
 * calibration kernel for HBM byte counters.  No reference counterpart.    */
int mhit_copy_probe(mhit_ctx* ctx, int64_t bytes, int iters, double* gbps);
/* Pointwise streaming over `reads` FP64 input and `writes` FP64 output
 * arrays of bytes_per_stream bytes each (1..8 of each), repeated iters
 * times; *gbps = achieved (read + write) GB/s.  The ceiling for a pass
 * with that mix of concurrent streams and no stencil.  No reference
 * counterpart.                                                            */
int mhit_stream_probe(mhit_ctx* ctx, int64_t bytes_per_stream, int reads, int writes, int iters,
                      double* gbps);

/* Library version string.                                                 */
const char* mhit_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MHIT36_H */
