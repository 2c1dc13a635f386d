/*
 * mhit_oracle.c -- CPU FP64 restatement of one MHIT36 timestep.
 *
 * TEST INFRASTRUCTURE (the oracle), not product code: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every loop follows the reference loop it cites (aroccon/MHIT36_mixedp,
 * paths relative to the reference root), at one rank, with the halo layers
 * of the X-pencil replaced by periodic index wrap in y and z.  Expressions
 * keep the reference's operand order and parenthesisation, and the file is
 * compiled without FP contraction, so results are reproducible to the bit
 * for a given thread count (loops are per-cell; reductions are sequential).
 *
 * Frames (multi/main.f90:326-331, 709-714): x(i) = (i-1) dx, built by
 * cumulative addition (main.f90:211-214); local y/z index j maps to x(jg)
 * with jg = j (the halo-shifted frame), so interior 0-based index jj has
 * coordinate x_ext(jj+2) = (jj+1) dx.  The one out-of-range read at
 * jj = n-1 (x(nz+1), mysin(nz+1)) takes the extended grid value
 * x_ext(n+1) = lx (main.f90:216-217).
 */
#include "mhit_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cd;
typedef struct { float re, im; } cf;

struct oracle_ctx {
    mhit_params P;
    int n;
    size_t N3;
    /* readinput.f90:39-49 */
    double lx, dx, dxi, ddxi, rhoi, eps, epsi, enm;
    /* main.f90:210-250 */
    double *x_ext;          /* n+1 */
    double *kx;             /* n   */
    double *mysin, *mycos;  /* n+1: index t <-> Fortran mysin(t+1) */
    /* module.f90:51-71 */
    double *u, *v, *w, *rhsu, *rhsv, *rhsw, *rhsu_o, *rhsv_o, *rhsw_o;
    double *rhsp, *p;
    double *phi, *rhsphi, *rhsphi_o, *psidi, *tanh_psi;
    double *normx, *normy, *normz, *fxst, *fyst, *fzst;
    /* particles.f90:11-22 (SoA columns) */
    int64_t np;
    double *ptag, *px, *py, *pz, *pu, *pv, *pw, *pu1, *pv1, *pw1;
    /* main.f90:440-444, 998-1003 */
    double alpha, beta, gumax, cou, umean, vmean, wmean;
    int64_t step;
    int halted;
};

/* ------------------------------------------------------------------ */
/* indexing: interior n^3, C order [z][y][x], periodic wrap           */
static inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
#define IDX(c, i, j, k) ((((size_t)(k)) * (c)->n + (size_t)(j)) * (c)->n + (size_t)(i))

static double* dalloc(size_t n) {
    double* a = (double*)calloc(n ? n : 1, sizeof(double));
    if (!a) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return a;
}

int oracle_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* 1-D FFT: recursive decimation in time over the smallest prime factor,
 * naive DFT butterflies, twiddles exp(sign 2 pi i j / N) from cos/sin in
 * double.  Textbook Cooley-Tukey; pinned against numpy.fft and a
 * brute-force DFT in tests/test_oracle_fft.py.                        */
static int smallest_factor(int n) {
    for (int p = 2; p * p <= n; p++) if (n % p == 0) return p;
    return n;
}

#define MAXP 64

static void fft_rec_d(const cd* x, cd* y, int n, int s, const cd* tw, int N) {
    if (n == 1) { y[0] = x[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; r++) fft_rec_d(x + (size_t)r * s, y + (size_t)r * m, m, s * p, tw, N);
    int step_n = N / n, step_p = N / p;
    cd t[MAXP];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            cd a = y[r * m + k];
            cd w = tw[((size_t)r * k * step_n) % N];
            t[r].re = a.re * w.re - a.im * w.im;
            t[r].im = a.re * w.im + a.im * w.re;
        }
        for (int q = 0; q < p; q++) {
            double re = 0.0, im = 0.0;
            for (int r = 0; r < p; r++) {
                cd w = tw[(((size_t)r * q) % p) * step_p];
                re += t[r].re * w.re - t[r].im * w.im;
                im += t[r].re * w.im + t[r].im * w.re;
            }
            y[k + m * q].re = re;
            y[k + m * q].im = im;
        }
    }
}

static void fft_rec_f(const cf* x, cf* y, int n, int s, const cf* tw, int N) {
    if (n == 1) { y[0] = x[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; r++) fft_rec_f(x + (size_t)r * s, y + (size_t)r * m, m, s * p, tw, N);
    int step_n = N / n, step_p = N / p;
    cf t[MAXP];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            cf a = y[r * m + k];
            cf w = tw[((size_t)r * k * step_n) % N];
            t[r].re = a.re * w.re - a.im * w.im;
            t[r].im = a.re * w.im + a.im * w.re;
        }
        for (int q = 0; q < p; q++) {
            float re = 0.0f, im = 0.0f;
            for (int r = 0; r < p; r++) {
                cf w = tw[(((size_t)r * q) % p) * step_p];
                re += t[r].re * w.re - t[r].im * w.im;
                im += t[r].re * w.im + t[r].im * w.re;
            }
            y[k + m * q].re = re;
            y[k + m * q].im = im;
        }
    }
}

static cd* twiddles_d(int N, int sign) {
    cd* tw = (cd*)malloc(sizeof(cd) * N);
    const double twopi = 8.0 * atan(1.0);
    for (int j = 0; j < N; j++) {
        double a = twopi * (double)j / (double)N;
        tw[j].re = cos(a);
        tw[j].im = sign * sin(a);
    }
    return tw;
}

static cf* twiddles_f(int N, int sign) {
    cf* tw = (cf*)malloc(sizeof(cf) * N);
    const double twopi = 8.0 * atan(1.0);
    for (int j = 0; j < N; j++) {
        double a = twopi * (double)j / (double)N;
        tw[j].re = (float)cos(a);
        tw[j].im = (float)(sign * sin(a));
    }
    return tw;
}

void oracle_fft1d(int n, const double* in, double* out, int sign) {
    cd* tw = twiddles_d(n, sign);
    cd* tmp = (cd*)malloc(sizeof(cd) * n);
    fft_rec_d((const cd*)in, tmp, n, 1, tw, n);
    memcpy(out, tmp, sizeof(cd) * n);
    free(tmp); free(tw);
}

void oracle_fft1d_f32(int n, const float* in, float* out, int sign) {
    cf* tw = twiddles_f(n, sign);
    cf* tmp = (cf*)malloc(sizeof(cf) * n);
    fft_rec_f((const cf*)in, tmp, n, 1, tw, n);
    memcpy(out, tmp, sizeof(cf) * n);
    free(tmp); free(tw);
}

/* 3-D transform of an n^3 complex array [z][y][x]: lines along x, y, z. */
static void fft3d_d(int n, cd* a, int sign) {
    cd* tw = twiddles_d(n, sign);
    size_t n2 = (size_t)n * n;
    for (int axis = 0; axis < 3; axis++) {
        size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)n : n2);
        #pragma omp parallel
        {
            cd* lin = (cd*)malloc(sizeof(cd) * n);
            cd* lout = (cd*)malloc(sizeof(cd) * n);
            #pragma omp for schedule(static)
            for (long line = 0; line < (long)n2; line++) {
                size_t base;
                if (axis == 0) base = (size_t)line * n;
                else if (axis == 1) base = (size_t)(line / n) * n2 + (size_t)(line % n);
                else base = (size_t)line;
                for (int t = 0; t < n; t++) lin[t] = a[base + t * stride];
                fft_rec_d(lin, lout, n, 1, tw, n);
                for (int t = 0; t < n; t++) a[base + t * stride] = lout[t];
            }
            free(lin); free(lout);
        }
    }
    free(tw);
}

static void fft3d_f(int n, cf* a, int sign) {
    cf* tw = twiddles_f(n, sign);
    size_t n2 = (size_t)n * n;
    for (int axis = 0; axis < 3; axis++) {
        size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)n : n2);
        #pragma omp parallel
        {
            cf* lin = (cf*)malloc(sizeof(cf) * n);
            cf* lout = (cf*)malloc(sizeof(cf) * n);
            #pragma omp for schedule(static)
            for (long line = 0; line < (long)n2; line++) {
                size_t base;
                if (axis == 0) base = (size_t)line * n;
                else if (axis == 1) base = (size_t)(line / n) * n2 + (size_t)(line % n);
                else base = (size_t)line;
                for (int t = 0; t < n; t++) lin[t] = a[base + t * stride];
                fft_rec_f(lin, lout, n, 1, tw, n);
                for (int t = 0; t < n; t++) a[base + t * stride] = lout[t];
            }
            free(lin); free(lout);
        }
    }
    free(tw);
}

void oracle_fft3d(int n, const double* in, double* out, int sign) {
    size_t N3 = (size_t)n * n * n;
    cd* a = (cd*)malloc(sizeof(cd) * N3);
    memcpy(a, in, sizeof(cd) * N3);
    fft3d_d(n, a, sign);
    memcpy(out, a, sizeof(cd) * N3);
    free(a);
}

/* ------------------------------------------------------------------ */
/* Poisson solve, main.f90:853-912: forward transform, p^ = -r^/k2/N with
 * k2 from the wavenumber table kx (main.f90:234-240, 889-890), zero mode
 * pinned to 0 (895), inverse transform.  The half-spectrum R2C/C2R of the
 * reference is done here as a full complex transform of the real input
 * (identical in exact arithmetic).  FP32 form: rhspp = real(rhsp,4)
 * (844); transforms in FP32; the divide in FP64 stored back in FP32
 * (complex(4) phi3d, real(8) k2); p = real(pp,8) (F1 of the survey: the
 * correction is meant to read the FP32 pressure).                        */
static double k2_of(const oracle_ctx* c, int i, int j, int k) {
    if (c->P.k2_mode == MHIT_K2_MODIFIED) {
        double sx = 2.0 * sin(c->kx[i] * c->dx * 0.5) * c->dxi;
        double sy = 2.0 * sin(c->kx[j] * c->dx * 0.5) * c->dxi;
        double sz = 2.0 * sin(c->kx[k] * c->dx * 0.5) * c->dxi;
        return sx * sx + sy * sy + sz * sz;
    }
    return c->kx[i] * c->kx[i] + c->kx[j] * c->kx[j] + c->kx[k] * c->kx[k];
}

static void poisson(oracle_ctx* c, const double* rhs, double* p) {
    int n = c->n;
    size_t N3 = c->N3;
    double ntot = (double)((int64_t)n * (int64_t)n * (int64_t)n);
    if (c->P.poisson_precision == MHIT_POISSON_FP32) {
        cf* a = (cf*)malloc(sizeof(cf) * N3);
        for (size_t q = 0; q < N3; q++) { a[q].re = (float)rhs[q]; a[q].im = 0.0f; }
        fft3d_f(n, a, -1);
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t q = IDX(c, i, j, k);
                    double k2 = k2_of(c, i, j, k);
                    double re = -(double)a[q].re, im = -(double)a[q].im;
                    a[q].re = (float)(re / k2 / ntot);
                    a[q].im = (float)(im / k2 / ntot);
                }
        a[0].re = 0.0f; a[0].im = 0.0f;
        fft3d_f(n, a, +1);
        for (size_t q = 0; q < N3; q++) p[q] = (double)a[q].re;
        free(a);
    } else {
        cd* a = (cd*)malloc(sizeof(cd) * N3);
        for (size_t q = 0; q < N3; q++) { a[q].re = rhs[q]; a[q].im = 0.0; }
        fft3d_d(n, a, -1);
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t q = IDX(c, i, j, k);
                    double k2 = k2_of(c, i, j, k);
                    a[q].re = -a[q].re / k2 / ntot;
                    a[q].im = -a[q].im / k2 / ntot;
                }
        a[0].re = 0.0; a[0].im = 0.0;
        fft3d_d(n, a, +1);
        for (size_t q = 0; q < N3; q++) p[q] = a[q].re;
        free(a);
    }
}

int oracle_poisson(oracle_ctx* c, const double* rhs, double* p) {
    poisson(c, rhs, p);
    return MHIT_OK;
}

/* ------------------------------------------------------------------ */
oracle_ctx* oracle_create(const mhit_params* P) {
    if (!P || P->nx < 4) return NULL;
    oracle_ctx* c = (oracle_ctx*)calloc(1, sizeof(oracle_ctx));
    c->P = *P;
    int n = c->n = P->nx;
    c->N3 = (size_t)n * n * n;
    /* readinput.f90:39-49 */
    double twopi = 8.0 * atan(1.0);
    c->lx = twopi;
    c->dx = c->lx / (double)n;
    c->dxi = 1.0 / c->dx;
    c->ddxi = 1.0 / c->dx / c->dx;
    c->rhoi = 1.0 / P->rho;
    c->eps = P->epsr * c->dx;
    c->epsi = 1.0 / c->eps;
    c->enm = (double)1.e-16f;   /* `enum=1.e-16` is a default-real literal */
    /* main.f90:210-217: cumulative grid, extended by lx */
    c->x_ext = dalloc(n + 1);
    c->x_ext[0] = 0.0;
    for (int i = 1; i < n; i++) c->x_ext[i] = c->x_ext[i - 1] + c->dx;
    c->x_ext[n] = c->lx;
    /* main.f90:234-240 */
    c->kx = dalloc(n);
    for (int i = 1; i <= n / 2; i++) c->kx[i - 1] = (double)(i - 1) * (twopi / c->lx);
    for (int i = n / 2 + 1; i <= n; i++) c->kx[i - 1] = (double)(i - 1 - n) * (twopi / c->lx);
    /* main.f90:244-250 (extended to n+1 points, see header) */
    c->mysin = dalloc(n + 1);
    c->mycos = dalloc(n + 1);
    for (int t = 0; t <= n; t++) {
        c->mysin[t] = sin(P->k0 * c->x_ext[t] + c->dx / 2);
        c->mycos[t] = cos(P->k0 * c->x_ext[t] + c->dx / 2);
    }
    size_t N3 = c->N3;
    c->u = dalloc(N3); c->v = dalloc(N3); c->w = dalloc(N3);
    c->rhsu = dalloc(N3); c->rhsv = dalloc(N3); c->rhsw = dalloc(N3);
    c->rhsu_o = dalloc(N3); c->rhsv_o = dalloc(N3); c->rhsw_o = dalloc(N3);
    c->rhsp = dalloc(N3); c->p = dalloc(N3);
    if (P->phiflag) {
        c->phi = dalloc(N3); c->rhsphi = dalloc(N3); c->rhsphi_o = dalloc(N3);
        c->psidi = dalloc(N3); c->tanh_psi = dalloc(N3);
        c->normx = dalloc(N3); c->normy = dalloc(N3); c->normz = dalloc(N3);
        c->fxst = dalloc(N3); c->fyst = dalloc(N3); c->fzst = dalloc(N3);
    }
    /* main.f90:440-444 */
    c->alpha = 1.0; c->beta = 0.0; c->gumax = 1.0;
    return c;
}

void oracle_destroy(oracle_ctx* c) {
    if (!c) return;
    double* arrs[] = {c->x_ext, c->kx, c->mysin, c->mycos, c->u, c->v, c->w, c->rhsu, c->rhsv,
                      c->rhsw, c->rhsu_o, c->rhsv_o, c->rhsw_o, c->rhsp, c->p, c->phi, c->rhsphi,
                      c->rhsphi_o, c->psidi, c->tanh_psi, c->normx, c->normy, c->normz, c->fxst,
                      c->fyst, c->fzst, c->ptag, c->px, c->py, c->pz, c->pu, c->pv, c->pw, c->pu1,
                      c->pv1, c->pw1};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); q++) free(arrs[q]);
    free(c);
}

static double* field_ptr(oracle_ctx* c, int fieldn) {
    switch (fieldn) {
        case MHIT_FIELD_U: return c->u;
        case MHIT_FIELD_V: return c->v;
        case MHIT_FIELD_W: return c->w;
        case MHIT_FIELD_P: return c->p;
        case MHIT_FIELD_PHI: return c->phi;
        default: return NULL;
    }
}

int oracle_set_field(oracle_ctx* c, int fieldn, const double* in) {
    double* f = field_ptr(c, fieldn);
    if (!f) return MHIT_ERR_ARG;
    memcpy(f, in, sizeof(double) * c->N3);
    return MHIT_OK;
}

int oracle_get_field(oracle_ctx* c, int fieldn, double* out) {
    double* f = field_ptr(c, fieldn);
    if (!f) return MHIT_ERR_ARG;
    memcpy(out, f, sizeof(double) * c->N3);
    return MHIT_OK;
}

int oracle_set_ab2_state(oracle_ctx* c, double alpha, double beta, double gumax) {
    c->alpha = alpha; c->beta = beta; c->gumax = gumax;
    return MHIT_OK;
}

int oracle_set_tracers(oracle_ctx* c, int64_t n, const double* x, const double* y,
                       const double* z) {
    double** cols[] = {&c->ptag, &c->px, &c->py, &c->pz, &c->pu, &c->pv, &c->pw,
                       &c->pu1, &c->pv1, &c->pw1};
    for (int q = 0; q < 10; q++) { free(*cols[q]); *cols[q] = dalloc((size_t)n); }
    c->np = n;
    for (int64_t i = 0; i < n; i++) {
        c->ptag[i] = (double)(i + 1);   /* particles.f90:46-49 at rank 0 */
        c->px[i] = x[i]; c->py[i] = y[i]; c->pz[i] = z[i];
    }
    return MHIT_OK;
}

int oracle_get_tracers(oracle_ctx* c, double* x, double* y, double* z, double* up, double* vp,
                       double* wp) {
    size_t b = sizeof(double) * (size_t)c->np;
    if (x) memcpy(x, c->px, b);
    if (y) memcpy(y, c->py, b);
    if (z) memcpy(z, c->pz, b);
    if (up) memcpy(up, c->pu, b);
    if (vp) memcpy(vp, c->pv, b);
    if (wp) memcpy(wp, c->pw, b);
    return MHIT_OK;
}

/* ------------------------------------------------------------------ */
/* NS right-hand side, main.f90:658-717.  parts selects the convective
 * (672-694), viscous (696-707) and ABC forcing (709-714) terms.       */
static void ns_rhs(oracle_ctx* c, int parts, double* ru, double* rv, double* rw) {
    const int n = c->n;
    const double* u = c->u; const double* v = c->v; const double* w = c->w;
    const double dxi = c->dxi, ddxi = c->ddxi, mu = c->P.mu, rhoi = c->rhoi;
    const double f1 = c->P.f1, f2 = c->P.f2, f3 = c->P.f3;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
        for (int j = 0; j < n; j++) {
            int jp = wrapi(j + 1, n), jm = wrapi(j - 1, n);
            for (int i = 0; i < n; i++) {
                int ip = wrapi(i + 1, n), im = wrapi(i - 1, n);
#define U(a, b, d) u[IDX(c, a, b, d)]
#define V(a, b, d) v[IDX(c, a, b, d)]
#define W(a, b, d) w[IDX(c, a, b, d)]
                double h11, h12, h13, h21, h22, h23, h31, h32, h33;
                double r_u = 0.0, r_v = 0.0, r_w = 0.0;
                if (parts & ORACLE_NS_CONV) {
                    h11 = (U(ip,j,k)+U(i,j,k))*(U(ip,j,k)+U(i,j,k))     - (U(i,j,k)+U(im,j,k))*(U(i,j,k)+U(im,j,k));
                    h12 = (U(i,jp,k)+U(i,j,k))*(V(i,jp,k)+V(im,jp,k))   - (U(i,j,k)+U(i,jm,k))*(V(i,j,k)+V(im,j,k));
                    h13 = (U(i,j,kp)+U(i,j,k))*(W(i,j,kp)+W(im,j,kp))   - (U(i,j,k)+U(i,j,km))*(W(i,j,k)+W(im,j,k));
                    h21 = (U(ip,j,k)+U(ip,jm,k))*(V(ip,j,k)+V(i,j,k))   - (U(i,j,k)+U(i,jm,k))*(V(i,j,k)+V(im,j,k));
                    h22 = (V(i,jp,k)+V(i,j,k))*(V(i,jp,k)+V(i,j,k))     - (V(i,j,k)+V(i,jm,k))*(V(i,j,k)+V(i,jm,k));
                    h23 = (W(i,j,kp)+W(i,jm,kp))*(V(i,j,kp)+V(i,j,k))   - (W(i,j,k)+W(i,jm,k))*(V(i,j,k)+V(i,j,km));
                    h31 = (W(ip,j,k)+W(i,j,k))*(U(ip,j,k)+U(ip,j,km))   - (W(i,j,k)+W(im,j,k))*(U(i,j,k)+U(i,j,km));
                    h32 = (V(i,jp,k)+V(i,jp,km))*(W(i,jp,k)+W(i,j,k))   - (V(i,j,k)+V(i,j,km))*(W(i,j,k)+W(i,jm,k));
                    h33 = (W(i,j,kp)+W(i,j,k))*(W(i,j,kp)+W(i,j,k))     - (W(i,j,k)+W(i,j,km))*(W(i,j,k)+W(i,j,km));
                    h11 = h11*0.25*dxi; h12 = h12*0.25*dxi; h13 = h13*0.25*dxi;
                    h21 = h21*0.25*dxi; h22 = h22*0.25*dxi; h23 = h23*0.25*dxi;
                    h31 = h31*0.25*dxi; h32 = h32*0.25*dxi; h33 = h33*0.25*dxi;
                    r_u = -(h11+h12+h13);
                    r_v = -(h21+h22+h23);
                    r_w = -(h31+h32+h33);
                }
                if (parts & ORACLE_NS_VISC) {
                    h11 = mu*(U(ip,j,k)-2.0*U(i,j,k)+U(im,j,k))*ddxi;
                    h12 = mu*(U(i,jp,k)-2.0*U(i,j,k)+U(i,jm,k))*ddxi;
                    h13 = mu*(U(i,j,kp)-2.0*U(i,j,k)+U(i,j,km))*ddxi;
                    h21 = mu*(V(ip,j,k)-2.0*V(i,j,k)+V(im,j,k))*ddxi;
                    h22 = mu*(V(i,jp,k)-2.0*V(i,j,k)+V(i,jm,k))*ddxi;
                    h23 = mu*(V(i,j,kp)-2.0*V(i,j,k)+V(i,j,km))*ddxi;
                    h31 = mu*(W(ip,j,k)-2.0*W(i,j,k)+W(im,j,k))*ddxi;
                    h32 = mu*(W(i,jp,k)-2.0*W(i,j,k)+W(i,jm,k))*ddxi;
                    h33 = mu*(W(i,j,kp)-2.0*W(i,j,k)+W(i,j,km))*ddxi;
                    r_u = r_u+(h11+h12+h13)*rhoi;
                    r_v = r_v+(h21+h22+h23)*rhoi;
                    r_w = r_w+(h31+h32+h33)*rhoi;
                }
                if (parts & ORACLE_NS_FORCE) {
                    /* kg, jg of main.f90:709-710 -> table index kk+1, jj+1 */
                    r_u = r_u + f3*c->mysin[k + 1]+f2*c->mycos[j + 1];
                    r_v = r_v + f1*c->mysin[i]+f3*c->mycos[k + 1];
                    r_w = r_w + f2*c->mysin[j + 1]+f1*c->mycos[i];
                }
#undef U
#undef V
#undef W
                size_t q = IDX(c, i, j, k);
                ru[q] = r_u; rv[q] = r_v; rw[q] = r_w;
            }
        }
    }
}

int oracle_ns_rhs(oracle_ctx* c, int parts, double* ru, double* rv, double* rw) {
    ns_rhs(c, parts, ru, rv, rw);
    return MHIT_OK;
}

/* ------------------------------------------------------------------ */
/* ACDI, main.f90:522-630.                                             */
static void acdi_psi(oracle_ctx* c) {
    /* main.f90:522-534, all points */
    const double eps = c->eps, epsi = c->epsi, enm = c->enm;
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)c->N3; q++) {
        double val = fmin(c->phi[q], 1.0);
        c->psidi[q] = eps*log((val+enm)/(1.0-val+enm));
        c->tanh_psi[q] = tanh(0.5*c->psidi[q]*epsi);
    }
}

/* main.f90:537-564 (conv + diff into rphi; raw normals) */
static void acdi_rhs_conv_diff(oracle_ctx* c, int parts, double gamma, double* rphi) {
    const int n = c->n;
    const double* u = c->u; const double* v = c->v; const double* w = c->w;
    const double* phi = c->phi; const double* psi = c->psidi;
    const double dxi = c->dxi, ddxi = c->ddxi, eps = c->eps;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
        for (int j = 0; j < n; j++) {
            int jp = wrapi(j + 1, n), jm = wrapi(j - 1, n);
            for (int i = 0; i < n; i++) {
                int ip = wrapi(i + 1, n), im = wrapi(i - 1, n);
#define F(A, a, b, d) A[IDX(c, a, b, d)]
                double conv = 0.0, diff = 0.0;
                if (parts & ORACLE_PHI_CONV) {
                    conv = - (F(u,ip,j,k)*0.5*(F(phi,ip,j,k)+F(phi,i,j,k)) - F(u,i,j,k)*0.5*(F(phi,i,j,k)+F(phi,im,j,k)))*dxi
                           - (F(v,i,jp,k)*0.5*(F(phi,i,jp,k)+F(phi,i,j,k)) - F(v,i,j,k)*0.5*(F(phi,i,j,k)+F(phi,i,jm,k)))*dxi
                           - (F(w,i,j,kp)*0.5*(F(phi,i,j,kp)+F(phi,i,j,k)) - F(w,i,j,k)*0.5*(F(phi,i,j,k)+F(phi,i,j,km)))*dxi;
                }
                if (parts & ORACLE_PHI_DIFF) {
                    diff = gamma*(eps*(F(phi,ip,j,k)-2.0*F(phi,i,j,k)+F(phi,im,j,k))*ddxi +
                                  eps*(F(phi,i,jp,k)-2.0*F(phi,i,j,k)+F(phi,i,jm,k))*ddxi +
                                  eps*(F(phi,i,j,kp)-2.0*F(phi,i,j,k)+F(phi,i,j,km))*ddxi);
                }
                size_t q = IDX(c, i, j, k);
                /* the reference writes one expression: -(..)-(..)-(..) + gamma*(..) */
                if ((parts & ORACLE_PHI_CONV) && (parts & ORACLE_PHI_DIFF)) rphi[q] = conv + diff;
                else if (parts & ORACLE_PHI_CONV) rphi[q] = conv;
                else rphi[q] = diff;
                c->normx[q] = (F(psi,ip,j,k) - F(psi,im,j,k));
                c->normy[q] = (F(psi,i,jp,k) - F(psi,i,jm,k));
                c->normz[q] = (F(psi,i,j,kp) - F(psi,i,j,km));
#undef F
            }
        }
    }
}

/* main.f90:582-593 */
static void acdi_normalize(oracle_ctx* c) {
    const double enm = c->enm;
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)c->N3; q++) {
        double normag = 1.0/(sqrt(c->normx[q]*c->normx[q] + c->normy[q]*c->normy[q] + c->normz[q]*c->normz[q]) + enm);
        c->normx[q] = c->normx[q]*normag;
        c->normy[q] = c->normy[q]*normag;
        c->normz[q] = c->normz[q]*normag;
    }
}

/* main.f90:596-618 */
static void acdi_sharpen(oracle_ctx* c, double gamma, double* rphi) {
    const int n = c->n;
    const double dxi = c->dxi;
    const double* T = c->tanh_psi;
    const double* nx = c->normx; const double* ny = c->normy; const double* nz = c->normz;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
        for (int j = 0; j < n; j++) {
            int jp = wrapi(j + 1, n), jm = wrapi(j - 1, n);
            for (int i = 0; i < n; i++) {
                int ip = wrapi(i + 1, n), im = wrapi(i - 1, n);
#define F(A, a, b, d) A[IDX(c, a, b, d)]
                size_t q = IDX(c, i, j, k);
                rphi[q] = rphi[q]-gamma*((0.25*(1.0-F(T,ip,j,k)*F(T,ip,j,k))*F(nx,ip,j,k) -
                                          0.25*(1.0-F(T,im,j,k)*F(T,im,j,k))*F(nx,im,j,k))*0.5*dxi +
                                         (0.25*(1.0-F(T,i,jp,k)*F(T,i,jp,k))*F(ny,i,jp,k) -
                                          0.25*(1.0-F(T,i,jm,k)*F(T,i,jm,k))*F(ny,i,jm,k))*0.5*dxi +
                                         (0.25*(1.0-F(T,i,j,kp)*F(T,i,j,kp))*F(nz,i,j,kp) -
                                          0.25*(1.0-F(T,i,j,km)*F(T,i,j,km))*F(nz,i,j,km))*0.5*dxi);
#undef F
            }
        }
    }
}

int oracle_acdi_rhs(oracle_ctx* c, int parts, double gamma, double* rphi) {
    if (!c->P.phiflag) return MHIT_ERR_STATE;
    acdi_psi(c);
    int cd_parts = parts & (ORACLE_PHI_CONV | ORACLE_PHI_DIFF);
    if (cd_parts) acdi_rhs_conv_diff(c, cd_parts, gamma, rphi);
    else { acdi_rhs_conv_diff(c, ORACLE_PHI_DIFF, gamma, rphi); memset(rphi, 0, sizeof(double) * c->N3); }
    if (parts & ORACLE_PHI_SHARPEN) {
        acdi_normalize(c);
        acdi_sharpen(c, gamma, rphi);
    }
    return MHIT_OK;
}

/* CSF, main.f90:721-743 */
static void csf(oracle_ctx* c, double* fx, double* fy, double* fz) {
    const int n = c->n;
    const double* phi = c->phi;
    const double dxi = c->dxi, ddxi = c->ddxi, eps = c->eps, epsi = c->epsi, sigma = c->P.sigma;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
        for (int j = 0; j < n; j++) {
            int jp = wrapi(j + 1, n), jm = wrapi(j - 1, n);
            for (int i = 0; i < n; i++) {
                int ip = wrapi(i + 1, n), im = wrapi(i - 1, n);
#define F(A, a, b, d) A[IDX(c, a, b, d)]
                double ph = F(phi,i,j,k);
                double chempot = ph*(1.0-ph)*(1.0-2.0*ph)*epsi-eps*(F(phi,ip,j,k)+F(phi,im,j,k)+F(phi,i,jp,k)+F(phi,i,jm,k)+F(phi,i,j,kp)+F(phi,i,j,km)- 6.0*ph)*ddxi;
                size_t q = IDX(c, i, j, k);
                fx[q] = 6.0*sigma*chempot*0.5*(F(phi,ip,j,k)-F(phi,im,j,k))*dxi;
                fy[q] = 6.0*sigma*chempot*0.5*(F(phi,i,jp,k)-F(phi,i,jm,k))*dxi;
                fz[q] = 6.0*sigma*chempot*0.5*(F(phi,i,j,kp)-F(phi,i,j,km))*dxi;
#undef F
            }
        }
    }
}

int oracle_csf(oracle_ctx* c, double* fx, double* fy, double* fz) {
    if (!c->P.phiflag) return MHIT_ERR_STATE;
    csf(c, fx, fy, fz);
    return MHIT_OK;
}

/* Poisson right-hand side, main.f90:833-848 */
static void divergence(oracle_ctx* c, double* rhsp) {
    const int n = c->n;
    const double* u = c->u; const double* v = c->v; const double* w = c->w;
    const double cdiv = c->P.rho*c->dxi/c->P.dt;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = wrapi(k + 1, n);
        for (int j = 0; j < n; j++) {
            int jp = wrapi(j + 1, n);
            for (int i = 0; i < n; i++) {
                int ip = wrapi(i + 1, n);
                size_t q = IDX(c, i, j, k);
                double r = cdiv*(u[IDX(c, ip, j, k)]-u[q]);
                r = r + cdiv*(v[IDX(c, i, jp, k)]-v[q]);
                r = r + cdiv*(w[IDX(c, i, j, kp)]-w[q]);
                rhsp[q] = r;
            }
        }
    }
}

int oracle_divergence(oracle_ctx* c, double* rhsp) {
    divergence(c, rhsp);
    return MHIT_OK;
}

/* ------------------------------------------------------------------ */
/* Tracer interpolation, particles.f90:88-118: 1-D linear per component
 * along its own axis.  y/z cell indices are the halo-shifted local ones
 * (iy = 1 + floor(y/dx) on an array whose index 1 is the halo), i.e.
 * interior 0-based (floor(y/dx) - 1) mod n.  The x index has no halo.   */
static void interp_reference(oracle_ctx* c) {
    const int n = c->n;
    const double dx = c->dx;
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)c->np; p++) {
        double xpt = c->px[p], ypt = c->py[p], zpt = c->pz[p];
        long fx = (long)floor(xpt / dx), fy = (long)floor(ypt / dx), fz = (long)floor(zpt / dx);
        double tx = xpt/dx - (double)fx;
        double ty = ypt/dx - (double)fy;
        double tz = zpt/dx - (double)fz;
        int i0 = (int)(((fx % n) + n) % n), i1 = (int)((((fx + 1) % n) + n) % n);
        int j0 = (int)((((fy - 1) % n) + n) % n), j1 = (int)(((fy % n) + n) % n);
        int k0 = (int)((((fz - 1) % n) + n) % n), k1 = (int)(((fz % n) + n) % n);
        c->pu[p] = (1.0 - tx)*c->u[IDX(c, i0, j0, k0)] + tx*c->u[IDX(c, i1, j0, k0)];
        c->pv[p] = (1.0 - ty)*c->v[IDX(c, i0, j0, k0)] + ty*c->v[IDX(c, i0, j1, k0)];
        c->pw[p] = (1.0 - tz)*c->w[IDX(c, i0, j0, k0)] + tz*c->w[IDX(c, i0, j0, k1)];
    }
}

/* Trilinear option (not in the reference): each component interpolated
 * at its staggered location in the time-stepper frame: u at
 * ((i-1/2)dx, (j+1)dx, (k+1)dx), v at (i dx, (j+1/2)dx, (k+1)dx), w at
 * (i dx, (j+1)dx, (k+1/2)dx) for 0-based storage index (i,j,k).        */
static double trilin(const oracle_ctx* c, const double* f, double xs, double ys, double zs) {
    const int n = c->n;
    double fxl = floor(xs), fyl = floor(ys), fzl = floor(zs);
    double tx = xs - fxl, ty = ys - fyl, tz = zs - fzl;
    long ix = (long)fxl, iy = (long)fyl, iz = (long)fzl;
    int i0 = (int)(((ix % n) + n) % n), i1 = (int)((((ix + 1) % n) + n) % n);
    int j0 = (int)(((iy % n) + n) % n), j1 = (int)((((iy + 1) % n) + n) % n);
    int k0 = (int)(((iz % n) + n) % n), k1 = (int)((((iz + 1) % n) + n) % n);
    double c00 = (1.0 - tx) * f[IDX(c, i0, j0, k0)] + tx * f[IDX(c, i1, j0, k0)];
    double c10 = (1.0 - tx) * f[IDX(c, i0, j1, k0)] + tx * f[IDX(c, i1, j1, k0)];
    double c01 = (1.0 - tx) * f[IDX(c, i0, j0, k1)] + tx * f[IDX(c, i1, j0, k1)];
    double c11 = (1.0 - tx) * f[IDX(c, i0, j1, k1)] + tx * f[IDX(c, i1, j1, k1)];
    double c0 = (1.0 - ty) * c00 + ty * c10;
    double c1 = (1.0 - ty) * c01 + ty * c11;
    return (1.0 - tz) * c0 + tz * c1;
}

static void interp_trilinear(oracle_ctx* c) {
    const double dx = c->dx;
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)c->np; p++) {
        double xs = c->px[p] / dx, ys = c->py[p] / dx, zs = c->pz[p] / dx;
        c->pu[p] = trilin(c, c->u, xs + 0.5, ys - 1.0, zs - 1.0);
        c->pv[p] = trilin(c, c->v, xs, ys - 0.5, zs - 1.0);
        c->pw[p] = trilin(c, c->w, xs, ys - 1.0, zs - 0.5);
    }
}

int oracle_interp(oracle_ctx* c, double* up, double* vp, double* wp) {
    if (c->P.interp_mode == MHIT_INTERP_TRILINEAR) interp_trilinear(c);
    else interp_reference(c);
    oracle_get_tracers(c, NULL, NULL, NULL, up, vp, wp);
    return MHIT_OK;
}

/* ------------------------------------------------------------------ */
static void tracers_step(oracle_ctx* c) {
    /* main.f90:477 */
    if (c->P.interp_mode == MHIT_INTERP_TRILINEAR) interp_trilinear(c);
    else interp_reference(c);
    const double dt = c->P.dt, alpha = c->alpha, beta = c->beta, lx = c->lx;
    for (long p = 0; p < (long)c->np; p++) {
        /* main.f90:480-486 */
        c->px[p] = c->px[p]+dt*(alpha*c->pu[p]-beta*c->pu1[p]);
        c->py[p] = c->py[p]+dt*(alpha*c->pv[p]-beta*c->pv1[p]);
        c->pz[p] = c->pz[p]+dt*(alpha*c->pw[p]-beta*c->pw1[p]);
        /* main.f90:489-498 at one rank only reorder particles (sort and
         * send-to-self); values are unchanged, so they are skipped here. */
        /* particles.f90:372-376 (FLOOR returns a default integer) */
        c->px[p] = c->px[p]-lx*(double)(int)floor(c->px[p]/lx);
        c->py[p] = c->py[p]-lx*(double)(int)floor(c->py[p]/lx);
        c->pz[p] = c->pz[p]-lx*(double)(int)floor(c->pz[p]/lx);
        /* main.f90:503-508 */
        c->pu1[p] = c->pu[p]; c->pv1[p] = c->pv[p]; c->pw1[p] = c->pw[p];
    }
}

/* ------------------------------------------------------------------ */
static int one_step(oracle_ctx* c) {
    const int n = c->n;
    const size_t N3 = c->N3;
    const double dt = c->P.dt;
    double alpha = c->alpha, beta = c->beta;

    if (c->P.partflag && c->np > 0) tracers_step(c);

    if (c->P.phiflag) {
        acdi_psi(c);                                      /* 522-534 */
        double gamma = 1.0*c->gumax;                      /* 536 */
        acdi_rhs_conv_diff(c, ORACLE_PHI_CONV | ORACLE_PHI_DIFF, gamma, c->rhsphi); /* 537-564 */
        acdi_normalize(c);                                /* 582-593 */
        acdi_sharpen(c, gamma, c->rhsphi);                /* 596-618 */
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)N3; q++) {             /* 621-630 */
            c->phi[q] = c->phi[q] + dt*(alpha*c->rhsphi[q]-beta*c->rhsphi_o[q]);
            c->rhsphi_o[q] = c->rhsphi[q];
        }
    }

    ns_rhs(c, ORACLE_NS_CONV | ORACLE_NS_VISC | ORACLE_NS_FORCE, c->rhsu, c->rhsv, c->rhsw); /* 658-717 */

    if (c->P.phiflag) {
        csf(c, c->fxst, c->fyst, c->fzst);                /* 721-743 */
        const double rhoi = c->rhoi;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++) {                     /* 760-780 */
            int km = wrapi(k - 1, n);
            for (int j = 0; j < n; j++) {
                int jm = wrapi(j - 1, n);
                for (int i = 0; i < n; i++) {
                    int im = wrapi(i - 1, n);
                    size_t q = IDX(c, i, j, k);
                    c->rhsu[q] = c->rhsu[q] + 0.5*(c->fxst[IDX(c, im, j, k)]+c->fxst[q])*rhoi;
                    c->rhsv[q] = c->rhsv[q] + 0.5*(c->fyst[IDX(c, i, jm, k)]+c->fyst[q])*rhoi;
                    c->rhsw[q] = c->rhsw[q] + 0.5*(c->fzst[IDX(c, i, j, km)]+c->fzst[q])*rhoi;
                    c->u[q] = c->u[q] + dt*(alpha*c->rhsu[q]-beta*c->rhsu_o[q]);
                    c->v[q] = c->v[q] + dt*(alpha*c->rhsv[q]-beta*c->rhsv_o[q]);
                    c->w[q] = c->w[q] + dt*(alpha*c->rhsw[q]-beta*c->rhsw_o[q]);
                    c->rhsu_o[q] = c->rhsu[q];
                    c->rhsv_o[q] = c->rhsv[q];
                    c->rhsw_o[q] = c->rhsw[q];
                }
            }
        }
    } else {
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)N3; q++) {             /* 783-796 */
            c->u[q] = c->u[q] + dt*(alpha*c->rhsu[q]-beta*c->rhsu_o[q]);
            c->v[q] = c->v[q] + dt*(alpha*c->rhsv[q]-beta*c->rhsv_o[q]);
            c->w[q] = c->w[q] + dt*(alpha*c->rhsw[q]-beta*c->rhsw_o[q]);
            c->rhsu_o[q] = c->rhsu[q];
            c->rhsv_o[q] = c->rhsv[q];
            c->rhsw_o[q] = c->rhsw[q];
        }
    }
    c->alpha = 1.5;                                       /* 801-802 */
    c->beta = 0.5;

    divergence(c, c->rhsp);                               /* 833-848 */
    poisson(c, c->rhsp, c->p);                            /* 853-912 */

    /* correction and local sums, 942-962 */
    const double dxi = c->dxi, rho = c->P.rho;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int km = wrapi(k - 1, n);
        for (int j = 0; j < n; j++) {
            int jm = wrapi(j - 1, n);
            for (int i = 0; i < n; i++) {
                int im = wrapi(i - 1, n);
                size_t q = IDX(c, i, j, k);
                double pc = c->p[q];
                c->u[q] = c->u[q] - dt/rho*(pc-c->p[IDX(c, im, j, k)])*dxi;
                c->v[q] = c->v[q] - dt/rho*(pc-c->p[IDX(c, i, jm, k)])*dxi;
                c->w[q] = c->w[q] - dt/rho*(pc-c->p[IDX(c, i, j, km)])*dxi;
            }
        }
    }
    double umean = 0.0, vmean = 0.0, wmean = 0.0;
    for (size_t q = 0; q < N3; q++) { umean = umean + c->u[q]; vmean = vmean + c->v[q]; wmean = wmean + c->w[q]; }
    umean = umean/n/n/n;                                  /* 967-969 */
    vmean = vmean/n/n/n;
    wmean = wmean/n/n/n;
    double gumean = umean, gvmean = vmean, gwmean = wmean;/* 972-974, 1 rank */
    const double ranks = 1.0;
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)N3; q++) {                 /* 977-981 */
        c->u[q] = c->u[q]-(gumean/ranks);
        c->v[q] = c->v[q]-(gvmean/ranks);
        c->w[q] = c->w[q]-(gwmean/ranks);
    }
    c->umean = gumean; c->vmean = gvmean; c->wmean = gwmean;
    double uc = c->u[0], vc = c->v[0], wc = c->w[0];      /* 998-1001 */
    for (size_t q = 1; q < N3; q++) {
        if (c->u[q] > uc) uc = c->u[q];
        if (c->v[q] > vc) vc = c->v[q];
        if (c->w[q] > wc) wc = c->w[q];
    }
    double umax = fmax(wc, fmax(uc, vc));
    c->gumax = umax;                                      /* 1002 */
    c->cou = c->gumax*dt*dxi;                             /* 1003 */
    c->step++;
    if (c->P.cfl_abort && c->cou > 7) { c->halted = 1; return MHIT_ERR_CFL; }  /* 1006 */
    return MHIT_OK;
}

int oracle_step(oracle_ctx* c, int nsteps) {
    for (int s = 0; s < nsteps; s++) {
        if (c->halted) return MHIT_ERR_CFL;
        int rc = one_step(c);
        if (rc) return rc;
    }
    return MHIT_OK;
}

/* ------------------------------------------------------------------ */
/* Diagnostics (not in the solver; dissipation follows
 * post/compute_diss/main.f90:72-117 with dx = lx/n and nu = mu/rho).   */
/* Compensated (Neumaier) running sum: the diagnostics sum up to 512^3
 * terms of one sign, where a plain running sum loses up to ~1e-10 relative;
 * this keeps the oracle's sums exact to a few ulp, so that they check the
 * GPU's tree reductions rather than the other way round.                 */
typedef struct { double s, c; } ksum;
static inline void ksum_add(ksum* a, double x) {
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}
static inline double ksum_value(const ksum* a) { return a->s + a->c; }

int oracle_diag(oracle_ctx* c, mhit_diag* d) {
    const int n = c->n;
    const double dxi = c->dxi, nu = c->P.mu / c->P.rho;
    memset(d, 0, sizeof(*d));
    d->step = c->step; d->alpha = c->alpha; d->beta = c->beta; d->gumax = c->gumax;
    d->cou = c->cou; d->umean = c->umean; d->vmean = c->vmean; d->wmean = c->wmean;
    ksum ke = {0.0, 0.0}, diss = {0.0, 0.0}, mass = {0.0, 0.0};
    double maxdiv = 0.0;
    double pmin = 0.0, pmax = 0.0;
    int64_t nans = 0;
    for (int k = 0; k < n; k++) {
        int kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
        for (int j = 0; j < n; j++) {
            int jp = wrapi(j + 1, n), jm = wrapi(j - 1, n);
            for (int i = 0; i < n; i++) {
                int ip = wrapi(i + 1, n), im = wrapi(i - 1, n);
                size_t q = IDX(c, i, j, k);
                double uu = c->u[q], vv = c->v[q], ww = c->w[q];
                if (!isfinite(uu) || !isfinite(vv) || !isfinite(ww)) nans++;
                ksum_add(&ke, uu * uu + vv * vv + ww * ww);
                double dudx = (c->u[IDX(c, ip, j, k)] - c->u[IDX(c, im, j, k)]) * 0.5 * dxi;
                double dudy = (c->u[IDX(c, i, jp, k)] - c->u[IDX(c, i, jm, k)]) * 0.5 * dxi;
                double dudz = (c->u[IDX(c, i, j, kp)] - c->u[IDX(c, i, j, km)]) * 0.5 * dxi;
                double dvdx = (c->v[IDX(c, ip, j, k)] - c->v[IDX(c, im, j, k)]) * 0.5 * dxi;
                double dvdy = (c->v[IDX(c, i, jp, k)] - c->v[IDX(c, i, jm, k)]) * 0.5 * dxi;
                double dvdz = (c->v[IDX(c, i, j, kp)] - c->v[IDX(c, i, j, km)]) * 0.5 * dxi;
                double dwdx = (c->w[IDX(c, ip, j, k)] - c->w[IDX(c, im, j, k)]) * 0.5 * dxi;
                double dwdy = (c->w[IDX(c, i, jp, k)] - c->w[IDX(c, i, jm, k)]) * 0.5 * dxi;
                double dwdz = (c->w[IDX(c, i, j, kp)] - c->w[IDX(c, i, j, km)]) * 0.5 * dxi;
                double S11 = dudx, S22 = dvdy, S33 = dwdz;
                double S12 = 0.5 * (dudy + dvdx), S13 = 0.5 * (dudz + dwdx), S23 = 0.5 * (dvdz + dwdy);
                ksum_add(&diss, S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23));
                double dv = ((c->u[IDX(c, ip, j, k)] - uu) + (c->v[IDX(c, i, jp, k)] - vv) +
                             (c->w[IDX(c, i, j, kp)] - ww)) * dxi;
                if (fabs(dv) > maxdiv) maxdiv = fabs(dv);
                if (c->phi) {
                    double ph = c->phi[q];
                    if (!isfinite(ph)) nans++;
                    ksum_add(&mass, ph);
                    if (q == 0 || ph < pmin) pmin = ph;
                    if (q == 0 || ph > pmax) pmax = ph;
                }
            }
        }
    }
    double N3 = (double)c->N3;
    d->ke = 0.5 * ksum_value(&ke) / N3;
    d->dissipation = 2.0 * nu * ksum_value(&diss) / N3;
    d->maxdiv = maxdiv;
    d->phi_mass = ksum_value(&mass) * c->dx * c->dx * c->dx;
    d->phi_min = pmin; d->phi_max = pmax;
    d->nan_count = nans;
    return MHIT_OK;
}

/* ------------------------------------------------------------------ */
/* drop statistics, post/unionfind_periodic/main.f90:97-441 and
 * post/DropCountSize/calculate_deq.f90.  The reference labels by iterated
 * minimum-label propagation; a flood fill from each unvisited masked cell
 * in scan order gives the same components with the same numbering (a
 * component's first cell in scan order is its minimum label).            */
int oracle_drop_stats(int n, const double* phi, const double* u, const double* v,
                      const double* w, double thr, int conn, int64_t maxd, double* out,
                      int64_t* nd) {
    if (conn != 6 && conn != 26) return 1;
    const size_t N3 = (size_t)n * n * n;
    const double pi = 4.0 * atan(1.0), lx = 2.0 * pi, dx = lx / (double)n;
    double* x = (double*)malloc(sizeof(double) * n);   /* cell centres, main.f90:295-299 */
    unsigned char* seen = (unsigned char*)calloc(N3, 1);
    int64_t* queue = (int64_t*)malloc(sizeof(int64_t) * N3);
    if (!x || !seen || !queue) { free(x); free(seen); free(queue); return 2; }
    x[0] = dx / 2.0;
    for (int i = 1; i < n; i++) x[i] = x[i - 1] + dx;
    int64_t count = 0;
    for (size_t q0 = 0; q0 < N3; q0++) {
        if (seen[q0] || !(phi[q0] >= thr)) continue;
        double cells = 0, sx = 0, cx = 0, sy = 0, cy = 0, sz = 0, cz = 0, su = 0, sv = 0, sw = 0;
        size_t head = 0, tail = 0;
        queue[tail++] = (int64_t)q0;
        seen[q0] = 1;
        while (head < tail) {
            const size_t q = (size_t)queue[head++];
            const int i = (int)(q % n), j = (int)((q / n) % n), k = (int)(q / ((size_t)n * n));
            cells += 1.0;
            sx += sin(2.0 * pi * x[i] / lx); cx += cos(2.0 * pi * x[i] / lx);
            sy += sin(2.0 * pi * x[j] / lx); cy += cos(2.0 * pi * x[j] / lx);
            sz += sin(2.0 * pi * x[k] / lx); cz += cos(2.0 * pi * x[k] / lx);
            if (u) { su += u[q]; sv += v[q]; sw += w[q]; }
            for (int dk = -1; dk <= 1; dk++)
                for (int dj = -1; dj <= 1; dj++)
                    for (int di = -1; di <= 1; di++) {
                        const int m = abs(di) + abs(dj) + abs(dk);
                        if (m == 0 || (conn == 6 && m != 1)) continue;
                        const size_t r = ((size_t)wrapi(k + dk, n) * n + wrapi(j + dj, n)) * n +
                                         wrapi(i + di, n);
                        if (!seen[r] && phi[r] >= thr) { seen[r] = 1; queue[tail++] = (int64_t)r; }
                    }
        }
        if (count < maxd) {
            double* o = out + (size_t)count * 10;
            double xc = lx * atan2(sx, cx) / (2.0 * pi), yc = lx * atan2(sy, cy) / (2.0 * pi),
                   zc = lx * atan2(sz, cz) / (2.0 * pi);
            if (xc < 0.0) xc += lx;
            if (yc < 0.0) yc += lx;
            if (zc < 0.0) zc += lx;
            o[0] = cells; o[1] = (double)q0;
            o[2] = xc; o[3] = yc; o[4] = zc;
            o[5] = su / cells; o[6] = sv / cells; o[7] = sw / cells;
            o[8] = cells * dx * dx * dx;
            o[9] = cbrt(6.0 * o[8] / pi);
        }
        count++;
    }
    free(x); free(seen); free(queue);
    *nd = count;
    return 0;
}
