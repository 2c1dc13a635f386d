// stencil.hip -- finite-difference passes of the MHIT36 timestep (gfx950).
//
// Each kernel restates the reference loop it cites (aroccon/MHIT36_mixedp,
// multi/main.f90) with the reference's operand order; the library is built
// without FP contraction.  Halo exchanges are replaced by periodic index
// wrap.  The velocity state is stored with a deferred mean: the reference
// removes the mean from u,v,w after the correction (main.f90:977-981); here
// the correction writes u + mean and every reader subtracts mean on load,
// which yields the same FP64 values without a second pass over u,v,w.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mhit {

#define MHIT_CELL_INDEX(g)                                   \
    const int n = (g).n;                                     \
    const int i = blockIdx.x * 64 + threadIdx.x;             \
    const int j = blockIdx.y * 4 + threadIdx.y;              \
    const int k = blockIdx.z;                                \
    if (i >= n || j >= n) return;                            \
    const int ip = wrap_p(i, n), im = wrap_m(i, n);          \
    const int jp = wrap_p(j, n), jm = wrap_m(j, n);          \
    const int kp = wrap_p(k, n), km = wrap_m(k, n);          \
    (void)ip; (void)im; (void)jp; (void)jm; (void)kp; (void)km; \
    const size_t q = idx3(i, j, k, n)

// ----------------------------------------------------------------- ACDI
// P1, main.f90:522-534 (all points)
__global__ __launch_bounds__(256)
void k_acdi_psi(const double* __restrict__ phi, double* __restrict__ psi,
                double* __restrict__ tpsi, Grid g, const DevScalars* __restrict__ S) {
    if (S->halted) return;
    MHIT_CELL_INDEX(g);
    const double val = fmin(phi[q], 1.0);
    const double ps = g.eps * log((val + g.enm) / (1.0 - val + g.enm));
    psi[q] = ps;
    tpsi[q] = tanh(0.5 * ps * g.epsi);
}

// P2, main.f90:536-564: convective + diffusive flux, raw normals
__global__ __launch_bounds__(256)
void k_acdi_rhs(const double* __restrict__ u, const double* __restrict__ v,
                const double* __restrict__ w, const double* __restrict__ phi,
                const double* __restrict__ psi, double* __restrict__ rphi,
                double* __restrict__ nx, double* __restrict__ ny, double* __restrict__ nz,
                Grid g, const DevScalars* __restrict__ S) {
    if (S->halted) return;
    MHIT_CELL_INDEX(g);
    const double m0 = S->mean[0], m1 = S->mean[1], m2 = S->mean[2];
    const double gamma = 1.0 * S->gumax;
    const double dxi = g.dxi, ddxi = g.ddxi, eps = g.eps;
    const double uc = u[q] - m0, up = u[idx3(ip, j, k, n)] - m0;
    const double vc = v[q] - m1, vp = v[idx3(i, jp, k, n)] - m1;
    const double wc = w[q] - m2, wp = w[idx3(i, j, kp, n)] - m2;
    const double pc = phi[q];
    const double pxp = phi[idx3(ip, j, k, n)], pxm = phi[idx3(im, j, k, n)];
    const double pyp = phi[idx3(i, jp, k, n)], pym = phi[idx3(i, jm, k, n)];
    const double pzp = phi[idx3(i, j, kp, n)], pzm = phi[idx3(i, j, km, n)];
    double r = -(up * 0.5 * (pxp + pc) - uc * 0.5 * (pc + pxm)) * dxi
               - (vp * 0.5 * (pyp + pc) - vc * 0.5 * (pc + pym)) * dxi
               - (wp * 0.5 * (pzp + pc) - wc * 0.5 * (pc + pzm)) * dxi
               + gamma * (eps * (pxp - 2.0 * pc + pxm) * ddxi +
                          eps * (pyp - 2.0 * pc + pym) * ddxi +
                          eps * (pzp - 2.0 * pc + pzm) * ddxi);
    rphi[q] = r;
    nx[q] = (psi[idx3(ip, j, k, n)] - psi[idx3(im, j, k, n)]);
    ny[q] = (psi[idx3(i, jp, k, n)] - psi[idx3(i, jm, k, n)]);
    nz[q] = (psi[idx3(i, j, kp, n)] - psi[idx3(i, j, km, n)]);
}

// P3, main.f90:582-593
__global__ __launch_bounds__(256)
void k_acdi_normalize(double* __restrict__ nx, double* __restrict__ ny, double* __restrict__ nz,
                      Grid g, const DevScalars* __restrict__ S) {
    if (S->halted) return;
    MHIT_CELL_INDEX(g);
    const double a = nx[q], b = ny[q], c = nz[q];
    const double normag = 1.0 / (sqrt(a * a + b * b + c * c) + g.enm);
    nx[q] = a * normag;
    ny[q] = b * normag;
    nz[q] = c * normag;
}

// P4 + P5, main.f90:596-630: sharpening flux and AB2 of phi (in place)
__global__ __launch_bounds__(256)
void k_acdi_update(double* __restrict__ phi, const double* __restrict__ rphi_in,
                   double* __restrict__ rphi_o, const double* __restrict__ T,
                   const double* __restrict__ nx, const double* __restrict__ ny,
                   const double* __restrict__ nz, Grid g, const DevScalars* __restrict__ S) {
    if (S->halted) return;
    MHIT_CELL_INDEX(g);
    const double gamma = 1.0 * S->gumax;
    const double alpha = S->alpha, beta = S->beta, dxi = g.dxi;
    const double Txp = T[idx3(ip, j, k, n)], Txm = T[idx3(im, j, k, n)];
    const double Typ = T[idx3(i, jp, k, n)], Tym = T[idx3(i, jm, k, n)];
    const double Tzp = T[idx3(i, j, kp, n)], Tzm = T[idx3(i, j, km, n)];
    double r = rphi_in[q] -
               gamma * ((0.25 * (1.0 - Txp * Txp) * nx[idx3(ip, j, k, n)] -
                         0.25 * (1.0 - Txm * Txm) * nx[idx3(im, j, k, n)]) * 0.5 * dxi +
                        (0.25 * (1.0 - Typ * Typ) * ny[idx3(i, jp, k, n)] -
                         0.25 * (1.0 - Tym * Tym) * ny[idx3(i, jm, k, n)]) * 0.5 * dxi +
                        (0.25 * (1.0 - Tzp * Tzp) * nz[idx3(i, j, kp, n)] -
                         0.25 * (1.0 - Tzm * Tzm) * nz[idx3(i, j, km, n)]) * 0.5 * dxi);
    phi[q] = phi[q] + g.dt * (alpha * r - beta * rphi_o[q]);
    rphi_o[q] = r;
}

int launch_acdi(mhit_ctx* c) {
    CellLaunch L(c->n);
    k_acdi_psi<<<L.grid, L.block, 0, c->stream>>>(c->phi, c->psi, c->tanh_psi, c->g, c->S);
    prof_mark(c, ST_ACDI_PSI);
    k_acdi_rhs<<<L.grid, L.block, 0, c->stream>>>(c->u, c->v, c->w, c->phi, c->psi, c->rhsphi,
                                                  c->normx, c->normy, c->normz, c->g, c->S);
    prof_mark(c, ST_ACDI_RHS);
    k_acdi_normalize<<<L.grid, L.block, 0, c->stream>>>(c->normx, c->normy, c->normz, c->g, c->S);
    prof_mark(c, ST_ACDI_NORM);
    k_acdi_update<<<L.grid, L.block, 0, c->stream>>>(c->phi, c->rhsphi, c->rhsphi_o, c->tanh_psi,
                                                     c->normx, c->normy, c->normz, c->g, c->S);
    prof_mark(c, ST_ACDI_UPDATE);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

// ------------------------------------------------------------------ CSF
// S1, main.f90:721-743 (phi already advanced)
__global__ __launch_bounds__(256)
void k_csf(const double* __restrict__ phi, double* __restrict__ fx, double* __restrict__ fy,
           double* __restrict__ fz, Grid g, const DevScalars* __restrict__ S) {
    if (S->halted) return;
    MHIT_CELL_INDEX(g);
    const double ph = phi[q];
    const double pxp = phi[idx3(ip, j, k, n)], pxm = phi[idx3(im, j, k, n)];
    const double pyp = phi[idx3(i, jp, k, n)], pym = phi[idx3(i, jm, k, n)];
    const double pzp = phi[idx3(i, j, kp, n)], pzm = phi[idx3(i, j, km, n)];
    const double chempot = ph * (1.0 - ph) * (1.0 - 2.0 * ph) * g.epsi -
                           g.eps * (pxp + pxm + pyp + pym + pzp + pzm - 6.0 * ph) * g.ddxi;
    fx[q] = 6.0 * g.sigma * chempot * 0.5 * (pxp - pxm) * g.dxi;
    fy[q] = 6.0 * g.sigma * chempot * 0.5 * (pyp - pym) * g.dxi;
    fz[q] = 6.0 * g.sigma * chempot * 0.5 * (pzp - pzm) * g.dxi;
}

// ------------------------------------------------------------------- NS
// N1 (main.f90:658-717) fused with the face force + AB2 of S2
// (main.f90:760-780) or with N2 (783-796).  Writes u* to us,vs,ws.
template <bool PHI>
__global__ __launch_bounds__(256)
void k_ns(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
          const double* __restrict__ fx, const double* __restrict__ fy, const double* __restrict__ fz,
          double* __restrict__ ru_o, double* __restrict__ rv_o, double* __restrict__ rw_o,
          double* __restrict__ us, double* __restrict__ vs, double* __restrict__ ws,
          const double* __restrict__ mysin, const double* __restrict__ mycos,
          Grid g, const DevScalars* __restrict__ S) {
    if (S->halted) return;
    MHIT_CELL_INDEX(g);
    const double m0 = S->mean[0], m1 = S->mean[1], m2 = S->mean[2];
    const double alpha = S->alpha, beta = S->beta;
    const double dxi = g.dxi, ddxi = g.ddxi, mu = g.mu, rhoi = g.rhoi;
#define U(a, b, c) (u[idx3(a, b, c, n)] - m0)
#define V(a, b, c) (v[idx3(a, b, c, n)] - m1)
#define W(a, b, c) (w[idx3(a, b, c, n)] - m2)
    const double u0 = U(i, j, k), v0 = V(i, j, k), w0 = W(i, j, k);
    const double uxp = U(ip, j, k), uxm = U(im, j, k), uyp = U(i, jp, k), uym = U(i, jm, k);
    const double uzp = U(i, j, kp), uzm = U(i, j, km);
    const double vxp = V(ip, j, k), vxm = V(im, j, k), vyp = V(i, jp, k), vym = V(i, jm, k);
    const double vzp = V(i, j, kp), vzm = V(i, j, km);
    const double wxp = W(ip, j, k), wxm = W(im, j, k), wyp = W(i, jp, k), wym = W(i, jm, k);
    const double wzp = W(i, j, kp), wzm = W(i, j, km);
    double h11 = (uxp + u0) * (uxp + u0) - (u0 + uxm) * (u0 + uxm);
    double h12 = (uyp + u0) * (vyp + V(im, jp, k)) - (u0 + uym) * (v0 + vxm);
    double h13 = (uzp + u0) * (wzp + W(im, j, kp)) - (u0 + uzm) * (w0 + wxm);
    double h21 = (uxp + U(ip, jm, k)) * (vxp + v0) - (u0 + uym) * (v0 + vxm);
    double h22 = (vyp + v0) * (vyp + v0) - (v0 + vym) * (v0 + vym);
    double h23 = (wzp + W(i, jm, kp)) * (vzp + v0) - (w0 + wym) * (v0 + vzm);
    double h31 = (wxp + w0) * (uxp + U(ip, j, km)) - (w0 + wxm) * (u0 + uzm);
    double h32 = (vyp + V(i, jp, km)) * (wyp + w0) - (v0 + vzm) * (w0 + wym);
    double h33 = (wzp + w0) * (wzp + w0) - (w0 + wzm) * (w0 + wzm);
#undef U
#undef V
#undef W
    h11 = h11 * 0.25 * dxi; h12 = h12 * 0.25 * dxi; h13 = h13 * 0.25 * dxi;
    h21 = h21 * 0.25 * dxi; h22 = h22 * 0.25 * dxi; h23 = h23 * 0.25 * dxi;
    h31 = h31 * 0.25 * dxi; h32 = h32 * 0.25 * dxi; h33 = h33 * 0.25 * dxi;
    double ru = -(h11 + h12 + h13);
    double rv = -(h21 + h22 + h23);
    double rw = -(h31 + h32 + h33);
    h11 = mu * (uxp - 2.0 * u0 + uxm) * ddxi;
    h12 = mu * (uyp - 2.0 * u0 + uym) * ddxi;
    h13 = mu * (uzp - 2.0 * u0 + uzm) * ddxi;
    h21 = mu * (vxp - 2.0 * v0 + vxm) * ddxi;
    h22 = mu * (vyp - 2.0 * v0 + vym) * ddxi;
    h23 = mu * (vzp - 2.0 * v0 + vzm) * ddxi;
    h31 = mu * (wxp - 2.0 * w0 + wxm) * ddxi;
    h32 = mu * (wyp - 2.0 * w0 + wym) * ddxi;
    h33 = mu * (wzp - 2.0 * w0 + wzm) * ddxi;
    ru = ru + (h11 + h12 + h13) * rhoi;
    rv = rv + (h21 + h22 + h23) * rhoi;
    rw = rw + (h31 + h32 + h33) * rhoi;
    // ABC forcing: kg = k+2, jg = j+2 in Fortran indexing -> table k+1, j+1
    ru = ru + g.f3 * mysin[k + 1] + g.f2 * mycos[j + 1];
    rv = rv + g.f1 * mysin[i] + g.f3 * mycos[k + 1];
    rw = rw + g.f2 * mysin[j + 1] + g.f1 * mycos[i];
    if (PHI) {
        ru = ru + 0.5 * (fx[idx3(im, j, k, n)] + fx[q]) * rhoi;
        rv = rv + 0.5 * (fy[idx3(i, jm, k, n)] + fy[q]) * rhoi;
        rw = rw + 0.5 * (fz[idx3(i, j, km, n)] + fz[q]) * rhoi;
    }
    const double dt = g.dt;
    us[q] = u0 + dt * (alpha * ru - beta * ru_o[q]);
    vs[q] = v0 + dt * (alpha * rv - beta * rv_o[q]);
    ws[q] = w0 + dt * (alpha * rw - beta * rw_o[q]);
    ru_o[q] = ru;
    rv_o[q] = rv;
    rw_o[q] = rw;
}

int launch_ns(mhit_ctx* c) {
    CellLaunch L(c->n);
    if (c->P.phiflag) {
        k_csf<<<L.grid, L.block, 0, c->stream>>>(c->phi, c->fxst, c->fyst, c->fzst, c->g, c->S);
        prof_mark(c, ST_CSF);
        k_ns<true><<<L.grid, L.block, 0, c->stream>>>(
            c->u, c->v, c->w, c->fxst, c->fyst, c->fzst, c->rhsu_o, c->rhsv_o, c->rhsw_o,
            c->us, c->vs, c->ws, c->mysin, c->mycos, c->g, c->S);
    } else {
        k_ns<false><<<L.grid, L.block, 0, c->stream>>>(
            c->u, c->v, c->w, nullptr, nullptr, nullptr, c->rhsu_o, c->rhsv_o, c->rhsw_o,
            c->us, c->vs, c->ws, c->mysin, c->mycos, c->g, c->S);
    }
    prof_mark(c, ST_NS);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

// ----------------------------------------------------------- correction
// C, main.f90:942-962: u = u* - dt/rho (p - p(-1)) dxi with the pressure
// promoted to FP64, plus per-block sums and maxima of the corrected u,v,w.
// A workgroup owns a 64 x 8 (or 32 x 16) tile of columns and marches a chunk of z
// planes (correct_zc): p(z-1) of its column stays in a register, the
// plane's p with a one-point halo at -x and -y is staged in LDS, and the
// next plane is fetched into registers during the update.  The number of partials is
// fixed by the grid, so the reduction order is deterministic.
constexpr int kCorrZ = 16;                    // z split of the diagnostics pass
constexpr int CO_NT = 512;                    // threads per workgroup
// tile columns: 64, or 32 where n is not a multiple of 64 (96^3: a
// 64-column tile would leave a third of the threads without a cell); rows
// make up the workgroup
static int co_mx(int n) { return (n > 64 && n % 64 != 0 && n % 32 == 0) ? 32 : 64; }
// z planes per workgroup: the longest of 32, 16, ..., 2 that gives a box
// at least 256 workgroups (256^3: 32, 128^3: 16, 96^3: 8, 64^3 and below:
// 2; at 64^3 8-plane chunks left 192 CUs idle, 14.8 us), and 16 from
// 512^3 (1.36 -> 1.31 ms; 8 planes gain no more once the larger finalize
// is counted; at 256^3 16 and 8 planes measured the same as 32).  A
// function of n only, never of the device: the number of partials, and so
// the reduction order, is fixed per grid size.
static int correct_zc(int n) {
    if (n >= 512) return 16;
    const long long tiles = (long long)div_up(n, co_mx(n)) * div_up(n, CO_NT / co_mx(n));
    int zc = 32;
    while (zc > 2 && tiles * div_up(n, zc) < 256) zc /= 2;
    return zc;
}

static dim3 correct_grid(int n) {
    const int zc = correct_zc(n) < n ? correct_zc(n) : n;
    return dim3(div_up(n, co_mx(n)), div_up(n, CO_NT / co_mx(n)), div_up(n, zc));
}

// Arguments of one box (a batched launch reads a table of them).
template <typename T>
struct CorrArgs {
    const double *us, *vs, *ws;
    const T* p;
    double *u, *v, *w, *partials;
    Grid g;
    const DevScalars* S;
    int zc, xcd;
};

template <typename T, int CMX>
__global__ __launch_bounds__(CO_NT)
void k_correct(CorrArgs<T> one, const CorrArgs<T>* __restrict__ tab, int nbox) {
    const Blk blk = launch_blk(nbox);
    CorrArgs<T> A = one;
    if (tab) A = tab[blk.box];
    const MHIT_GLOBAL double* __restrict__ us = gmem(A.us);
    const MHIT_GLOBAL double* __restrict__ vs = gmem(A.vs);
    const MHIT_GLOBAL double* __restrict__ ws = gmem(A.ws);
    const MHIT_GLOBAL T* __restrict__ p = gmem(A.p);
    MHIT_GLOBAL double* __restrict__ u = gmem(A.u);
    MHIT_GLOBAL double* __restrict__ v = gmem(A.v);
    MHIT_GLOBAL double* __restrict__ w = gmem(A.w);
    MHIT_GLOBAL double* __restrict__ partials = gmem(A.partials);
    const Grid g = A.g;
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    const int zc = A.zc, xcd = A.xcd;
    constexpr int MX = CMX, MY = CO_NT / CMX, NT = MX * MY, W = MX + 1, H = MY + 1;
    constexpr int NP = (W * H + NT - 1) / NT;
    // p of the plane, x = -1..MX-1, y = -1..MY-1; two buffers, so that the
    // next plane's commit needs no barrier after this plane's reads
    __shared__ T sp2[2][W * H];
    __shared__ double red[6][NT / 64];
    const int n = g.n;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * MX + tx;
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz, blk, xcd);
    const int i0 = tbx * MX, j0 = tby * MY, z0 = tbz * zc;
    const int z1 = min(z0 + zc, n);
    const int i = i0 + tx, j = j0 + ty;
    const bool halted = S->halted != 0;
    const bool live = !halted && i < n && j < n;
    const double dtr = g.dt / g.rho;
    double su = 0.0, sv = 0.0, sw = 0.0;
    double mu_ = -INFINITY, mv_ = -INFINITY, mw_ = -INFINITY;
    if (!halted) {
        // in-plane offsets of the thread's halo points, fixed for the whole march
        int off[NP];
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int e = tid + t * NT;
            const int lx = e % W, ly = e / W;
            off[t] = (e < W * H) ? ((j0 + ly - 1 + n) % n) * n + (i0 + lx - 1 + n) % n : 0;
        }
        T pf[NP];
        auto fetch = [&](int k) {
            const MHIT_GLOBAL T* __restrict__ pl = p + (size_t)k * n * n;
#pragma unroll
            for (int t = 0; t < NP; t++)
                if ((W * H) % NT == 0 || tid + t * NT < W * H) pf[t] = pl[off[t]];
        };
        double pkm = live ? (double)p[idx3(i, j, wrap_m(z0, n), n)] : 0.0;
        fetch(z0);
        for (int k = z0; k < z1; k++) {
            const size_t q = idx3(live ? i : 0, live ? j : 0, k, n);
            double usq = 0.0, vsq = 0.0, wsq = 0.0;
            if (live) { usq = ld_once(&us[q]); vsq = ld_once(&vs[q]); wsq = ld_once(&ws[q]); }   // in flight across the barrier
            T* const sp = sp2[(k - z0) & 1];
#pragma unroll
            for (int t = 0; t < NP; t++) {
                const int e = tid + t * NT;
                if (e < W * H) sp[e] = pf[t];
            }
            lds_sync();
            if (k + 1 < z1) fetch(k + 1);
            if (live) {
                const int c = (ty + 1) * W + tx + 1;
                const double pc = (double)sp[c];
                const double uu = usq - dtr * (pc - (double)sp[c - 1]) * g.dxi;
                const double vv = vsq - dtr * (pc - (double)sp[c - W]) * g.dxi;
                const double ww = wsq - dtr * (pc - pkm) * g.dxi;
                st_once(&u[q], uu); st_once(&v[q], vv); st_once(&w[q], ww);
                su += uu; sv += vv; sw += ww;
                mu_ = fmax(mu_, uu); mv_ = fmax(mv_, vv); mw_ = fmax(mw_, ww);
                pkm = pc;
            }
            // (the next commit goes to the other buffer; the one read here
            // is written again only after the next plane's barrier)
        }
    }
    su = wave_sum(su); sv = wave_sum(sv); sw = wave_sum(sw);
    mu_ = wave_max(mu_); mv_ = wave_max(mv_); mw_ = wave_max(mw_);
    const int wv = tid >> 6;
    if ((tid & 63) == 0) {
        red[0][wv] = su; red[1][wv] = sv; red[2][wv] = sw;
        red[3][wv] = mu_; red[4][wv] = mv_; red[5][wv] = mw_;
    }
    __syncthreads();
    if (tid < 6) {
        const int f = tid;
        double r = red[f][0];
        for (int t = 1; t < NT / 64; t++) r = (f < 3) ? r + red[f][t] : fmax(r, red[f][t]);
        // partials in tile order, so the reduction order does not depend on
        // the workgroup order (xcd_tile)
        const size_t b = ((size_t)tbz * blk.ny + tby) * blk.nx + tbx;
        partials[b * 6 + f] = r;
    }
}

template <typename T>
static CorrArgs<T> corr_args(mhit_ctx* c) {
    CorrArgs<T> a;
    a.us = c->us; a.vs = c->vs; a.ws = c->ws;
    a.p = (const T*)c->pres;
    a.u = c->u; a.v = c->v; a.w = c->w;
    a.partials = c->partials;
    a.g = c->g;
    a.S = c->S;
    a.zc = correct_zc(c->n) < c->n ? correct_zc(c->n) : c->n;
    // XCD-aware tile order always: neighbouring tiles' p halo rows then hit
    // in the same L2 (1.39 -> 1.36 ms at 512^3; the partials are indexed by
    // tile, so the order changes no bit)
    a.xcd = 1;
    return a;
}

template <typename T>
static void go_correct(int n, const CorrArgs<T>& one, const CorrArgs<T>* tab, int nb, hipStream_t st) {
    dim3 grid = correct_grid(n);
    grid.z *= nb;
    if (co_mx(n) == 32) k_correct<T, 32><<<grid, dim3(32, CO_NT / 32, 1), 0, st>>>(one, tab, nb);
    else k_correct<T, 64><<<grid, dim3(64, CO_NT / 64, 1), 0, st>>>(one, tab, nb);
}

int launch_correct(mhit_ctx* c) {
    if (c->fp32) go_correct<float>(c->n, corr_args<float>(c), nullptr, 1, c->stream);
    else go_correct<double>(c->n, corr_args<double>(c), nullptr, 1, c->stream);
    prof_mark(c, ST_CORRECT);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

template <typename T>
static int correct_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    std::vector<CorrArgs<T>> t(nb);
    for (int b = 0; b < nb; b++) t[b] = corr_args<T>(boxes[b]);
    const CorrArgs<T>* tab = arena.put(t.data(), nb);
    if (!tab) { set_error("batched launch: argument arena full"); return MHIT_ERR_STATE; }
    go_correct<T>(boxes[0]->n, t[0], tab, nb, st);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

int launch_correct_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    mhit_ctx* c = boxes[0];
    for (int b = 1; b < nb; b++)
        if (boxes[b]->n != c->n || boxes[b]->fp32 != c->fp32) {
            set_error("batched correction: boxes differ in size or precision");
            return MHIT_ERR_ARG;
        }
    return c->fp32 ? correct_batch<float>(boxes, nb, arena, st) : correct_batch<double>(boxes, nb, arena, st);
}

int correct_block_count(int n) {
    const dim3 g = correct_grid(n);
    return (int)(g.x * g.y * g.z);
}

int diag_block_count(int n) {
    return div_up(n, 64) * div_up(n, 4) * (kCorrZ < n ? kCorrZ : n);
}

// Finalize: means (main.f90:964-974), signed max after the mean removal
// (998-1002, max(u - m) = max(u) - m as rounding is monotone), CFL
// (1003-1006), AB2 switch (801-802).  One block, fixed reduction order.
struct FinArgs {
    const double* partials;
    int nblocks, cfl_abort;
    Grid g;
    DevScalars* S;
};

// One workgroup per box (blockIdx.x is the box of a batched launch).
__global__ __launch_bounds__(1024)
void k_finalize(FinArgs one, const FinArgs* __restrict__ tab) {
    FinArgs A = one;
    if (tab) A = tab[blockIdx.x];
    const MHIT_GLOBAL double* __restrict__ partials = gmem(A.partials);
    const int nblocks = A.nblocks, cfl_abort = A.cfl_abort;
    const Grid g = A.g;
    MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    __shared__ double red[6][1024];
    const int t = threadIdx.x;
    if (S->halted) {
        if (t == 0) S->skipped += 1;   // a step whose passes all returned at once
        return;
    }
    double a[6] = {0.0, 0.0, 0.0, -INFINITY, -INFINITY, -INFINITY};
    for (int b = t; b < nblocks; b += 1024) {
        for (int f = 0; f < 3; f++) a[f] += partials[(size_t)b * 6 + f];
        for (int f = 3; f < 6; f++) a[f] = fmax(a[f], partials[(size_t)b * 6 + f]);
    }
    for (int f = 0; f < 6; f++) red[f][t] = a[f];
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            for (int f = 0; f < 3; f++) red[f][t] += red[f][t + s];
            for (int f = 3; f < 6; f++) red[f][t] = fmax(red[f][t], red[f][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const double nn = (double)g.n;
        double m[3], mx[3];
        for (int f = 0; f < 3; f++) {
            m[f] = red[f][0] / nn / nn / nn;
            mx[f] = red[f + 3][0] - m[f];
        }
        const double umax = fmax(mx[2], fmax(mx[0], mx[1]));
        S->mean[0] = m[0]; S->mean[1] = m[1]; S->mean[2] = m[2];
        S->gumax = umax;
        S->cou = umax * g.dt * g.dxi;
        S->alpha = 1.5;
        S->beta = 0.5;
        S->step += 1;
        if (cfl_abort && S->cou > 7.0) S->halted = 1;
    }
}

static FinArgs fin_args(mhit_ctx* c) {
    FinArgs a;
    a.partials = c->partials;
    a.nblocks = c->npart_blocks;
    a.cfl_abort = c->P.cfl_abort;
    a.g = c->g;
    a.S = c->S;
    return a;
}

int launch_finalize(mhit_ctx* c) {
    k_finalize<<<1, 1024, 0, c->stream>>>(fin_args(c), nullptr);
    prof_mark(c, ST_FINALIZE);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

int launch_finalize_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    std::vector<FinArgs> t(nb);
    for (int b = 0; b < nb; b++) t[b] = fin_args(boxes[b]);
    const FinArgs* tab = arena.put(t.data(), nb);
    if (!tab) { set_error("batched launch: argument arena full"); return MHIT_ERR_STATE; }
    k_finalize<<<nb, 1024, 0, st>>>(t[0], tab);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

// ---------------------------------------------------------- diagnostics
// KE, dissipation 2 nu <S_ij S_ij> with central differences
// (post/compute_diss/main.f90:72-117, dx = lx/n, nu = mu/rho), max |D.u|
// with the divergence stencil of main.f90:841-843, phi mass and range.
__global__ __launch_bounds__(256)
void k_diag(const double* __restrict__ u, const double* __restrict__ v,
            const double* __restrict__ w, const double* __restrict__ phi,
            double* __restrict__ partials, Grid g, const DevScalars* __restrict__ S) {
    constexpr int NF = 8;
    __shared__ double red[NF][4];
    const int n = g.n;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    double acc[NF] = {0.0, 0.0, 0.0, 0.0, 0.0, INFINITY, -INFINITY, 0.0};
    if (i < n && j < n) {
        const double m0 = S->mean[0], m1 = S->mean[1], m2 = S->mean[2];
        const int ip = wrap_p(i, n), im = wrap_m(i, n), jp = wrap_p(j, n), jm = wrap_m(j, n);
        const double h = 0.5 * g.dxi;
        for (int k = blockIdx.z; k < n; k += gridDim.z) {
            const int kp = wrap_p(k, n), km = wrap_m(k, n);
            const size_t q = idx3(i, j, k, n);
#define U(a, b, c) (u[idx3(a, b, c, n)] - m0)
#define V(a, b, c) (v[idx3(a, b, c, n)] - m1)
#define W(a, b, c) (w[idx3(a, b, c, n)] - m2)
            const double uu = U(i, j, k), vv = V(i, j, k), ww = W(i, j, k);
            if (!isfinite(uu) || !isfinite(vv) || !isfinite(ww)) acc[7] += 1.0;
            acc[0] += uu * uu + vv * vv + ww * ww;
            const double dudx = (U(ip, j, k) - U(im, j, k)) * h, dudy = (U(i, jp, k) - U(i, jm, k)) * h;
            const double dudz = (U(i, j, kp) - U(i, j, km)) * h;
            const double dvdx = (V(ip, j, k) - V(im, j, k)) * h, dvdy = (V(i, jp, k) - V(i, jm, k)) * h;
            const double dvdz = (V(i, j, kp) - V(i, j, km)) * h;
            const double dwdx = (W(ip, j, k) - W(im, j, k)) * h, dwdy = (W(i, jp, k) - W(i, jm, k)) * h;
            const double dwdz = (W(i, j, kp) - W(i, j, km)) * h;
            const double S12 = 0.5 * (dudy + dvdx), S13 = 0.5 * (dudz + dwdx), S23 = 0.5 * (dvdz + dwdy);
            acc[1] += dudx * dudx + dvdy * dvdy + dwdz * dwdz + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23);
            const double dv = ((U(ip, j, k) - uu) + (V(i, jp, k) - vv) + (W(i, j, kp) - ww)) * g.dxi;
#undef U
#undef V
#undef W
            acc[2] = fmax(acc[2], fabs(dv));
            if (phi) {
                const double ph = phi[q];
                if (!isfinite(ph)) acc[7] += 1.0;
                acc[3] += ph;
                acc[5] = fmin(acc[5], ph);
                acc[6] = fmax(acc[6], ph);
            }
        }
    }
    for (int f = 0; f < NF; f++) {
        double x = acc[f];
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_xor(x, off, 64);
            x = (f == 2 || f == 6) ? fmax(x, y) : (f == 5 ? fmin(x, y) : x + y);
        }
        acc[f] = x;
    }
    if ((threadIdx.x & 63) == 0)
        for (int f = 0; f < NF; f++) red[f][threadIdx.y] = acc[f];
    __syncthreads();
    if (threadIdx.y == 0 && threadIdx.x < NF) {
        const int f = threadIdx.x;
        double r = red[f][0];
        for (int t = 1; t < 4; t++)
            r = (f == 2 || f == 6) ? fmax(r, red[f][t]) : (f == 5 ? fmin(r, red[f][t]) : r + red[f][t]);
        const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        partials[b * NF + f] = r;
    }
}

// Streaming copy at the practical HBM rate: one 16-byte vector per thread,
// nontemporal load and store, one pass over the buffer (measured 6.1 TB/s
// against 5.7 TB/s for four plain vectors per thread).
__global__ __launch_bounds__(256)
void k_copy_probe(const double2* __restrict__ src, double2* __restrict__ dst, size_t n2) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n2) st_once((MHIT_GLOBAL d2*)(dst + q), ld_once((const MHIT_GLOBAL d2*)(src + q)));
}

int launch_copy_probe(mhit_ctx* c, const void* src, void* dst, size_t bytes) {
    const size_t n2 = bytes / sizeof(double2);
    k_copy_probe<<<div_up((long long)n2, 256), 256, 0, c->stream>>>((const double2*)src, (double2*)dst, n2);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

// Pointwise streaming over R read and W write FP64 arrays (R, W <= 8): the
// rate a kernel with a given mix of concurrent streams can reach without
// any stencil, halo or barrier (a practical ceiling for the fused passes),
// with the nontemporal stores the passes use.
struct StreamProbe {
    const double* in[8];
    double* out[8];
};

__global__ __launch_bounds__(256)
void k_stream_probe(StreamProbe p, int R, int W, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < 8; r++)
            if (r < R) s += p.in[r][i];
#pragma unroll
        for (int w = 0; w < 8; w++)
            if (w < W) st_once(gmem(p.out[w]) + i, s + (double)w);
    }
}

int launch_stream_probe(mhit_ctx* c, const double* const* in, int R, double* const* out, int W, size_t n) {
    StreamProbe p{};
    for (int r = 0; r < R; r++) p.in[r] = in[r];
    for (int w = 0; w < W; w++) p.out[w] = out[w];
    const int blocks = (int)std::min<size_t>(32768, (n + 255) / 256);
    k_stream_probe<<<blocks, 256, 0, c->stream>>>(p, R, W, n);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

int launch_diag(mhit_ctx* c, double* out) {
    const int n = c->n;
    dim3 grid(div_up(n, 64), div_up(n, 4), kCorrZ < n ? kCorrZ : n), block(64, 4, 1);
    k_diag<<<grid, block, 0, c->stream>>>(c->u, c->v, c->w, c->P.phiflag ? c->phi : nullptr,
                                          c->dpartials, c->g, c->S);
    MHIT_HIP_TRY(hipGetLastError());
    const int nb = grid.x * grid.y * grid.z;
    std::vector<double> h((size_t)nb * 8);
    MHIT_HIP_TRY(hipMemcpyAsync(h.data(), c->dpartials, h.size() * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
    MHIT_HIP_TRY(hipStreamSynchronize(c->stream));
    double r[8] = {0.0, 0.0, 0.0, 0.0, 0.0, INFINITY, -INFINITY, 0.0};
    for (int b = 0; b < nb; b++)
        for (int f = 0; f < 8; f++) {
            const double x = h[(size_t)b * 8 + f];
            r[f] = (f == 2 || f == 6) ? std::fmax(r[f], x) : (f == 5 ? std::fmin(r[f], x) : r[f] + x);
        }
    for (int f = 0; f < 8; f++) out[f] = r[f];
    return MHIT_OK;
}

}  // namespace mhit
