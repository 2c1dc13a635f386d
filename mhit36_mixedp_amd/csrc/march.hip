// march.hip -- fused z-marching stencil passes (gfx950).
//
// A workgroup owns a tile of columns (MX x MY, one column per thread) and
// marches up a chunk of z planes.  Planes enter rings in LDS with their x/y
// halo; values needed only on the thread's own column stay in registers.
// Every array is then read from HBM once per pass (halo rows and the few
// planes before and after a chunk come mostly from L2/MALL), instead of
// once per neighbour as in a cell-per-thread kernel.
//
// NS pass: N1 (main.f90:658-717) + S1 (721-743) + S2 (760-780) or N2
//   (783-796) in one kernel: reads u, v, w, phi, rhs_o; writes u*, rhs_o.
// ACDI pass: P1-P5 (main.f90:522-630) in one kernel: reads phi, u, v, w,
//   rhsphi_o; writes phi', rhsphi_o.  psi = eps log(..) is computed once per
//   plane on the tile + 2-point halo, the normalised normals and the
//   sharpening fluxes F_d = 1/4 (1 - tanh^2) n_d once per point of the tile
//   + 1-point halo.
// Arithmetic follows the reference expressions term by term; only where
// each intermediate lives changes.
#include "internal.hpp"

#include <type_traits>
#include <vector>

namespace mhit {

__device__ __forceinline__ int wrapn(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}
__device__ __forceinline__ int ring3(int k) { return ((k % 3) + 3) % 3; }
__device__ __forceinline__ int ring4(int k) { return ((k % 4) + 4) % 4; }

// =====================================================================
// NS + CSF + AB2 + divergence of u*
// =====================================================================
// Besides u* and rhs_o, the pass writes the divergence of u* that opens
// the Poisson solve (main.f90:833-848), so that the x R2C pass reads one
// scalar field instead of u*, v*, w*.  The divergence needs u*(i+1),
// v*(j+1) and w*(k+1): u* and v* of the plane are exchanged through LDS,
// with the column x = MX and the row y = MY just past the tile evaluated by
// two waves (the same expressions as the neighbouring tile's own, so the
// same bits); each thread carries its partial sum to the next plane, where
// w*(k+1) is known, so a chunk marches one plane past its end, computing
// only w* there.  Threads of a partial tile past the domain edge compute
// the periodic image of their cell (their results feed the divergence of
// the last cells but are not stored).  The velocity planes carry a
// two-point halo on the +x/+y sides for these evaluations, and the fused
// pass reads its neighbours' previous rhs, so the new AB2 history goes to
// the other buffer of a ping-pong pair.
// Tile: TMX columns in x by TMY rows in y (Tuning::ns_my: 64 x 16, 64 x 8,
// 32 x 32 or 96 x 10), one thread per column; z chunk: Tuning::ns_mz.  The velocity planes carry x, y in
// [-1, M + 1], phi a two-point halo.

// Arguments of one box.  A launch covers one box (`one`) or, for an
// ensemble, a table of boxes of the same size stacked along the grid's z.
template <typename TD>
struct NsArgs {
    const double *u, *v, *w, *phi, *ru_o, *rv_o, *rw_o;
    double *ru_n, *rv_n, *rw_n, *us, *vs, *ws;
    TD* div;
    double cdiv;
    const double *mysin, *mycos;
    Grid g;
    const DevScalars* S;
    int zc, xcd;
};

template <bool PHI, typename TD, int TMY, int TMX>
__global__ __launch_bounds__(TMX * TMY)
void k_ns_march(NsArgs<TD> one, const NsArgs<TD>* __restrict__ tab, int nbox) {
    const Blk blk = launch_blk(nbox);
    NsArgs<TD> A = one;
    if (tab) A = tab[blk.box];
    const MHIT_GLOBAL double* __restrict__ u = gmem(A.u);
    const MHIT_GLOBAL double* __restrict__ v = gmem(A.v);
    const MHIT_GLOBAL double* __restrict__ w = gmem(A.w);
    const MHIT_GLOBAL double* __restrict__ phi = gmem(A.phi);
    const MHIT_GLOBAL double* __restrict__ ru_o = gmem(A.ru_o);
    const MHIT_GLOBAL double* __restrict__ rv_o = gmem(A.rv_o);
    const MHIT_GLOBAL double* __restrict__ rw_o = gmem(A.rw_o);
    MHIT_GLOBAL double* __restrict__ ru_n = gmem(A.ru_n);
    MHIT_GLOBAL double* __restrict__ rv_n = gmem(A.rv_n);
    MHIT_GLOBAL double* __restrict__ rw_n = gmem(A.rw_n);
    MHIT_GLOBAL double* __restrict__ us = gmem(A.us);
    MHIT_GLOBAL double* __restrict__ vs = gmem(A.vs);
    MHIT_GLOBAL double* __restrict__ ws = gmem(A.ws);
    MHIT_GLOBAL TD* __restrict__ div = gmem(A.div);
    const double cdiv = A.cdiv;
    const MHIT_GLOBAL double* __restrict__ mysin = gmem(A.mysin);
    const MHIT_GLOBAL double* __restrict__ mycos = gmem(A.mycos);
    const Grid g = A.g;
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    const int zc = A.zc, xcd = A.xcd;
    constexpr int MX = TMX, MY = TMY, W1 = MX + 3, H1 = MY + 3, W2 = MX + 4, H2 = MY + 4;
    constexpr int NT = MX * MY;
    constexpr int FXW = MX + 2;                          // fx at x = -1..MX
    __shared__ double su[3][H1 * W1], sv[3][H1 * W1], sw[3][H1 * W1];
    __shared__ double sp[PHI ? 3 : 1][PHI ? H2 * W2 : 1];
    __shared__ double sfx[PHI ? MY * FXW : 1];
    __shared__ double sfy[PHI ? (MY + 2) * MX : 1];     // fy at y = -1..MY
    __shared__ double sus[MY * (MX + 1)];               // u* at x = 0..MX
    __shared__ double svs[(MY + 1) * MX];               // v* at y = 0..MY

    const int n = g.n;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * MX + tx;
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz, blk, xcd);
    const int i0 = tbx * MX, j0 = tby * MY;
    const int z0 = tbz * zc;
    const int z1 = min(z0 + zc, n);
    const int i = i0 + tx, j = j0 + ty;
    const bool live = i < n && j < n;
    // A stopped run writes nothing; its AB2 history stays in the buffer it
    // was in, and the host undoes an odd number of buffer swaps afterwards
    // (the finalize pass counts the skipped steps).  No copy here: any code
    // in this branch costs the marching loop registers.
    if (S->halted) return;
    // global indices of the cell, or of its periodic image past the domain edge
    const int ic = wrapn(i, n), jc = wrapn(j, n);
    const double m0 = S->mean[0], m1 = S->mean[1], m2 = S->mean[2];
    const double alpha = S->alpha, beta = S->beta;
    const double dxi = g.dxi, ddxi = g.ddxi, mu = g.mu, rhoi = g.rhoi, dt = g.dt;
    const double eps = g.eps, epsi = g.epsi, sig6 = 6.0 * g.sigma;

    // one plane in flight: global -> registers (fetch), registers -> ring (commit)
    constexpr int NV = (H1 * W1 + NT - 1) / NT, NPH = (H2 * W2 + NT - 1) / NT;
    const size_t plane = (size_t)n * n;
    // in-plane offsets of the thread's halo points, fixed for the whole march
    int voff[NV], poff[PHI ? NPH : 1];
#pragma unroll
    for (int t = 0; t < NV; t++) {
        const int e = tid + t * NT;
        const int lx = e % W1, ly = e / W1;
        voff[t] = (e < H1 * W1) ? wrapn(j0 - 1 + ly, n) * n + wrapn(i0 - 1 + lx, n) : 0;
    }
    if (PHI) {
#pragma unroll
        for (int t = 0; t < NPH; t++) {
            const int e = tid + t * NT;
            const int lx = e % W2, ly = e / W2;
            poff[t] = (e < H2 * W2) ? wrapn(j0 - 2 + ly, n) * n + wrapn(i0 - 2 + lx, n) : 0;
        }
    }
    double pu[NV], pv[NV], pw[NV], pp[NPH];
    auto fetch = [&](int k) {
        const size_t base = (size_t)wrapn(k, n) * plane;
#pragma unroll
        for (int t = 0; t < NV; t++) {
            if ((H1 * W1) % NT == 0 || tid + t * NT < H1 * W1) {
                const size_t q = base + voff[t];
                pu[t] = u[q]; pv[t] = v[q]; pw[t] = w[q];
            }
        }
        if (PHI) {
#pragma unroll
            for (int t = 0; t < NPH; t++)
                if ((H2 * W2) % NT == 0 || tid + t * NT < H2 * W2) pp[t] = phi[base + poff[t]];
        }
    };
    auto commit = [&](int k) {
        const int s = ring3(k);
#pragma unroll
        for (int t = 0; t < NV; t++) {
            const int e = tid + t * NT;
            if (e < H1 * W1) { su[s][e] = pu[t] - m0; sv[s][e] = pv[t] - m1; sw[s][e] = pw[t] - m2; }
        }
        if (PHI) {
#pragma unroll
            for (int t = 0; t < NPH; t++) {
                const int e = tid + t * NT;
                if (e < H2 * W2) sp[s][e] = pp[t];
            }
        }
    };
    // phi at tile coordinates (a, b) of plane k (a, b from -2)
    auto P = [&](int k, int a, int b) -> double { return sp[ring3(k)][(b + 2) * W2 + (a + 2)]; };
    // chemical potential at (a, b, k), main.f90:735
    auto chem = [&](int k, int a, int b) -> double {
        const double ph = P(k, a, b);
        return ph * (1.0 - ph) * (1.0 - 2.0 * ph) * epsi -
               eps * (P(k, a + 1, b) + P(k, a - 1, b) + P(k, a, b + 1) + P(k, a, b - 1) +
                      P(k + 1, a, b) + P(k - 1, a, b) - 6.0 * ph) * ddxi;
    };
    // velocity ring entry at tile coordinates (a, b), a, b from -1
    auto at = [&](int a, int b) { return (b + 1) * W1 + (a + 1); };

    // x-momentum RHS at tile (a, b) of plane k with global row jg
    // (main.f90:664-668, 691-693, 711, 760-764), in the reference's order
    auto rhs_u = [&](int sm, int sc, int sq, int a, int b, int jg, int k) -> double {
        const int c = at(a, b);
        const double u0 = su[sc][c], uxp = su[sc][c + 1], uxm = su[sc][c - 1];
        const double uyp = su[sc][c + W1], uym = su[sc][c - W1], uzp = su[sq][c], uzm = su[sm][c];
        double h11 = (uxp + u0) * (uxp + u0) - (u0 + uxm) * (u0 + uxm);
        double h12 = (uyp + u0) * (sv[sc][c + W1] + sv[sc][c + W1 - 1]) - (u0 + uym) * (sv[sc][c] + sv[sc][c - 1]);
        double h13 = (uzp + u0) * (sw[sq][c] + sw[sq][c - 1]) - (u0 + uzm) * (sw[sc][c] + sw[sc][c - 1]);
        h11 = h11 * 0.25 * dxi; h12 = h12 * 0.25 * dxi; h13 = h13 * 0.25 * dxi;
        double ru = -(h11 + h12 + h13);
        h11 = mu * (uxp - 2.0 * u0 + uxm) * ddxi;
        h12 = mu * (uyp - 2.0 * u0 + uym) * ddxi;
        h13 = mu * (uzp - 2.0 * u0 + uzm) * ddxi;
        ru = ru + (h11 + h12 + h13) * rhoi;
        ru = ru + g.f3 * mysin[k + 1] + g.f2 * mycos[jg + 1];
        if (PHI) ru = ru + 0.5 * (sfx[b * FXW + a] + sfx[b * FXW + a + 1]) * rhoi;
        return ru;
    };
    // y-momentum RHS at tile (a, b) of plane k with global column ig
    auto rhs_v = [&](int sm, int sc, int sq, int a, int b, int ig, int k) -> double {
        const int c = at(a, b);
        const double v0 = sv[sc][c], vxp = sv[sc][c + 1], vxm = sv[sc][c - 1];
        const double vyp = sv[sc][c + W1], vym = sv[sc][c - W1], vzp = sv[sq][c], vzm = sv[sm][c];
        const double u0 = su[sc][c], uym = su[sc][c - W1];
        double h21 = (su[sc][c + 1] + su[sc][c + 1 - W1]) * (vxp + v0) - (u0 + uym) * (v0 + vxm);
        double h22 = (vyp + v0) * (vyp + v0) - (v0 + vym) * (v0 + vym);
        double h23 = (sw[sq][c] + sw[sq][c - W1]) * (vzp + v0) - (sw[sc][c] + sw[sc][c - W1]) * (v0 + vzm);
        h21 = h21 * 0.25 * dxi; h22 = h22 * 0.25 * dxi; h23 = h23 * 0.25 * dxi;
        double rv = -(h21 + h22 + h23);
        h21 = mu * (vxp - 2.0 * v0 + vxm) * ddxi;
        h22 = mu * (vyp - 2.0 * v0 + vym) * ddxi;
        h23 = mu * (vzp - 2.0 * v0 + vzm) * ddxi;
        rv = rv + (h21 + h22 + h23) * rhoi;
        rv = rv + g.f1 * mysin[ig] + g.f3 * mycos[k + 1];
        if (PHI) rv = rv + 0.5 * (sfy[b * MX + a] + sfy[(b + 1) * MX + a]) * rhoi;
        return rv;
    };
    // z-momentum RHS at the thread's own column of plane k
    auto rhs_w = [&](int sm, int sc, int sq, double fzm, double fzc) -> double {
        const int c = at(tx, ty);
        const double w0 = sw[sc][c], wxp = sw[sc][c + 1], wxm = sw[sc][c - 1];
        const double wyp = sw[sc][c + W1], wym = sw[sc][c - W1], wzp = sw[sq][c], wzm = sw[sm][c];
        const double u0 = su[sc][c], v0 = sv[sc][c];
        double h31 = (wxp + w0) * (su[sc][c + 1] + su[sm][c + 1]) - (w0 + wxm) * (u0 + su[sm][c]);
        double h32 = (sv[sc][c + W1] + sv[sm][c + W1]) * (wyp + w0) - (v0 + sv[sm][c]) * (w0 + wym);
        double h33 = (wzp + w0) * (wzp + w0) - (w0 + wzm) * (w0 + wzm);
        h31 = h31 * 0.25 * dxi; h32 = h32 * 0.25 * dxi; h33 = h33 * 0.25 * dxi;
        double rw = -(h31 + h32 + h33);
        h31 = mu * (wxp - 2.0 * w0 + wxm) * ddxi;
        h32 = mu * (wyp - 2.0 * w0 + wym) * ddxi;
        h33 = mu * (wzp - 2.0 * w0 + wzm) * ddxi;
        rw = rw + (h31 + h32 + h33) * rhoi;
        rw = rw + g.f2 * mysin[jc + 1] + g.f1 * mycos[ic];
        if (PHI) rw = rw + 0.5 * (fzm + fzc) * rhoi;
        return rw;
    };

    double fz_prev = 0.0, fz_cur = 0.0;
    // warm-up: planes z0-1, z0 of u,v,w; phi planes z0-2..z0 and fz(z0-1)
    if (PHI) {
        fetch(z0 - 2);
        commit(z0 - 2);   // also fills the velocity slot of z0-2 = z0+1, rewritten below
    }
    fetch(z0 - 1);
    commit(z0 - 1);
    fetch(z0);
    commit(z0);
    lds_sync();
    if (PHI) fz_prev = sig6 * chem(z0 - 1, tx, ty) * 0.5 * (P(z0, tx, ty) - P(z0 - 2, tx, ty)) * dxi;
    lds_sync();
    fetch(z0 + 1);

    double dprev = 0.0, wsprev = 0.0;   // divergence partial and w* of the previous plane
    for (int k = z0; k <= z1; k++) {
        const bool last = k == z1;      // one plane past the chunk: w* only
        const int kg = last ? wrapn(k, n) : k;
        // AB2 history of the own column: issued before the barriers
        const size_t q = idx3(ic, jc, kg, n);
        const double row = rw_o[q];
        double rou = 0.0, rov = 0.0;
        if (!last) { rou = ru_o[q]; rov = rv_o[q]; }
        // the cell just past the tile, for the edge waves: x = MX (wave 0),
        // y = MY (wave 1)
        const bool xedge = tid < MY, yedge = tid >= 64 && tid < 64 + MX;
        const int eb = tid, ea = tid - 64;
        commit(k + 1);
        lds_sync();
        if (PHI) {
            // S1 at plane k: fx on x = -1..MX, fy on y = -1..MY, fz on the column
            const double c0 = chem(k, tx, ty);
            fz_cur = sig6 * c0 * 0.5 * (P(k + 1, tx, ty) - P(k - 1, tx, ty)) * dxi;
            if (!last) {
                sfx[ty * FXW + tx + 1] = sig6 * c0 * 0.5 * (P(k, tx + 1, ty) - P(k, tx - 1, ty)) * dxi;
                sfy[(ty + 1) * MX + tx] = sig6 * c0 * 0.5 * (P(k, tx, ty + 1) - P(k, tx, ty - 1)) * dxi;
                // one edge point per thread of waves 0-2 (no wave takes
                // both branches): y = -1 and y = MY rows on threads
                // [0, 2 MX), x = -1 and x = MX columns on [2 MX, 2 MX + 2 MY)
                if (tid < 2 * MX) {
                    const int a = tid >> 1, b = (tid & 1) ? MY : -1;
                    sfy[(b + 1) * MX + a] = sig6 * chem(k, a, b) * 0.5 * (P(k, a, b + 1) - P(k, a, b - 1)) * dxi;
                } else if (tid < 2 * MX + 2 * MY) {
                    const int e = tid - 2 * MX;
                    const int b = e >> 1, a = (e & 1) ? MX : -1;
                    sfx[b * FXW + a + 1] = sig6 * chem(k, a, b) * 0.5 * (P(k, a + 1, b) - P(k, a - 1, b)) * dxi;
                }
            }
            lds_sync();
        }
        if (k + 2 <= z1 + 1) fetch(k + 2);   // next plane in flight during the compute
        const int sm = ring3(k - 1), sc = ring3(k), sq = ring3(k + 1);
        const int c = at(tx, ty);
        const double rw = rhs_w(sm, sc, sq, fz_prev, fz_cur);
        const double wsk = sw[sc][c] + dt * (alpha * rw - beta * row);
        if (live && k > z0) {
            // main.f90:841-844: rhsp = rho/(dt dx) sum_d (u*_d(+1) - u*_d)
            const double r = dprev + cdiv * (wsk - wsprev);
            st_once(&div[idx3(i, j, k - 1, n)], (TD)r);
        }
        double usk = 0.0, vsk = 0.0;
        if (!last) {
            if (live) { st_once(&ws[q], wsk); st_once(&rw_n[q], rw); }
            const double ru = rhs_u(sm, sc, sq, tx, ty, jc, k);
            usk = su[sc][c] + dt * (alpha * ru - beta * rou);
            if (live) { st_once(&us[q], usk); st_once(&ru_n[q], ru); }
            const double rv = rhs_v(sm, sc, sq, tx, ty, ic, k);
            vsk = sv[sc][c] + dt * (alpha * rv - beta * rov);
            if (live) { st_once(&vs[q], vsk); st_once(&rv_n[q], rv); }
            sus[ty * (MX + 1) + tx] = usk;
            svs[ty * MX + tx] = vsk;
            if (xedge) {
                const double roe = ru_o[idx3(wrapn(i0 + MX, n), wrapn(j0 + eb, n), k, n)];
                const double re = rhs_u(sm, sc, sq, MX, eb, wrapn(j0 + eb, n), k);
                sus[eb * (MX + 1) + MX] = su[sc][at(MX, eb)] + dt * (alpha * re - beta * roe);
            } else if (yedge) {
                const double roe = rv_o[idx3(wrapn(i0 + ea, n), wrapn(j0 + MY, n), k, n)];
                const double re = rhs_v(sm, sc, sq, ea, MY, wrapn(i0 + ea, n), k);
                svs[MY * MX + ea] = sv[sc][at(ea, MY)] + dt * (alpha * re - beta * roe);
            }
        }
        wsprev = wsk;
        fz_prev = fz_cur;
        lds_sync();
        if (!last) {
            double r = cdiv * (sus[ty * (MX + 1) + tx + 1] - usk);
            r = r + cdiv * (svs[(ty + 1) * MX + tx] - vsk);
            dprev = r;
        }
    }
}

// =====================================================================
// ACDI
// =====================================================================
// Tile: TMX columns by TMY rows (Tuning::ac_my: 64 x 16, 64 x 8, 96 x 10 or
// 32 x 32); z chunk: Tuning::ac_mz.  phi ring with a one-point halo, psi
// ring with two.

struct AcdiArgs {
    double* phi_out;
    const double *phi, *u, *v, *w;
    double* rphi_o;
    Grid g;
    const DevScalars* S;
    int zc, xcd;
};

template <int TMY, int TMX>
__global__ __launch_bounds__(TMX * TMY)
void k_acdi_march(AcdiArgs one, const AcdiArgs* __restrict__ tab, int nbox) {
    const Blk blk = launch_blk(nbox);
    AcdiArgs A = one;
    if (tab) A = tab[blk.box];
    MHIT_GLOBAL double* __restrict__ phi_out = gmem(A.phi_out);
    const MHIT_GLOBAL double* __restrict__ phi = gmem(A.phi);
    const MHIT_GLOBAL double* __restrict__ u = gmem(A.u);
    const MHIT_GLOBAL double* __restrict__ v = gmem(A.v);
    const MHIT_GLOBAL double* __restrict__ w = gmem(A.w);
    MHIT_GLOBAL double* __restrict__ rphi_o = gmem(A.rphi_o);
    const Grid g = A.g;
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    const int zc = A.zc, xcd = A.xcd;
    constexpr int MX = TMX, MY = TMY, W1 = MX + 2, H1 = MY + 2, W2 = MX + 4, H2 = MY + 4;
    constexpr int NT = MX * MY;
    __shared__ double spsi[3][H2 * W2];
    __shared__ double sphi[4][H1 * W1];
    __shared__ double sFx[2][MY * (MX + 2)];   // x = -1..MX
    __shared__ double sFy[2][(MY + 2) * MX];   // y = -1..MY

    const int n = g.n;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * MX + tx;
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz, blk, xcd);
    const int i0 = tbx * MX, j0 = tby * MY;
    const int z0 = tbz * zc;
    const int z1 = min(z0 + zc, n);
    const int i = i0 + tx, j = j0 + ty;
    const bool live = i < n && j < n;
    if (S->halted) {
        // a stopped run keeps its state: phi_out becomes a copy of phi, so
        // the host-side ping-pong of the two buffers stays consistent
        if (live)
            for (int k = z0; k < z1; k++) phi_out[idx3(i, j, k, n)] = phi[idx3(i, j, k, n)];
        return;
    }
    const double m0 = S->mean[0], m1 = S->mean[1], m2 = S->mean[2];
    const double alpha = S->alpha, beta = S->beta;
    const double gamma = 1.0 * S->gumax;               // main.f90:536
    const double dxi = g.dxi, ddxi = g.ddxi, eps = g.eps, enm = g.enm, dt = g.dt;
    const double hdx = 0.5 * dxi, ced = eps * ddxi;

    // phi of one plane in flight: global -> registers
    constexpr int NPL = (H2 * W2 + NT - 1) / NT;
    const size_t plane = (size_t)n * n;
    // in-plane offsets of the thread's halo points, fixed for the whole march
    int poff[NPL];
#pragma unroll
    for (int t = 0; t < NPL; t++) {
        const int e = tid + t * NT;
        const int lx = e % W2, ly = e / W2;
        poff[t] = (e < H2 * W2) ? wrapn(j0 - 2 + ly, n) * n + wrapn(i0 - 2 + lx, n) : 0;
    }
    double pf[NPL];
    auto fetch = [&](int k) {
        const MHIT_GLOBAL double* __restrict__ pl = phi + (size_t)wrapn(k, n) * plane;
#pragma unroll
        for (int t = 0; t < NPL; t++)
            if ((H2 * W2) % NT == 0 || tid + t * NT < H2 * W2) pf[t] = pl[poff[t]];
    };
    // registers -> psi ring (halo 2) and phi ring (halo 1), main.f90:522-534
    auto commit = [&](int k) {
        const int sps = ring3(k), sph = ring4(k);
#pragma unroll
        for (int t = 0; t < NPL; t++) {
            const int e = tid + t * NT;
            if (e >= H2 * W2) break;
            const int lx = e % W2, ly = e / W2;
            const double ph = pf[t];
            const double val = fmin(ph, 1.0);
            spsi[sps][e] = eps * ln_ratio(val + enm, 1.0 - val + enm);
            if (lx >= 1 && lx <= MX + 2 && ly >= 1 && ly <= MY + 2)
                sphi[sph][(ly - 1) * W1 + (lx - 1)] = ph;
        }
    };
    auto PS = [&](int k, int a, int b) -> double { return spsi[ring3(k)][(b + 2) * W2 + (a + 2)]; };
    auto PH = [&](int k, int a, int b) -> double { return sphi[ring4(k)][(b + 1) * W1 + (a + 1)]; };
    // normalised normal and 1/4 (1 - tanh^2(psi/(2 eps))) at (a, b, k):
    // main.f90:530, 559-561, 586-589.  With r = (v+e)/(1-v+e) and
    // psi = eps ln r, tanh(ln(r)/2) = (r-1)/(r+1) = (2v-1)/(1+2e), hence
    // 1/4 (1 - tanh^2) = (v+e)(1-v+e)/(1+2e)^2: the same quantity without a
    // second transcendental, and without the cancellation of 1 - T*T where
    // phi is near 0 or 1.
    const double g_scale = 1.0 / ((1.0 + 2.0 * enm) * (1.0 + 2.0 * enm));
    auto flux = [&](int k, int a, int b, double& fx, double& fy, double& fz) {
        const double nx = (PS(k, a + 1, b) - PS(k, a - 1, b));
        const double ny = (PS(k, a, b + 1) - PS(k, a, b - 1));
        const double nz = (PS(k + 1, a, b) - PS(k - 1, a, b));
        // 1/(|grad psi| + 1e-16), main.f90:559-561: |grad psi| as s/sqrt(s)
        // with the refined hardware reciprocal square root, then the
        // refined hardware reciprocal (each within 1 ulp)
        const double s2 = fma(nx, nx, fma(ny, ny, nz * nz));
        const double normag = rcp_refined((s2 > 0.0 ? s2 * rsq_refined(s2) : 0.0) + enm);
        const double val = fmin(PH(k, a, b), 1.0);
        const double Gn = (val + enm) * (1.0 - val + enm) * g_scale * normag;
        fx = Gn * nx;
        fy = Gn * ny;
        fz = Gn * nz;
    };
    // sharpening fluxes of plane k into buffer bf (tile + 1-point halo in x / y);
    // returns Fz at the thread's column
    auto fluxes = [&](int k, int bf) -> double {
        double fx, fy, fz;
        flux(k, tx, ty, fx, fy, fz);
        sFx[bf][ty * (MX + 2) + tx + 1] = fx;
        sFy[bf][(ty + 1) * MX + tx] = fy;
        // one edge point per thread, on waves that do not take the second
        // psi point of the commit (threads [0, 336) at 64 x 16), so that
        // the extra work spreads over the SIMDs: y = -1 and y = MY rows on
        // threads [EY, EY + 2 MX), x = -1 and x = MX columns on
        // [EX, EX + 2 MY)
        constexpr bool SPREAD = MX == 64 && MY == 16;
        constexpr int EY = SPREAD ? 384 : 0, EX = SPREAD ? 640 : 2 * MX;
        if (tid >= EY && tid < EY + 2 * MX) {
            const int e = tid - EY;
            const int a = e >> 1, b = (e & 1) ? MY : -1;
            double gx, gy, gz;
            flux(k, a, b, gx, gy, gz);
            sFy[bf][(b + 1) * MX + a] = gy;
        } else if (tid >= EX && tid < EX + 2 * MY) {
            const int e = tid - EX;
            const int b = e >> 1, a = (e & 1) ? MX : -1;
            double gx, gy, gz;
            flux(k, a, b, gx, gy, gz);
            sFx[bf][b * (MX + 2) + a + 1] = gx;
        }
        return fz;
    };

    // warm-up: psi planes z0-2..z0 -> Fz(z0-1); plane z0+1 -> fluxes of z0
    for (int k = z0 - 2; k <= z0; k++) {
        fetch(k);
        commit(k);
    }
    lds_sync();
    double fz_m1 = 0.0, fz_0 = 0.0, fz_p1 = 0.0;
    {
        double fx, fy;
        flux(z0 - 1, tx, ty, fx, fy, fz_m1);
    }
    lds_sync();
    fetch(z0 + 1);
    commit(z0 + 1);
    lds_sync();
    int bf = 0;
    fz_0 = fluxes(z0, bf);
    fetch(z0 + 2);
    lds_sync();

    double wk = 0.0;
    if (live) wk = w[idx3(i, j, z0, n)] - m2;
    for (int k = z0; k < z1; k++) {
        // velocity and AB2 history of the own column: issued before the barriers
        const size_t q = idx3(live ? i : 0, live ? j : 0, k, n);
        double uc = 0.0, up = 0.0, vc = 0.0, vp = 0.0, wp = 0.0, ro = 0.0;
        if (live) {
            uc = u[q] - m0;
            up = u[idx3(wrap_p(i, n), j, k, n)] - m0;
            vc = v[q] - m1;
            vp = v[idx3(i, wrap_p(j, n), k, n)] - m1;
            wp = w[idx3(i, j, wrap_p(k, n), n)] - m2;
            ro = rphi_o[q];
        }
        commit(k + 2);
        lds_sync();
        if (k + 3 <= z1 + 1) fetch(k + 3);   // next plane in flight during the compute
        fz_p1 = fluxes(k + 1, bf ^ 1);
        if (live) {
            const double pc = PH(k, tx, ty);
            const double pxp = PH(k, tx + 1, ty), pxm = PH(k, tx - 1, ty);
            const double pyp = PH(k, tx, ty + 1), pym = PH(k, tx, ty - 1);
            const double pzp = PH(k + 1, tx, ty), pzm = PH(k - 1, tx, ty);
            // P2 conv + diff (main.f90:551-557) and P4 sharpening
            // (609-614) with the common factors taken out of the sums:
            // 0.5/dx of the face fluxes and of the sharpening divergence,
            // eps/dx^2 of the diffusion.  Each face flux is still formed
            // from the same operands by the two cells that share it;
            // the sum differs from the term-by-term form by a few ulp.
            const double* Fx = sFx[bf];
            const double* Fy = sFy[bf];
            const double conv = ((up * (pxp + pc) - uc * (pc + pxm)) + (vp * (pyp + pc) - vc * (pc + pym))) +
                                (wp * (pzp + pc) - wk * (pc + pzm));
            const double lap = ((pxp + pxm) + (pyp + pym)) + (pzp + pzm) - 6.0 * pc;
            const double shp = ((Fx[ty * (MX + 2) + tx + 2] - Fx[ty * (MX + 2) + tx]) +
                                (Fy[(ty + 2) * MX + tx] - Fy[ty * MX + tx])) + (fz_p1 - fz_m1);
            const double r = gamma * (ced * lap - hdx * shp) - hdx * conv;
            // P5 AB2, main.f90:625-626
            st_once(&phi_out[q], pc + dt * (alpha * r - beta * ro));
            st_once(&rphi_o[q], r);
            wk = wp;
        }
        fz_m1 = fz_0;
        fz_0 = fz_p1;
        bf ^= 1;
        lds_sync();
    }
}

// ------------------------------------------------------------- launchers
template <typename TD>
static NsArgs<TD> ns_args(mhit_ctx* c, TD* div) {
    NsArgs<TD> a;
    a.u = c->u; a.v = c->v; a.w = c->w;
    a.phi = c->P.phiflag ? c->phi : nullptr;
    a.ru_o = c->rhsu_o; a.rv_o = c->rhsv_o; a.rw_o = c->rhsw_o;
    a.ru_n = c->rhsu_n; a.rv_n = c->rhsv_n; a.rw_n = c->rhsw_n;
    a.us = c->us; a.vs = c->vs; a.ws = c->ws;
    a.div = div;
    a.cdiv = c->g.rho * c->g.dxi / c->g.dt;
    a.mysin = c->mysin; a.mycos = c->mycos;
    a.g = c->g;
    a.S = c->S;
    a.zc = c->n < c->tune.ns_mz ? c->n : c->tune.ns_mz;
    a.xcd = c->tune.xcd;
    return a;
}

static AcdiArgs acdi_args(mhit_ctx* c) {
    AcdiArgs a;
    a.phi_out = c->phi_next;
    a.phi = c->phi;
    a.u = c->u; a.v = c->v; a.w = c->w;
    a.rphi_o = c->rhsphi_o;
    a.g = c->g;
    a.S = c->S;
    a.zc = c->n < c->tune.ac_mz ? c->n : c->tune.ac_mz;
    a.xcd = c->tune.xcd;
    return a;
}

// the new AB2 history is in the other buffers (neighbours' old values are
// read during the pass)
void ns_swap(mhit_ctx* c) {
    std::swap(c->rhsu_o, c->rhsu_n);
    std::swap(c->rhsv_o, c->rhsv_n);
    std::swap(c->rhsw_o, c->rhsw_n);
}

// NS pass over boxes[0..nb) of one size, phase flag, Poisson precision and
// z chunk, one launch; tab is their argument table (nullptr: one box).
template <bool PH, typename TD>
static void go_ns(mhit_ctx* c, const NsArgs<TD>& one, const NsArgs<TD>* tab, int nb, hipStream_t st) {
    const int n = c->n;
    auto go = [&](auto my, auto mx) {
        constexpr int MY = decltype(my)::value, MX = decltype(mx)::value;
        dim3 grid(div_up(n, MX), div_up(n, MY), div_up(n, one.zc) * nb), block(MX, MY, 1);
        k_ns_march<PH, TD, MY, MX><<<grid, block, 0, st>>>(one, tab, nb);
    };
    if (c->tune.ns_my == 8) go(std::integral_constant<int, 8>{}, std::integral_constant<int, 64>{});
    else if (c->tune.ns_my == 32) go(std::integral_constant<int, 32>{}, std::integral_constant<int, 32>{});
    else if (c->tune.ns_my == 10) go(std::integral_constant<int, 10>{}, std::integral_constant<int, 96>{});
    else go(std::integral_constant<int, 16>{}, std::integral_constant<int, 64>{});
}

// The divergence goes to the pressure buffer, which the x R2C pass reads
// and the x C2R pass overwrites with p.
int launch_ns_march(mhit_ctx* c) {
    if (c->P.phiflag) {
        if (c->fp32) go_ns<true>(c, ns_args(c, (float*)c->pres), (const NsArgs<float>*)nullptr, 1, c->stream);
        else go_ns<true>(c, ns_args(c, (double*)c->pres), (const NsArgs<double>*)nullptr, 1, c->stream);
    } else {
        if (c->fp32) go_ns<false>(c, ns_args(c, (float*)c->pres), (const NsArgs<float>*)nullptr, 1, c->stream);
        else go_ns<false>(c, ns_args(c, (double*)c->pres), (const NsArgs<double>*)nullptr, 1, c->stream);
    }
    prof_mark(c, ST_NS);
    MHIT_HIP_TRY(hipGetLastError());
    ns_swap(c);
    return MHIT_OK;
}

template <bool PH, typename TD>
static int ns_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    std::vector<NsArgs<TD>> t(nb);
    for (int b = 0; b < nb; b++) t[b] = ns_args(boxes[b], (TD*)boxes[b]->pres);
    const NsArgs<TD>* tab = arena.put(t.data(), nb);
    if (!tab) { set_error("batched launch: argument arena full"); return MHIT_ERR_STATE; }
    go_ns<PH, TD>(boxes[0], t[0], tab, nb, st);
    MHIT_HIP_TRY(hipGetLastError());
    for (int b = 0; b < nb; b++) ns_swap(boxes[b]);
    return MHIT_OK;
}

int launch_ns_march_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    mhit_ctx* c = boxes[0];
    for (int b = 1; b < nb; b++)
        if (boxes[b]->n != c->n || boxes[b]->P.phiflag != c->P.phiflag || boxes[b]->fp32 != c->fp32 ||
            boxes[b]->tune.ns_mz != c->tune.ns_mz || boxes[b]->tune.ns_my != c->tune.ns_my ||
            boxes[b]->tune.xcd != c->tune.xcd) {
            set_error("batched NS pass: boxes differ in size, phase flag, precision or launch shape");
            return MHIT_ERR_ARG;
        }
    if (c->P.phiflag) return c->fp32 ? ns_batch<true, float>(boxes, nb, arena, st)
                                     : ns_batch<true, double>(boxes, nb, arena, st);
    return c->fp32 ? ns_batch<false, float>(boxes, nb, arena, st)
                   : ns_batch<false, double>(boxes, nb, arena, st);
}

static void go_acdi(mhit_ctx* c, const AcdiArgs& one, const AcdiArgs* tab, int nb, hipStream_t st) {
    const int n = c->n;
    auto go = [&](auto my, auto mx) {
        constexpr int MY = decltype(my)::value, MX = decltype(mx)::value;
        dim3 grid(div_up(n, MX), div_up(n, MY), div_up(n, one.zc) * nb), block(MX, MY, 1);
        k_acdi_march<MY, MX><<<grid, block, 0, st>>>(one, tab, nb);
    };
    if (c->tune.ac_my == 8) go(std::integral_constant<int, 8>{}, std::integral_constant<int, 64>{});
    else if (c->tune.ac_my == 32) go(std::integral_constant<int, 32>{}, std::integral_constant<int, 32>{});
    else if (c->tune.ac_my == 10) go(std::integral_constant<int, 10>{}, std::integral_constant<int, 96>{});
    else go(std::integral_constant<int, 16>{}, std::integral_constant<int, 64>{});
}

int launch_acdi_march(mhit_ctx* c) {
    go_acdi(c, acdi_args(c), nullptr, 1, c->stream);
    prof_mark(c, ST_ACDI);
    MHIT_HIP_TRY(hipGetLastError());
    std::swap(c->phi, c->phi_next);
    return MHIT_OK;
}

int launch_acdi_march_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    mhit_ctx* c = boxes[0];
    std::vector<AcdiArgs> t(nb);
    for (int b = 0; b < nb; b++) {
        if (boxes[b]->n != c->n || !boxes[b]->P.phiflag || boxes[b]->tune.ac_mz != c->tune.ac_mz ||
            boxes[b]->tune.ac_my != c->tune.ac_my || boxes[b]->tune.xcd != c->tune.xcd) {
            set_error("batched ACDI pass: boxes differ in size or launch shape, or have no phase field");
            return MHIT_ERR_ARG;
        }
        t[b] = acdi_args(boxes[b]);
    }
    const AcdiArgs* tab = arena.put(t.data(), nb);
    if (!tab) { set_error("batched launch: argument arena full"); return MHIT_ERR_STATE; }
    go_acdi(c, t[0], tab, nb, st);
    MHIT_HIP_TRY(hipGetLastError());
    for (int b = 0; b < nb; b++) std::swap(boxes[b]->phi, boxes[b]->phi_next);
    return MHIT_OK;
}

}  // namespace mhit
