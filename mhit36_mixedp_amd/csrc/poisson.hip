// poisson.hip -- FFT Poisson projection on one GPU (gfx950).
//
// Replaces the reference's cuFFT + cuDecomp pipeline (multi/main.f90:833-923):
// divergence -> R2C x -> transpose -> C2C y -> transpose -> C2C z -> divide
// by -k^2 -> inverse z -> transpose -> inverse y -> transpose -> C2R x.
// On one GPU no transpose is needed: every pass reads the spectrum in
// place, in the layout that keeps its own axis strided and 16 (FP32) or 8
// (FP64) wavenumbers contiguous per row segment, and stages columns
// through LDS.
//
// Layout in HBM (C order):
//   spec[z][y][kx], kx < n/2   complex, the half spectrum without Nyquist
//   nyq [z][y]                 complex, the kx = n/2 (Nyquist) plane
//   pres[z][y][x]              real, the pressure
// Passes per projection:
//   X1 (x_fwd):   divergence of u* (main.f90:833-848) fused with the R2C
//                 along x, two real rows per complex transform
//   Y1:           forward C2C along y on spec and on nyq (one launch)
//   Z:            forward C2C along z, divide (main.f90:883-896), inverse,
//                 on spec and on nyq (one launch)
//   Y2:           inverse C2C along y (one launch)
//   X2 (x_inv):   C2R along x, writes pres
// For n <= 128 (FP32; 96 in FP64) X1 + Y1 and Y2 + X2 are one launch each,
// one workgroup per z plane (k_fft_xy_fwd, k_fft_yx_inv).
#include "fft.hpp"
#include "internal.hpp"

#include <vector>

namespace mhit {

// --------------------------------------------------------------- x forward
// One workgroup: PX pairs of y rows at one z.  Row pair (ya, ya+1) is
// transformed as one complex sequence a + i b and split with
// X_a[k] = (Z[k] + conj Z[N-k])/2, X_b[k] = (Z[k] - conj Z[N-k])/(2i).
// Input IN: X_IN_FIELD an FP64 field (us), X_IN_DIV the divergence of
// (us, vs, ws), X_IN_READY a field already in the transform precision
// (din: the divergence written by the fused NS pass).
//
// Every pass takes the arguments of one box, or a table of boxes of the
// same size (an ensemble) whose grids are stacked along z.
enum { X_IN_FIELD = 0, X_IN_DIV = 1, X_IN_READY = 2 };
template <typename T>
struct XFwdArgs {
    const double *us, *vs, *ws;
    const T* din;
    double cdiv;
    cplx_t<T> *spec, *nyq;
    const cplx_t<T>* tw;
    const DevScalars* S;
};
template <typename T, int N, int IN, int P, int PL>
__global__ __launch_bounds__((P * FftCfg<T, N>::TPC))
void k_fft_x_fwd(XFwdArgs<T> one, const XFwdArgs<T>* __restrict__ tab, int nbox) {
    using V = cplx_t<T>;
    using Cfg = FftCfg<T, N>;
    constexpr int TPC = Cfg::TPC, H = N / 2, LD = P, SW = P - 1;
    constexpr int NT = P * TPC;
    __shared__ V lds[N * LD];
    __shared__ V stw[N];
    const Blk blk = launch_blk(nbox);
    XFwdArgs<T> A = one;
    if (tab) A = tab[blk.box];
    const MHIT_GLOBAL double* __restrict__ us = gmem(A.us);
    const MHIT_GLOBAL double* __restrict__ vs = gmem(A.vs);
    const MHIT_GLOBAL double* __restrict__ ws = gmem(A.ws);
    const MHIT_GLOBAL T* __restrict__ din = gmem(A.din);
    const double cdiv = A.cdiv;
    MHIT_GLOBAL V* __restrict__ spec = gmem(A.spec);
    MHIT_GLOBAL V* __restrict__ nyq = gmem(A.nyq);
    const MHIT_GLOBAL V* __restrict__ tw = gmem(A.tw);
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    if (S && S->halted) return;
    const int y0 = blk.x * 2 * P;
    const int z = blk.y;
    const int tid = threadIdx.x;
    const int zp = (z + 1 == N) ? 0 : z + 1;
    stage_twiddles<V, N, NT>(stw, tw, tid);

    // compile-time trip counts below let the compiler issue all global loads
    // of a thread before the LDS writes that wait for them
#pragma unroll
    for (int r = 0; r < (N * P + NT - 1) / NT; r++) {
        const int e = tid + r * NT;
        if ((N * P) % NT != 0 && e >= N * P) break;
        const int a = e % N, c = e / N;
        V val;
        if (IN == X_IN_READY) {
            val.x = ld_once(&din[idx3(a, y0 + 2 * c, z, N)]);
            val.y = ld_once(&din[idx3(a, y0 + 2 * c + 1, z, N)]);
            lds[lidx<LD, SW>(a, c)] = val;
            continue;
        }
        double d[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int y = y0 + 2 * c + h;
            const size_t q = idx3(a, y, z, N);
            if (IN == X_IN_DIV) {
                // rhsp, main.f90:841-843
                const int ap = (a + 1 == N) ? 0 : a + 1;
                const int yp = (y + 1 == N) ? 0 : y + 1;
                double r = cdiv * (us[idx3(ap, y, z, N)] - us[q]);
                r = r + cdiv * (vs[idx3(a, yp, z, N)] - vs[q]);
                r = r + cdiv * (ws[idx3(a, y, zp, N)] - ws[q]);
                d[h] = r;
            } else {
                d[h] = us[q];
            }
        }
        val.x = (T)d[0];   // rhspp = real(rhsp,4) in the FP32 form (main.f90:844)
        val.y = (T)d[1];
        lds[lidx<LD, SW>(a, c)] = val;
    }
    __syncthreads();
    fft_stages<V, N, 1, P, TPC, LD, false, N, SW, PL>(lds, tid % P, tid / P, stw);

#pragma unroll
    for (int r = 0; r < (H * 2 * P + NT - 1) / NT; r++) {
        const int e = tid + r * NT;
        if ((H * 2 * P) % NT != 0 && e >= H * 2 * P) break;
        const int k = e % H, rr = e / H, cc = rr >> 1;
        const V zk = lds[lidx<LD, SW>(k, cc)];
        const V zm = lds[lidx<LD, SW>((N - k) % N, cc)];
        V out;
        if ((rr & 1) == 0) {
            out.x = (zk.x + zm.x) * (T)0.5;
            out.y = (zk.y - zm.y) * (T)0.5;
        } else {
            out.x = (zk.y + zm.y) * (T)0.5;
            out.y = (zm.x - zk.x) * (T)0.5;
        }
        gstore_once(spec + ((size_t)z * N + y0 + rr) * H + k, out);
    }
    for (int rr = tid; rr < 2 * P; rr += NT) {
        const V zh = lds[lidx<LD, SW>(H, rr >> 1)];
        V out;
        out.x = (rr & 1) ? zh.y : zh.x;
        out.y = (T)0;
        gstore_once(nyq + (size_t)z * N + y0 + rr, out);
    }
}

// --------------------------------------------------------------- x inverse
// C2R: Z[k] = X_a[k] + i X_b[k] over the Hermitian extension; the imaginary
// parts of the kx = 0 and kx = n/2 coefficients are ignored (a real
// transform has none).  Real part -> row ya, imaginary part -> row ya+1.
template <typename T>
struct XInvArgs {
    const cplx_t<T> *spec, *nyq;
    T* pres;
    const cplx_t<T>* tw;
    const DevScalars* S;
};
template <typename T, int N, int P, int PL>
__global__ __launch_bounds__((P * FftCfg<T, N>::TPC))
void k_fft_x_inv(XInvArgs<T> one, const XInvArgs<T>* __restrict__ tab, int nbox) {
    using V = cplx_t<T>;
    using Cfg = FftCfg<T, N>;
    constexpr int TPC = Cfg::TPC, H = N / 2, LD = P, SW = P - 1;
    constexpr int NT = P * TPC;
    __shared__ V lds[N * LD];
    __shared__ V stw[N];
    const Blk blk = launch_blk(nbox);
    XInvArgs<T> A = one;
    if (tab) A = tab[blk.box];
    const MHIT_GLOBAL V* __restrict__ spec = gmem(A.spec);
    const MHIT_GLOBAL V* __restrict__ nyq = gmem(A.nyq);
    MHIT_GLOBAL T* __restrict__ pres = gmem(A.pres);
    const MHIT_GLOBAL V* __restrict__ tw = gmem(A.tw);
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    if (S && S->halted) return;
    const int y0 = blk.x * 2 * P;
    const int z = blk.y;
    const int tid = threadIdx.x;
    stage_twiddles<V, N, NT>(stw, tw, tid);
    // row H: the kx = n/2 coefficients of each row pair (real parts)
    if (tid < P) {
        const V An = gload_once(nyq + (size_t)z * N + y0 + 2 * tid);
        const V Bn = gload_once(nyq + (size_t)z * N + y0 + 2 * tid + 1);
        V zh; zh.x = An.x; zh.y = Bn.x;
        lds[lidx<LD, SW>(H, tid)] = zh;
    }
    // rows k and n - k; kx = 0 (real parts only) by selects, without a
    // divergent branch: its two writes both go to row 0
#pragma unroll
    for (int r = 0; r < (H * P + NT - 1) / NT; r++) {
        const int e = tid + r * NT;
        if ((H * P) % NT != 0 && e >= H * P) break;
        const int k = e % H, c = e / H;
        const size_t ra = ((size_t)z * N + y0 + 2 * c) * H;
        const V A = gload_once(spec + ra + k);
        const V B = gload_once(spec + ra + H + k);
        const bool k0 = k == 0;
        V zk; zk.x = k0 ? A.x : A.x - B.y; zk.y = k0 ? B.x : A.y + B.x;
        V zm; zm.x = k0 ? A.x : A.x + B.y; zm.y = k0 ? B.x : B.x - A.y;
        lds[lidx<LD, SW>(k, c)] = zk;
        lds[lidx<LD, SW>(k0 ? 0 : N - k, c)] = zm;
    }
    __syncthreads();
    fft_stages<V, N, 1, P, TPC, LD, true, N, SW, PL>(lds, tid % P, tid / P, stw);

#pragma unroll
    for (int r = 0; r < (N * 2 * P + NT - 1) / NT; r++) {
        const int e = tid + r * NT;
        if ((N * 2 * P) % NT != 0 && e >= N * 2 * P) break;
        const int x = e % N, rr = e / N;
        const V zz = lds[lidx<LD, SW>(x, rr >> 1)];
        st_once(&pres[idx3(x, y0 + rr, z, N)], (rr & 1) ? zz.y : zz.x);
    }
}

// ------------------------------------------- x and y passes in one launch
// For n <= 128 a z plane's half spectrum fits in LDS, so one workgroup per
// plane runs the x transforms of all its rows and the y transforms of all
// its columns, and the spectrum goes to HBM once instead of twice
// (Tuning::fft_xy).  Forward: the N rows enter the x R2C as P = N/2 row
// pairs (columns of the LDS array, as in k_fft_x_fwd); the split rewrites
// the same array as the half spectrum sy[y][kx] (N/2 columns kx of length
// N; every thread holds its split values in registers across a barrier),
// and the y transforms run on those columns.  The kx = 0 and kx = n/2
// coefficients of every row are real, so column kx = 0 carries both,
// X[y][0] + i X[y][n/2]; after the y transform they are separated by
// Hermitian symmetry as the x pass separates a row pair.  Inverse: the
// same steps backwards (column 0 takes F[ky][0] + i F[ky][n/2], whose
// inverse along y has the two real columns as its real and imaginary
// parts).  One array of n (n/2 + 1) values; compiled where two workgroups
// fit a CU's LDS (n <= 128 FP32, n <= 96 FP64: FP64 at 128^3 measured
// slower with one, 0.190 -> 0.193 ms per multiphase step).
template <typename T, int N>
constexpr bool xy_fits() {
    return N % 8 == 0 && N >= 32 && N <= 128 &&
           (N * (N / 2 + 1) + N) * (int)sizeof(cplx_t<T>) <= 80 * 1024;
}

template <typename T, int N, int IN>
__global__ __launch_bounds__((N / 2) * (N / 8))
void k_fft_xy_fwd(XFwdArgs<T> one, const XFwdArgs<T>* __restrict__ tab, int nbox) {
    using V = cplx_t<T>;
    constexpr int P = N / 2, H = N / 2, TPC = N / 8, NT = P * TPC;
    constexpr bool POW2 = (P & (P - 1)) == 0;
    constexpr int LD = POW2 ? P : P + 1, SW = POW2 ? P - 1 : 0;
    constexpr int NSP = N * H / NT;   // split values per thread (8)
    static_assert(N * H % NT == 0 && H <= LD, "one plane per workgroup");
    __shared__ V lds[N * LD];   // x stage: row pair c as column c; then sy[y][kx]
    __shared__ V stw[N];
    V* const sy = lds;
    const Blk blk = launch_blk(nbox);
    XFwdArgs<T> A = one;
    if (tab) A = tab[blk.box];
    const MHIT_GLOBAL double* __restrict__ us = gmem(A.us);
    const MHIT_GLOBAL double* __restrict__ vs = gmem(A.vs);
    const MHIT_GLOBAL double* __restrict__ ws = gmem(A.ws);
    const MHIT_GLOBAL T* __restrict__ din = gmem(A.din);
    const double cdiv = A.cdiv;
    MHIT_GLOBAL V* __restrict__ spec = gmem(A.spec);
    MHIT_GLOBAL V* __restrict__ nyq = gmem(A.nyq);
    const MHIT_GLOBAL V* __restrict__ tw = gmem(A.tw);
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    if (S && S->halted) return;
    const int z = blk.y;
    const int tid = threadIdx.x;
    const int zp = (z + 1 == N) ? 0 : z + 1;
    stage_twiddles<V, N, NT>(stw, tw, tid);
#pragma unroll
    for (int r = 0; r < N * P / NT; r++) {
        const int e = tid + r * NT;
        const int a = e % N, c = e / N;
        V val;
        if (IN == X_IN_READY) {
            val.x = ld_once(&din[idx3(a, 2 * c, z, N)]);
            val.y = ld_once(&din[idx3(a, 2 * c + 1, z, N)]);
        } else {
            double d[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int y = 2 * c + h;
                const size_t q = idx3(a, y, z, N);
                if (IN == X_IN_DIV) {
                    // rhsp, main.f90:841-843
                    const int ap = (a + 1 == N) ? 0 : a + 1;
                    const int yp = (y + 1 == N) ? 0 : y + 1;
                    double rr = cdiv * (us[idx3(ap, y, z, N)] - us[q]);
                    rr = rr + cdiv * (vs[idx3(a, yp, z, N)] - vs[q]);
                    rr = rr + cdiv * (ws[idx3(a, y, zp, N)] - ws[q]);
                    d[h] = rr;
                } else {
                    d[h] = us[q];
                }
            }
            val.x = (T)d[0];   // rhspp = real(rhsp,4) in the FP32 form (main.f90:844)
            val.y = (T)d[1];
        }
        lds[lidx<LD, SW>(a, c)] = val;
    }
    __syncthreads();
    fft_stages<V, N, 1, P, TPC, LD, false, N, SW>(lds, tid % P, tid / P, stw);
    // split into rows y: X_y[k], k < H; column 0 takes X_y[0] + i X_y[H]
    V sp[NSP];
#pragma unroll
    for (int r = 0; r < NSP; r++) {
        const int e = tid + r * NT;
        const int k = e % H, y = e / H, cc = y >> 1;
        const V zk = lds[lidx<LD, SW>(k, cc)];
        const V zm = lds[lidx<LD, SW>((N - k) % N, cc)];
        const V zh = lds[lidx<LD, SW>(H, cc)];
        V out;
        if ((y & 1) == 0) {
            out.x = (zk.x + zm.x) * (T)0.5;
            out.y = (zk.y - zm.y) * (T)0.5;
        } else {
            out.x = (zk.y + zm.y) * (T)0.5;
            out.y = (zm.x - zk.x) * (T)0.5;
        }
        if (k == 0) out.y = (y & 1) ? zh.y : zh.x;   // X_y[0] is real: its imaginary slot takes X_y[H]
        sp[r] = out;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NSP; r++) sy[tid + r * NT] = sp[r];   // sy[y][k] at y H + k = e
    __syncthreads();
    fft_stages<V, N, 1, H, TPC, H, false>(sy, tid % H, tid / H, stw);
    // store spec[z][ky][kx]; kx = 0 and the Nyquist plane separated from column 0
#pragma unroll
    for (int r = 0; r < NSP; r++) {
        const int e = tid + r * NT;
        const int kx = e % H, ky = e / H;
        V out = sy[e];
        if (kx == 0) {
            const V cm = sy[((N - ky) % N) * H];
            const V c0 = out;
            out.x = (c0.x + cm.x) * (T)0.5;
            out.y = (c0.y - cm.y) * (T)0.5;
            V fn;
            fn.x = (c0.y + cm.y) * (T)0.5;
            fn.y = (cm.x - c0.x) * (T)0.5;
            gstore_once(nyq + (size_t)z * N + ky, fn);
        }
        gstore_once(spec + ((size_t)z * N + ky) * H + kx, out);
    }
}

template <typename T, int N>
__global__ __launch_bounds__((N / 2) * (N / 8))
void k_fft_yx_inv(XInvArgs<T> one, const XInvArgs<T>* __restrict__ tab, int nbox) {
    using V = cplx_t<T>;
    constexpr int P = N / 2, H = N / 2, TPC = N / 8, NT = P * TPC;
    constexpr bool POW2 = (P & (P - 1)) == 0;
    constexpr int LD = POW2 ? P : P + 1, SW = POW2 ? P - 1 : 0;
    constexpr int NSP = N * H / NT;   // spectrum values per thread (8)
    constexpr int NPR = H * P / NT;   // row-pair coefficients per thread (4)
    static_assert(N * H % NT == 0 && H * P % NT == 0 && P <= NT && H <= LD, "one plane per workgroup");
    __shared__ V lds[N * LD];   // sy[ky][kx]; then the x stage, row pair c as column c
    __shared__ V stw[N];
    V* const sy = lds;
    const Blk blk = launch_blk(nbox);
    XInvArgs<T> A = one;
    if (tab) A = tab[blk.box];
    const MHIT_GLOBAL V* __restrict__ spec = gmem(A.spec);
    const MHIT_GLOBAL V* __restrict__ nyq = gmem(A.nyq);
    MHIT_GLOBAL T* __restrict__ pres = gmem(A.pres);
    const MHIT_GLOBAL V* __restrict__ tw = gmem(A.tw);
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    if (S && S->halted) return;
    const int z = blk.y;
    const int tid = threadIdx.x;
    stage_twiddles<V, N, NT>(stw, tw, tid);
    // the plane's spectrum into sy[ky][kx]; column 0 takes F[ky][0] + i F[ky][H]
#pragma unroll
    for (int r = 0; r < NSP; r++) {
        const int e = tid + r * NT;
        const int kx = e % H, ky = e / H;
        V v = gload_once(spec + (size_t)z * N * H + e);
        const V fn = gload(nyq + (size_t)z * N + ky);   // one line per row, shared by its lanes
        if (kx == 0) {
            const V f0 = v;
            v.x = f0.x - fn.y;
            v.y = f0.y + fn.x;
        }
        sy[e] = v;
    }
    __syncthreads();
    fft_stages<V, N, 1, H, TPC, H, true>(sy, tid % H, tid / H, stw);
    // x C2R input of row pair c (as k_fft_x_inv): Z[k] = G[2c][k] + i G[2c+1][k]
    // over the Hermitian extension; G[y][0] = Re sy[y][0], G[y][H] = Im sy[y][0]
    V zh{};
    if (tid < P) {
        const V a0 = sy[(2 * tid) * H], b0 = sy[(2 * tid + 1) * H];
        zh.x = a0.y;
        zh.y = b0.y;
    }
    V zk[NPR], zm[NPR];
#pragma unroll
    for (int r = 0; r < NPR; r++) {
        const int e = tid + r * NT;
        const int k = e % H, c = e / H;
        const V Av = sy[(2 * c) * H + k];
        const V Bv = sy[(2 * c + 1) * H + k];
        const bool k0 = k == 0;
        zk[r].x = k0 ? Av.x : Av.x - Bv.y; zk[r].y = k0 ? Bv.x : Av.y + Bv.x;
        zm[r].x = k0 ? Av.x : Av.x + Bv.y; zm[r].y = k0 ? Bv.x : Bv.x - Av.y;
    }
    __syncthreads();
    if (tid < P) lds[lidx<LD, SW>(H, tid)] = zh;
#pragma unroll
    for (int r = 0; r < NPR; r++) {
        const int e = tid + r * NT;
        const int k = e % H, c = e / H;
        lds[lidx<LD, SW>(k, c)] = zk[r];
        lds[lidx<LD, SW>(k == 0 ? 0 : N - k, c)] = zm[r];
    }
    __syncthreads();
    fft_stages<V, N, 1, P, TPC, LD, true, N, SW>(lds, tid % P, tid / P, stw);
#pragma unroll
    for (int r = 0; r < N * 2 * P / NT; r++) {
        const int e = tid + r * NT;
        const int x = e % N, rr = e / N;
        const V zz = lds[lidx<LD, SW>(x, rr >> 1)];
        st_once(&pres[idx3(x, rr, z, N)], (rr & 1) ? zz.y : zz.x);
    }
}

// ------------------------------------------------------------ column pass
// C columns of length N, element (c, a) at data[base + a*sa + c*sc] with
// base = bxi*bx + byi*by.  MODE 0 forward, 1 inverse, 2 forward + spectral
// divide + inverse (the z pass).  AXIS_CONTIG picks the global access order
// (a fastest when the axis is contiguous).  NYQ marks the Nyquist plane,
// where the columns are y wavenumbers and kx = n/2; otherwise the columns
// are kx wavenumbers and byi is ky.
template <typename T>
struct ColArgs {
    cplx_t<T>* data;
    size_t bx, by, sa, sc;
    const cplx_t<T>* tw;
    const double* kk;
    double ntot;
    const DevScalars* S;
};
// The columns of one workgroup; lds holds N (C + 1) values, stw N, skk N
// (MODE 2).
template <typename T, int N, int C, bool AXIS_CONTIG, int MODE, bool WF>
__device__ __forceinline__ void cols_body(const ColArgs<T>& A, const int bxi, const int byi, const bool NYQ,
                                          cplx_t<T>* __restrict__ lds, cplx_t<T>* __restrict__ stw,
                                          double* __restrict__ skk) {
    using V = cplx_t<T>;
    constexpr int TPC = FftCfg<T, N>::TPC;
    constexpr int LD = AXIS_CONTIG ? C + 1 : C;
    constexpr int NT = C * TPC;
    MHIT_GLOBAL V* __restrict__ data = gmem(A.data);
    const size_t bx = A.bx, by = A.by, sa = A.sa, sc = A.sc;
    const MHIT_GLOBAL V* __restrict__ tw = gmem(A.tw);
    const MHIT_GLOBAL double* __restrict__ kk = gmem(A.kk);
    const double ntot = A.ntot;
    const int tid = threadIdx.x;
    MHIT_GLOBAL V* __restrict__ g = data + bxi * bx + byi * by;
    // whole 128-byte lines per row segment: stream the stores; a half-line
    // segment's other half is written by the paired workgroup
    constexpr bool LINES = C * (int)sizeof(V) >= 128;
    auto ld = [](const MHIT_GLOBAL V* p) -> V { return gload(p); };
    auto st = [](MHIT_GLOBAL V* p, const V& v) {
        if constexpr (LINES) gstore_once(p, v);
        else gstore(p, v);
    };

    static_assert((N * C) % NT == 0, "column tile must be a multiple of the workgroup");
    // Wavefront form: N = 8 M 8 (M = 1, 2, 4, 8), eight columns, N/8
    // threads per column, thread t = b + 8 w of column c in lane 8 b + c of
    // wave w.  With n = n0 + 8 n1 + 8M n2 the thread holds x[t + 8M r], so
    // n0 = b is a lane digit, n1 = w the wave and n2 = r the register.
    // Decimation in frequency, k = k0 + 8 k1 + 8M k2:
    //   radix 8 over n2 in registers -> k0; twiddle W_N^(k0 t);
    //   LDS exchange (the only cross-wave step): wave <- k0 (8/M values
    //   per wave), registers <- n1;
    //   radix M over n1 -> k1; twiddle W_N^(8 n0 k1);
    //   lane/register transpose inside the wave (xpose_lane_hi) -> n0;
    //   radix 8 over n0 -> k2.
    // X[k] is then in wave w, lane digit b, register r with
    // k = spec_k(w, b, r) below.  The inverse runs the same steps backwards
    // from that layout and ends in the natural one, so the z pass needs
    // two LDS exchanges (three barriers with the twiddle staging) instead of
    // four exchanges and seven barriers.  N = 64 has no cross-wave step.
    // Each wave writes its half of the inverse exchange to the same 512
    // values it read in the forward one, so no barrier separates them.
    constexpr bool WAVE = WF && wave_form_fits<T, N, C, AXIS_CONTIG>();
    if constexpr (WAVE) {
        constexpr int M = N / 64, S = 8 / M;
        const int c = tid & 7, b = (tid >> 3) & 7, w = tid >> 6, t = b + 8 * w;
        V v[8];
        auto spec_k = [&](int r) { return (w * S + b / M) + 8 * (b % M) + 8 * M * r; };
        if constexpr (MODE == 1) {
#pragma unroll
            for (int r = 0; r < 8; r++) v[r] = ld(g + (size_t)spec_k(r) * sa + c * sc);
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) v[r] = ld(g + (size_t)(t + r * TPC) * sa + c * sc);
        }
        stage_twiddles<V, N, NT>(stw, tw, tid);
        if constexpr (MODE == 2) {
            for (int e = tid; e < N; e += NT) skk[e] = kk[e];
        }
        lds_sync();
        auto twiddle = [&](V& x, int j, auto inv) {
            V wj = stw[j];
            if (decltype(inv)::value) wj.y = -wj.y;
            x = c_mul_fma(x, wj);
        };
        auto forward = [&]() {
            dft8<false>(v);
#pragma unroll
            for (int r = 1; r < 8; r++) twiddle(v[r], r * t, std::false_type{});
            if constexpr (M > 1) {
#pragma unroll
                for (int r = 0; r < 8; r++) lds[(r * TPC + t) * LD + c] = v[r];
                lds_sync();
#pragma unroll
                for (int s = 0; s < S; s++)
#pragma unroll
                    for (int q = 0; q < M; q++) v[s * M + q] = lds[((w * S + s) * TPC + q * 8 + b) * LD + c];
#pragma unroll
                for (int s = 0; s < S; s++) {
                    dft<M, false>(v + s * M);
#pragma unroll
                    for (int q = 1; q < M; q++) twiddle(v[s * M + q], 8 * b * q, std::false_type{});
                }
            }
            xpose_lane_hi(v);
            dft8<false>(v);
        };
        // the inverse from the layout the forward transform leaves; b is
        // then the digit s M + k1 and, after the transpose, n0
        auto inverse = [&]() {
            dft8<true>(v);
            if constexpr (M > 1) {
#pragma unroll
                for (int r = 1; r < 8; r++) twiddle(v[r], 8 * (b % M) * r, std::true_type{});
            }
            xpose_lane_hi(v);
            if constexpr (M > 1) {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    dft<M, true>(v + s * M);
#pragma unroll
                    for (int q = 0; q < M; q++) {
                        const int j = (w * S + s) * (b + 8 * q);
                        if (j) twiddle(v[s * M + q], j, std::true_type{});
                    }
                }
#pragma unroll
                for (int s = 0; s < S; s++)
#pragma unroll
                    for (int q = 0; q < M; q++) lds[((w * S + s) * TPC + q * 8 + b) * LD + c] = v[s * M + q];
                lds_sync();
#pragma unroll
                for (int r = 0; r < 8; r++) v[r] = lds[(r * TPC + t) * LD + c];
            } else {
#pragma unroll
                for (int r = 1; r < 8; r++) twiddle(v[r], r * b, std::true_type{});
            }
            dft8<true>(v);
        };
        if constexpr (MODE == 1) {
            inverse();
        } else {
            forward();
        }
        if constexpr (MODE == 2) {
            // main.f90:883-896: p^ = -r^/(k2 nx ny nz), k2 = kx^2+ky^2+kz^2
            // in FP64, stored back in the spectrum precision; zero mode 0.
            // One refined hardware reciprocal per element (within 1 ulp of
            // 1/x) replaces the reference's two divisions per component.
            const int kxi = NYQ ? N / 2 : bxi * C + c;
            const int kyi = NYQ ? bxi * C + c : byi;
            const double kxy = skk[kxi] + skk[kyi];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int kz = spec_k(r);
                if (!NYQ && kxi == 0 && kyi == 0 && kz == 0) {
                    v[r].x = (T)0; v[r].y = (T)0;
                } else {
                    const double s = -rcp_refined((kxy + skk[kz]) * ntot);
                    v[r].x = (T)((double)v[r].x * s);
                    v[r].y = (T)((double)v[r].y * s);
                }
            }
            inverse();
        }
        if constexpr (MODE == 0) {
#pragma unroll
            for (int r = 0; r < 8; r++) st(g + (size_t)spec_k(r) * sa + c * sc, v[r]);
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) st(g + (size_t)(t + r * TPC) * sa + c * sc, v[r]);
        }
        return;
    }
    // Register-direct form (strided columns, N/8 threads per column, radix-8
    // last stage): thread t of column c loads and stores elements t + r N/8,
    // which are the inputs of its first-stage butterfly and the outputs of
    // its last-stage butterfly.  The z pass keeps the spectrum in registers
    // from the last forward stage through the divide into the first inverse
    // stage.  Half of the LDS round trips and barriers of the general form.
    constexpr bool DIRECT = !AXIS_CONTIG && N % 8 == 0 && TPC == N / 8 && last_radix(N) == 8;
    if constexpr (DIRECT) {
        constexpr bool DIRECT_IN = N >= 64 && next_radix(N) == 8;
        constexpr int NS1 = DIRECT_IN ? 8 : 1;     // points combined after the first stage
        constexpr bool INV0 = MODE == 1;
        const int c = tid % C, t = tid / C;
        V v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = ld(g + (size_t)(t + r * TPC) * sa + c * sc);
        stage_twiddles<V, N, NT>(stw, tw, tid);
        // per-axis k^2, staged with the data
        if constexpr (MODE == 2) {
            for (int e = tid; e < N; e += NT) skk[e] = kk[e];
        }
        auto to_lds = [&](auto inv) {
            if constexpr (DIRECT_IN) {
                fft_first8_to_lds<V, N, LD, decltype(inv)::value>(lds, c, t, v);
            } else {
#pragma unroll
                for (int r = 0; r < 8; r++) lds[(t + r * TPC) * LD + c] = v[r];
            }
        };
        to_lds(std::integral_constant<bool, INV0>{});
        __syncthreads();
        fft_stages<V, N, NS1, C, TPC, LD, INV0, N / 8>(lds, c, t, stw);
        fft_last8_from_lds<V, N, LD, INV0>(lds, c, t, v, stw);
        if constexpr (MODE == 2) {
            // main.f90:883-896: p^ = -r^/(k2 nx ny nz), k2 = kx^2+ky^2+kz^2
            // in FP64, stored back in the spectrum precision; zero mode 0.
            // One FP64 reciprocal per element replaces the reference's two
            // divisions per component (at most 2 ulp apart in FP64).
            __syncthreads();   // every thread holds its last-stage inputs
            const int kxi = NYQ ? N / 2 : bxi * C + c;
            const int kyi = NYQ ? bxi * C + c : byi;
            const double kxy = skk[kxi] + skk[kyi];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int a = t + r * TPC;
                if (!NYQ && kxi == 0 && kyi == 0 && a == 0) {
                    v[r].x = (T)0; v[r].y = (T)0;
                } else {
                    const double s = -1.0 / ((kxy + skk[a]) * ntot);
                    v[r].x = (T)((double)v[r].x * s);
                    v[r].y = (T)((double)v[r].y * s);
                }
            }
            to_lds(std::integral_constant<bool, true>{});
            __syncthreads();
            fft_stages<V, N, NS1, C, TPC, LD, true, N / 8>(lds, c, t, stw);
            fft_last8_from_lds<V, N, LD, true>(lds, c, t, v, stw);
        }
#pragma unroll
        for (int r = 0; r < 8; r++) st(g + (size_t)(t + r * TPC) * sa + c * sc, v[r]);
        return;
    }
    {
        V buf[(N * C) / NT];   // all loads of the thread in flight before the LDS writes
#pragma unroll
        for (int r = 0; r < (N * C) / NT; r++) {
            const int e = tid + r * NT;
            int a, c;
            if (AXIS_CONTIG) { a = e % N; c = e / N; } else { c = e % C; a = e / C; }
            buf[r] = ld(g + a * sa + c * sc);
        }
        stage_twiddles<V, N, NT>(stw, tw, tid);
#pragma unroll
        for (int r = 0; r < (N * C) / NT; r++) {
            const int e = tid + r * NT;
            int a, c;
            if (AXIS_CONTIG) { a = e % N; c = e / N; } else { c = e % C; a = e / C; }
            lds[a * LD + c] = buf[r];
        }
    }
    const int c = tid % C, t = tid / C;
    // the z pass: wavenumbers of the thread's elements, loaded with the data
    constexpr int NA = (N + TPC - 1) / TPC;
    double k2a[MODE == 2 ? NA : 1];
    double kxy = 0.0;
    int kxi = 0, kyi = 0;
    if constexpr (MODE == 2) {
        kxi = NYQ ? N / 2 : bxi * C + c;
        kyi = NYQ ? bxi * C + c : byi;
        kxy = kk[kxi] + kk[kyi];
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int a = t + q * TPC;
            k2a[q] = (N % TPC == 0 || a < N) ? kk[a] : 0.0;
        }
    }
    __syncthreads();
    if (MODE == 0 || MODE == 2) fft_stages<V, N, 1, C, TPC, LD, false>(lds, c, t, stw);
    else fft_stages<V, N, 1, C, TPC, LD, true>(lds, c, t, stw);

    if constexpr (MODE == 2) {
        // the divide as in the register-direct form above
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int a = t + q * TPC;
            if (N % TPC != 0 && a >= N) break;
            V val = lds[a * LD + c];
            if (!NYQ && kxi == 0 && kyi == 0 && a == 0) {
                val.x = (T)0; val.y = (T)0;
            } else {
                const double s = -1.0 / ((kxy + k2a[q]) * ntot);
                val.x = (T)((double)val.x * s);
                val.y = (T)((double)val.y * s);
            }
            lds[a * LD + c] = val;
        }
        __syncthreads();
        fft_stages<V, N, 1, C, TPC, LD, true>(lds, c, t, stw);
    }

#pragma unroll
    for (int r = 0; r < (N * C) / NT; r++) {
        const int e = tid + r * NT;
        int a, cc;
        if (AXIS_CONTIG) { a = e % N; cc = e / N; } else { cc = e % C; a = e / C; }
        st(g + a * sa + cc * sc, lds[a * LD + cc]);
    }
}

// One launch per column pass.  Workgroup rows y < N hold the bulk
// spectrum's columns (AXIS_CONTIG, NYQ = false); the two rows y = N, N + 1
// hold the N / C workgroups of the Nyquist plane (NYQ_AC, NYQ = true), with
// their own arguments (`onen`, `tabn`), so the plane costs no launch of its
// own.
//
// Workgroups dispatch to the 8 XCDs round-robin, each XCD with its own L2.
// When a row segment is half a 128-byte line, the two workgroups that share
// the line are renumbered to land on the same XCD, 8 dispatches apart, so
// the second reads the line from L2: within each aligned group of 16
// workgroups of one part (bulk or Nyquist), where the part starts on a
// multiple of 16 in dispatch order.
template <typename T, int N, int C, bool AC, int MODE, bool NYQ_AC, bool WF>
__global__ __launch_bounds__((C * FftCfg<T, N>::TPC))
void k_fft_cols(ColArgs<T> one, ColArgs<T> onen, const ColArgs<T>* __restrict__ tab,
                const ColArgs<T>* __restrict__ tabn, int nbox) {
    using V = cplx_t<T>;
    __shared__ V lds[N * ((AC || NYQ_AC) ? C + 1 : C)];
    __shared__ V stw[N];
    __shared__ double skk[MODE == 2 ? N : 1];
    const Blk blk = launch_blk(nbox);
    const bool nyq = blk.y >= N;
    ColArgs<T> A = nyq ? onen : one;
    if (tab) A = nyq ? tabn[blk.box] : tab[blk.box];
    const MHIT_GLOBAL DevScalars* __restrict__ S = gmem(A.S);
    if (S && S->halted) return;
    const bool ac = nyq ? NYQ_AC : AC;
    const int start = nyq ? blk.nx * N : 0;                  // first workgroup of the part
    const int count = nyq ? 2 * blk.nx : blk.nx * N;
    int L = blk.x + blk.nx * blk.y - start;
    const long long first = (long long)blk.box * blk.nx * blk.ny + start;   // in dispatch order
    if (!ac && C * (int)sizeof(V) == 64 && blk.nx % 2 == 0 && count % 16 == 0 && first % 16 == 0) {
        const int r = L % 16;
        L = L - r + (r % 8) * 2 + r / 8;
    }
    const int bxi = nyq ? L : L % blk.nx, byi = nyq ? 0 : L / blk.nx;
    if constexpr (AC == NYQ_AC) {
        cols_body<T, N, C, AC, MODE, WF>(A, bxi, byi, nyq, lds, stw, skk);   // one body for both parts
    } else {
        if (nyq) cols_body<T, N, C, NYQ_AC, MODE, WF>(A, bxi, byi, true, lds, stw, skk);
        else cols_body<T, N, C, AC, MODE, WF>(A, bxi, byi, false, lds, stw, skk);
    }
}

// ---------------------------------------------------------------- drivers
// The argument of box 0 (`one`) and, for a batched launch (arena given),
// the device table of all nb boxes.
template <typename A, typename F>
static int with_table(mhit_ctx* const* boxes, int nb, ArgArena* arena, F build, A& one, const A*& tab) {
    one = build(boxes[0]);
    tab = nullptr;
    if (!arena) return MHIT_OK;
    std::vector<A> t(nb);
    for (int b = 0; b < nb; b++) t[b] = build(boxes[b]);
    tab = arena->put(t.data(), nb);
    if (!tab) { set_error("batched launch: argument arena full"); return MHIT_ERR_STATE; }
    return MHIT_OK;
}

// Passes over boxes[0..nb) of size N and precision T on stream st; a single
// box is nb = 1 without an arena.
template <typename T, int N>
struct PoissonPasses {
    using V = cplx_t<T>;
    using Cfg = FftCfg<T, N>;
    static constexpr int H = N / 2;
    using Boxes = mhit_ctx* const*;
    // launch widths compiled for this size (fft_widths lists the same):
    // the default first, alternatives from n = 64 up
    static constexpr bool ALT = N >= 64;
    static constexpr int Y0 = Cfg::CY, Y1 = Cfg::CY / 2;
    static constexpr int Z0 = sizeof(T) == 4 ? Cfg::CY / 2 : Cfg::CY, Z1 = sizeof(T) == 4 ? Cfg::CY : Cfg::CY / 2;
    static constexpr int X0 = Cfg::PX, X1 = Cfg::PX / 2, X2 = Cfg::PX * 2;
    static constexpr bool X2OK = H % X2 == 0;

    static int x_fwd(Boxes bx, int nb, ArgArena* ar, hipStream_t st, const double* a, const double* b,
                     const double* w, int in) {
        const int p = fft_width(bx[0], 2);
        if constexpr (ALT) {
            if (p == X1) return x_fwd_p<X1>(bx, nb, ar, st, a, b, w, in);
            if constexpr (X2OK) {
                if (p == X2) return x_fwd_p<X2>(bx, nb, ar, st, a, b, w, in);
            }
        }
        return x_fwd_p<X0>(bx, nb, ar, st, a, b, w, in);
    }
    template <int P>
    static int x_fwd_p(Boxes bx, int nb, ArgArena* ar, hipStream_t st, const double* a, const double* b,
                       const double* w, int in) {
        XFwdArgs<T> one;
        const XFwdArgs<T>* tab;
        int rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
            XFwdArgs<T> x;
            x.us = in == X_IN_READY ? nullptr : a;
            x.vs = in == X_IN_DIV ? b : nullptr;
            x.ws = in == X_IN_DIV ? w : nullptr;
            x.din = in == X_IN_READY ? (const T*)c->pres : nullptr;
            x.cdiv = in == X_IN_DIV ? c->g.rho * c->g.dxi / c->g.dt : 0.0;
            x.spec = (V*)c->spec;
            x.nyq = (V*)c->nyq;
            x.tw = (const V*)c->tw;
            x.S = c->S;
            return x;
        }, one, tab);
        if (rc) return rc;
        dim3 grid(N / (2 * P), N, nb), block(P * Cfg::TPC);
        if constexpr (plans_differ(N)) {
            if (bx[0]->tune.fft_xplan) {
                x_fwd_launch<P, 1>(grid, block, st, one, tab, nb, in);
                return MHIT_OK;
            }
        }
        x_fwd_launch<P, 0>(grid, block, st, one, tab, nb, in);
        return MHIT_OK;
    }
    template <int P, int PL>
    static void x_fwd_launch(dim3 grid, dim3 block, hipStream_t st, const XFwdArgs<T>& one,
                             const XFwdArgs<T>* tab, int nb, int in) {
        if (in == X_IN_READY) k_fft_x_fwd<T, N, X_IN_READY, P, PL><<<grid, block, 0, st>>>(one, tab, nb);
        else if (in == X_IN_DIV) k_fft_x_fwd<T, N, X_IN_DIV, P, PL><<<grid, block, 0, st>>>(one, tab, nb);
        else k_fft_x_fwd<T, N, X_IN_FIELD, P, PL><<<grid, block, 0, st>>>(one, tab, nb);
    }
    // A column pass in the context's transform form (Tuning::fft_wave):
    // the wavefront form where it fits, else the register-direct / LDS
    // form.  Boxes of one batched launch share the form.  The grid is the
    // bulk's (nxb, N) workgroups per box plus two rows for the Nyquist plane.
    template <int C, bool AC, int MODE, bool NAC>
    static void launch_cols(const mhit_ctx* c, int nxb, const ColArgs<T>& one, const ColArgs<T>& onen,
                            const ColArgs<T>* tab, const ColArgs<T>* tabn, int nb, hipStream_t st) {
        const dim3 grid(nxb, N + 2, nb), block(C * Cfg::TPC);
        if constexpr (wave_form_fits<T, N, C, AC>()) {
            if (c->tune.fft_wave) {
                k_fft_cols<T, N, C, AC, MODE, NAC, true><<<grid, block, 0, st>>>(one, onen, tab, tabn, nb);
                return;
            }
        }
        k_fft_cols<T, N, C, AC, MODE, NAC, false><<<grid, block, 0, st>>>(one, onen, tab, tabn, nb);
    }
    template <int MODE>
    static int y_pass(Boxes bx, int nb, ArgArena* ar, hipStream_t st) {
        if constexpr (ALT) {
            if (fft_width(bx[0], 0) == Y1) return y_pass_w<MODE, Y1>(bx, nb, ar, st);
        }
        return y_pass_w<MODE, Y0>(bx, nb, ar, st);
    }
    template <int MODE, int CY>
    static int y_pass_w(Boxes bx, int nb, ArgArena* ar, hipStream_t st) {
        ColArgs<T> one, onen;
        const ColArgs<T> *tab, *tabn;
        // spec: columns kx at fixed z; axis y (stride H)
        int rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
            return ColArgs<T>{(V*)c->spec, (size_t)CY, (size_t)N * H, (size_t)H, 1,
                              (const V*)c->tw, c->kk, 0.0, c->S};
        }, one, tab);
        if (rc) return rc;
        // nyq [z][y]: columns z, axis y contiguous
        rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
            return ColArgs<T>{(V*)c->nyq, (size_t)CY * N, 0, 1, (size_t)N,
                              (const V*)c->tw, c->kk, 0.0, c->S};
        }, onen, tabn);
        if (rc) return rc;
        launch_cols<CY, false, MODE, true>(bx[0], H / CY, one, onen, tab, tabn, nb, st);
        return MHIT_OK;
    }
    // Default z width, FP32: half the y pass's columns per workgroup
    // (64-byte row segments): the two FFTs of the z pass make it the
    // longest-running column kernel, and smaller workgroups (two per CU at
    // its register count) hide its barriers better (0.38 -> 0.33 ms at
    // 512^3).  FP64 keeps 8 columns (128-byte segments; 4 columns measured
    // 0.84 vs 0.55 ms).
    template <int MODE>
    static int z_pass(Boxes bx, int nb, ArgArena* ar, hipStream_t st) {
        if constexpr (ALT) {
            if (fft_width(bx[0], 1) == Z1) return z_pass_w<MODE, Z1>(bx, nb, ar, st);
        }
        return z_pass_w<MODE, Z0>(bx, nb, ar, st);
    }
    template <int MODE, int CY>
    static int z_pass_w(Boxes bx, int nb, ArgArena* ar, hipStream_t st) {
        const double ntot = (double)((long long)N * N * N);
        ColArgs<T> one, onen;
        const ColArgs<T> *tab, *tabn;
        // spec: columns kx at fixed y; axis z (stride N*H)
        int rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
            return ColArgs<T>{(V*)c->spec, (size_t)CY, (size_t)H, (size_t)N * H, 1,
                              (const V*)c->tw, c->kk, ntot, c->S};
        }, one, tab);
        if (rc) return rc;
        // nyq [z][y]: columns y, axis z (stride N)
        rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
            return ColArgs<T>{(V*)c->nyq, (size_t)CY, 0, (size_t)N, 1,
                              (const V*)c->tw, c->kk, ntot, c->S};
        }, onen, tabn);
        if (rc) return rc;
        launch_cols<CY, false, MODE, false>(bx[0], H / CY, one, onen, tab, tabn, nb, st);
        return MHIT_OK;
    }
    static int x_inv(Boxes bx, int nb, ArgArena* ar, hipStream_t st) {
        const int p = fft_width(bx[0], 2);
        if constexpr (ALT) {
            if (p == X1) return x_inv_p<X1>(bx, nb, ar, st);
            if constexpr (X2OK) {
                if (p == X2) return x_inv_p<X2>(bx, nb, ar, st);
            }
        }
        return x_inv_p<X0>(bx, nb, ar, st);
    }
    template <int P>
    static int x_inv_p(Boxes bx, int nb, ArgArena* ar, hipStream_t st) {
        XInvArgs<T> one;
        const XInvArgs<T>* tab;
        int rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
            return XInvArgs<T>{(const V*)c->spec, (const V*)c->nyq, (T*)c->pres, (const V*)c->tw, c->S};
        }, one, tab);
        if (rc) return rc;
        dim3 grid(N / (2 * P), N, nb), block(P * Cfg::TPC);
        if constexpr (plans_differ(N)) {
            if (bx[0]->tune.fft_xplan) {
                k_fft_x_inv<T, N, P, 1><<<grid, block, 0, st>>>(one, tab, nb);
                return MHIT_OK;
            }
        }
        k_fft_x_inv<T, N, P, 0><<<grid, block, 0, st>>>(one, tab, nb);
        return MHIT_OK;
    }
    // x and y passes in one launch per z plane (Tuning::fft_xy) where the
    // plane fits in LDS
    static constexpr bool XY = xy_fits<T, N>();
    static bool use_xy(const mhit_ctx* c) { return XY && c->tune.fft_xy; }
    static int xy_fwd(Boxes bx, int nb, ArgArena* ar, hipStream_t st, const double* a, const double* b,
                      const double* w, int in) {
        if constexpr (XY) {
            XFwdArgs<T> one;
            const XFwdArgs<T>* tab;
            int rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
                XFwdArgs<T> x;
                x.us = in == X_IN_READY ? nullptr : a;
                x.vs = in == X_IN_DIV ? b : nullptr;
                x.ws = in == X_IN_DIV ? w : nullptr;
                x.din = in == X_IN_READY ? (const T*)c->pres : nullptr;
                x.cdiv = in == X_IN_DIV ? c->g.rho * c->g.dxi / c->g.dt : 0.0;
                x.spec = (V*)c->spec;
                x.nyq = (V*)c->nyq;
                x.tw = (const V*)c->tw;
                x.S = c->S;
                return x;
            }, one, tab);
            if (rc) return rc;
            dim3 grid(1, N, nb), block((N / 2) * (N / 8));
            if (in == X_IN_READY) k_fft_xy_fwd<T, N, X_IN_READY><<<grid, block, 0, st>>>(one, tab, nb);
            else if (in == X_IN_DIV) k_fft_xy_fwd<T, N, X_IN_DIV><<<grid, block, 0, st>>>(one, tab, nb);
            else k_fft_xy_fwd<T, N, X_IN_FIELD><<<grid, block, 0, st>>>(one, tab, nb);
            return MHIT_OK;
        } else {
            set_error("fused x/y transform not compiled for this size");
            return MHIT_ERR_STATE;
        }
    }
    static int yx_inv(Boxes bx, int nb, ArgArena* ar, hipStream_t st) {
        if constexpr (XY) {
            XInvArgs<T> one;
            const XInvArgs<T>* tab;
            int rc = with_table(bx, nb, ar, [&](mhit_ctx* c) {
                return XInvArgs<T>{(const V*)c->spec, (const V*)c->nyq, (T*)c->pres, (const V*)c->tw, c->S};
            }, one, tab);
            if (rc) return rc;
            dim3 grid(1, N, nb), block((N / 2) * (N / 8));
            k_fft_yx_inv<T, N><<<grid, block, 0, st>>>(one, tab, nb);
            return MHIT_OK;
        } else {
            set_error("fused x/y transform not compiled for this size");
            return MHIT_ERR_STATE;
        }
    }
    // x then y forward, or y then x inverse, fused or as two passes
    // (mark: stage events for the per-stage profile of a single box; the
    // fused launch is timed as the x stage, its y stage as zero)
    static int fwd_xy(Boxes bx, int nb, ArgArena* ar, hipStream_t st, const double* a, const double* b,
                      const double* w, int in, bool mark) {
        if (use_xy(bx[0])) {
            const int rc = xy_fwd(bx, nb, ar, st, a, b, w, in);
            if (mark) { prof_mark(bx[0], ST_FFT_X_FWD); prof_mark(bx[0], ST_FFT_Y_FWD); }
            return rc;
        }
        int rc;
        if ((rc = x_fwd(bx, nb, ar, st, a, b, w, in))) return rc;
        if (mark) prof_mark(bx[0], ST_FFT_X_FWD);
        if ((rc = y_pass<0>(bx, nb, ar, st))) return rc;
        if (mark) prof_mark(bx[0], ST_FFT_Y_FWD);
        return MHIT_OK;
    }
    static int inv_yx(Boxes bx, int nb, ArgArena* ar, hipStream_t st, bool mark) {
        if (use_xy(bx[0])) {
            const int rc = yx_inv(bx, nb, ar, st);
            if (mark) { prof_mark(bx[0], ST_FFT_Y_INV); prof_mark(bx[0], ST_FFT_X_INV); }
            return rc;
        }
        int rc;
        if ((rc = y_pass<1>(bx, nb, ar, st))) return rc;
        if (mark) prof_mark(bx[0], ST_FFT_Y_INV);
        if ((rc = x_inv(bx, nb, ar, st))) return rc;
        if (mark) prof_mark(bx[0], ST_FFT_X_INV);
        return MHIT_OK;
    }
    static int project(mhit_ctx* c, const double* a, const double* b, const double* w, int in) {
        mhit_ctx* const bx[1] = {c};
        hipStream_t st = c->stream;
        int rc;
        if ((rc = fwd_xy(bx, 1, nullptr, st, a, b, w, in, true))) return rc;
        if ((rc = z_pass<2>(bx, 1, nullptr, st))) return rc;
        prof_mark(c, ST_FFT_Z);
        return inv_yx(bx, 1, nullptr, st, true);
    }
    static int project_batch(Boxes bx, int nb, ArgArena& ar, hipStream_t st) {
        int rc;
        if ((rc = fwd_xy(bx, nb, &ar, st, nullptr, nullptr, nullptr, X_IN_READY, false))) return rc;
        if ((rc = z_pass<2>(bx, nb, &ar, st))) return rc;
        return inv_yx(bx, nb, &ar, st, false);
    }
    static int r2c(mhit_ctx* c, const double* in) {
        mhit_ctx* const bx[1] = {c};
        int rc;
        if ((rc = fwd_xy(bx, 1, nullptr, c->stream, in, nullptr, nullptr, X_IN_FIELD, false))) return rc;
        return z_pass<0>(bx, 1, nullptr, c->stream);
    }
    static int c2r(mhit_ctx* c) {
        mhit_ctx* const bx[1] = {c};
        int rc;
        if ((rc = z_pass<1>(bx, 1, nullptr, c->stream))) return rc;
        return inv_yx(bx, 1, nullptr, c->stream, false);
    }
};

#define MHIT_FOR_EACH_N(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

bool fft_size_supported(int n) {
    switch (n) {
#define CASE(NN) case NN:
        MHIT_FOR_EACH_N(CASE)
#undef CASE
        return true;
        default: return false;
    }
}

bool fft_xplan_differs(int n) { return plans_differ(n); }

bool fft_xy_supported(int n, bool fp32) {
    switch (n) {
#define CASE(NN) \
    case NN: return fp32 ? xy_fits<float, NN>() : xy_fits<double, NN>();
        MHIT_FOR_EACH_N(CASE)
#undef CASE
        default: return false;
    }
}

int fft_widths(int n, bool fp32, int pass, int out[3]) {
    const int h = n / 2;
    const int cy = pow2_divisor_cap(h, fp32 ? 16 : 8), px = pow2_divisor_cap(h, fp32 ? 8 : 4);
    const bool alt = n >= 64;
    int k = 0;
    if (pass == 0) {
        out[k++] = cy;
        if (alt) out[k++] = cy / 2;
    } else if (pass == 1) {
        out[k++] = fp32 ? cy / 2 : cy;
        if (alt) out[k++] = fp32 ? cy : cy / 2;
    } else {
        out[k++] = px;
        if (alt) {
            out[k++] = px / 2;
            if (h % (2 * px) == 0) out[k++] = 2 * px;
        }
    }
    return k;
}

int fft_width(const mhit_ctx* c, int pass) {
    int w[3];
    const int k = fft_widths(c->n, c->fp32, pass, w);
    const int want = pass == 0 ? c->tune.fft_ycols : pass == 1 ? c->tune.fft_zcols : c->tune.fft_xrows;
    for (int i = 0; i < k; i++)
        if (w[i] == want) return want;
    return w[0];
}

enum class Op { Project, ProjectReady, Solve, R2C, C2R };

template <typename T>
static int dispatch(mhit_ctx* c, Op op, const double* a, const double* b, const double* w) {
    int rc = MHIT_OK;
    switch (c->n) {
#define CASE(NN)                                                                        \
    case NN:                                                                            \
        if (op == Op::Project) rc = PoissonPasses<T, NN>::project(c, a, b, w, X_IN_DIV);  \
        else if (op == Op::ProjectReady)                                               \
            rc = PoissonPasses<T, NN>::project(c, a, b, w, X_IN_READY);                \
        else if (op == Op::Solve) rc = PoissonPasses<T, NN>::project(c, a, b, w, X_IN_FIELD);\
        else if (op == Op::R2C) rc = PoissonPasses<T, NN>::r2c(c, a);                  \
        else rc = PoissonPasses<T, NN>::c2r(c);                                         \
        break;
        MHIT_FOR_EACH_N(CASE)
#undef CASE
        default:
            set_error("unsupported grid size for the FFT");
            return MHIT_ERR_ARG;
    }
    if (rc) return rc;
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

template <typename T>
static int dispatch_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    int rc = MHIT_OK;
    switch (boxes[0]->n) {
#define CASE(NN) case NN: rc = PoissonPasses<T, NN>::project_batch(boxes, nb, arena, st); break;
        MHIT_FOR_EACH_N(CASE)
#undef CASE
        default:
            set_error("unsupported grid size for the FFT");
            return MHIT_ERR_ARG;
    }
    if (rc) return rc;
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

static int run(mhit_ctx* c, Op op, const double* a, const double* b, const double* w) {
    return c->fp32 ? dispatch<float>(c, op, a, b, w) : dispatch<double>(c, op, a, b, w);
}

// The fused NS pass leaves the divergence of u* in the pressure buffer; the
// cell-per-thread path leaves u* only.
int poisson_project(mhit_ctx* c) {
    return c->march ? run(c, Op::ProjectReady, nullptr, nullptr, nullptr)
                    : run(c, Op::Project, c->us, c->vs, c->ws);
}
int poisson_project_batch(mhit_ctx* const* boxes, int nb, ArgArena& arena, hipStream_t st) {
    mhit_ctx* c = boxes[0];
    for (int b = 0; b < nb; b++)
        if (boxes[b]->n != c->n || boxes[b]->fp32 != c->fp32 || !boxes[b]->march ||
            boxes[b]->tune.fft_wave != c->tune.fft_wave || fft_width(boxes[b], 0) != fft_width(c, 0) ||
            fft_width(boxes[b], 1) != fft_width(c, 1) || fft_width(boxes[b], 2) != fft_width(c, 2)) {
            set_error("batched projection: boxes differ in size, precision, transform form or launch "
                      "widths, or do not use the fused passes");
            return MHIT_ERR_ARG;
        }
    return c->fp32 ? dispatch_batch<float>(boxes, nb, arena, st) : dispatch_batch<double>(boxes, nb, arena, st);
}
int poisson_solve_field(mhit_ctx* c, const double* rhs) { return run(c, Op::Solve, rhs, nullptr, nullptr); }
int fft_r2c_dev(mhit_ctx* c, const double* in) { return run(c, Op::R2C, in, nullptr, nullptr); }
int fft_c2r_dev(mhit_ctx* c) { return run(c, Op::C2R, nullptr, nullptr, nullptr); }

}  // namespace mhit
