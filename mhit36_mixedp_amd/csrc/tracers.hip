// tracers.hip -- Lagrangian tracer step (gfx950).
//
// One thread per tracer: interpolate u^n at the tracer (particles.f90:78-122),
// advance with AB2/Euler (main.f90:480-486), wrap into the box
// (particles.f90:364-380) and keep the interpolated velocity as the
// previous-step velocity (main.f90:503-508).  At one rank the reference's
// y/z sort and send-to-self (main.f90:489-498) only permute the tracers,
// so it has no counterpart here; tracers keep their order and their tag.
#include "internal.hpp"

#include <utility>

namespace mhit {

__device__ __forceinline__ int pmod(long long a, int n) {
    long long r = a % n;
    return (int)(r < 0 ? r + n : r);
}

// Reference interpolation: 1-D linear per component along its own axis.
// The y/z indices are those of the halo-shifted local array
// (iy = 1 + floor(y/dx) with index 1 the halo) -> interior (fy-1) mod n.
__device__ __forceinline__ void interp_reference(const double* __restrict__ u, const double* __restrict__ v,
                                                 const double* __restrict__ w, const double* m,
                                                 double xp, double yp, double zp, const Grid& g,
                                                 double& up, double& vp, double& wp) {
    const int n = g.n;
    const double dx = g.dx;
    const long long fx = (long long)floor(xp / dx), fy = (long long)floor(yp / dx),
                    fz = (long long)floor(zp / dx);
    const double tx = xp / dx - (double)fx;
    const double ty = yp / dx - (double)fy;
    const double tz = zp / dx - (double)fz;
    const int i0 = pmod(fx, n), i1 = pmod(fx + 1, n);
    const int j0 = pmod(fy - 1, n), j1 = pmod(fy, n);
    const int k0 = pmod(fz - 1, n), k1 = pmod(fz, n);
    const size_t q = idx3(i0, j0, k0, n);
    up = (1.0 - tx) * (u[q] - m[0]) + tx * (u[idx3(i1, j0, k0, n)] - m[0]);
    vp = (1.0 - ty) * (v[q] - m[1]) + ty * (v[idx3(i0, j1, k0, n)] - m[1]);
    wp = (1.0 - tz) * (w[q] - m[2]) + tz * (w[idx3(i0, j0, k1, n)] - m[2]);
}

// Trilinear option: each component at its staggered location in the
// time-stepper frame (u at ((i-1/2)dx,(j+1)dx,(k+1)dx), v at
// (i dx,(j+1/2)dx,(k+1)dx), w at (i dx,(j+1)dx,(k+1/2)dx)).
__device__ __forceinline__ double trilin(const double* __restrict__ f, double mean, double xs, double ys,
                                         double zs, int n) {
    const double fxl = floor(xs), fyl = floor(ys), fzl = floor(zs);
    const double tx = xs - fxl, ty = ys - fyl, tz = zs - fzl;
    const long long ix = (long long)fxl, iy = (long long)fyl, iz = (long long)fzl;
    const int i0 = pmod(ix, n), i1 = pmod(ix + 1, n);
    const int j0 = pmod(iy, n), j1 = pmod(iy + 1, n);
    const int k0 = pmod(iz, n), k1 = pmod(iz + 1, n);
    const double c00 = (1.0 - tx) * (f[idx3(i0, j0, k0, n)] - mean) + tx * (f[idx3(i1, j0, k0, n)] - mean);
    const double c10 = (1.0 - tx) * (f[idx3(i0, j1, k0, n)] - mean) + tx * (f[idx3(i1, j1, k0, n)] - mean);
    const double c01 = (1.0 - tx) * (f[idx3(i0, j0, k1, n)] - mean) + tx * (f[idx3(i1, j0, k1, n)] - mean);
    const double c11 = (1.0 - tx) * (f[idx3(i0, j1, k1, n)] - mean) + tx * (f[idx3(i1, j1, k1, n)] - mean);
    const double c0 = (1.0 - ty) * c00 + ty * c10;
    const double c1 = (1.0 - ty) * c01 + ty * c11;
    return (1.0 - tz) * c0 + tz * c1;
}

template <int MODE>
__global__ __launch_bounds__(256)
void k_tracers(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
               double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
               double* __restrict__ pu1, double* __restrict__ pv1, double* __restrict__ pw1,
               long long np, Grid g, const DevScalars* __restrict__ S) {
    if (S->halted) return;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double m[3] = {S->mean[0], S->mean[1], S->mean[2]};
    const double alpha = S->alpha, beta = S->beta, dt = g.dt, lx = g.lx;
    double x = px[p], y = py[p], z = pz[p];
    double up, vp, wp;
    if (MODE == MHIT_INTERP_TRILINEAR) {
        const double xs = x / g.dx, ys = y / g.dx, zs = z / g.dx;
        up = trilin(u, m[0], xs + 0.5, ys - 1.0, zs - 1.0, g.n);
        vp = trilin(v, m[1], xs, ys - 0.5, zs - 1.0, g.n);
        wp = trilin(w, m[2], xs, ys - 1.0, zs - 0.5, g.n);
    } else {
        interp_reference(u, v, w, m, x, y, z, g, up, vp, wp);
    }
    x = x + dt * (alpha * up - beta * pu1[p]);
    y = y + dt * (alpha * vp - beta * pv1[p]);
    z = z + dt * (alpha * wp - beta * pw1[p]);
    // FLOOR returns a default integer in the reference
    x = x - lx * (double)(int)floor(x / lx);
    y = y - lx * (double)(int)floor(y / lx);
    z = z - lx * (double)(int)floor(z / lx);
    px[p] = x; py[p] = y; pz[p] = z;
    pu1[p] = up; pv1[p] = vp; pw1[p] = wp;
}

// ----------------------------------------------------------- cell order
// Tracers are re-ordered every few steps by the index of their cell,
// (k n + j) n + i (stable radix sort), so that the gathers of consecutive
// tracers share cache lines and sweep u, v, w in memory order: the rows
// j + 1 and planes k + 1 that the v and w gathers also read come back a
// few hundred tracers later, from L2.  Measured at 512^3 with 4M tracers:
// tracer pass 0.53 ms in the Morton order of the cells, 0.44 ms in this
// one.  This is a permutation only: every tracer keeps its tag and its
// values, and the order is a deterministic function of the positions.

__global__ __launch_bounds__(256)
void k_tracer_keys(const double* __restrict__ px, const double* __restrict__ py,
                   const double* __restrict__ pz, unsigned* __restrict__ keys, int* __restrict__ idx,
                   long long np, Grid g) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int n = g.n;
    const int ci = pmod((long long)floor(px[p] / g.dx), n);
    const int cj = pmod((long long)floor(py[p] / g.dx), n);
    const int ck = pmod((long long)floor(pz[p] / g.dx), n);
    keys[p] = ((unsigned)ck * (unsigned)n + (unsigned)cj) * (unsigned)n + (unsigned)ci;
    idx[p] = (int)p;
}

struct Cols7 { double* c[7]; };

__global__ __launch_bounds__(256)
void k_tracer_permute(const int* __restrict__ perm, Cols7 src, Cols7 dst, long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int s = perm[p];
#pragma unroll
    for (int t = 0; t < 7; t++) dst.c[t][p] = src.c[t][s];
}

// Stable LSD radix sort of (key, index) pairs, 8 bits per pass.  A
// workgroup of 256 threads owns a tile of 4096 consecutive keys; wave w
// holds the quarter [1024 w, 1024 w + 1024) in 16 rows of 64 lanes, so the
// tile order is (wave, row, lane).  Per pass:
//   k_sort_hist     digit counts of each tile -> counts[digit][tile]
//   k_sort_scan     one workgroup per digit: exclusive prefix of its row
//                   over the tiles, in place, and the digit's total
//   k_sort_scatter  each key goes to (keys of smaller digits) + (its
//                   digit's keys in earlier tiles) + its rank among the
//                   tile's earlier keys of the same digit (lanes of one
//                   digit found by ballots over its 8 bits; running counts
//                   per wave in LDS, then summed over the earlier waves)
// Equal digits keep their order, so the passes compose to a stable sort.
constexpr int kSortThreads = 256, kSortRows = 16, kSortTile = kSortThreads * kSortRows;
constexpr int kSortDigits = 256, kSortWaves = kSortThreads / 64;

__global__ __launch_bounds__(kSortThreads)
void k_sort_hist(const unsigned* __restrict__ keys, unsigned* __restrict__ counts, long long np, int shift,
                 int ntile) {
    __shared__ unsigned h[kSortDigits];
    const int tid = threadIdx.x;
    h[tid] = 0u;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortRows; r++) {
        const long long p = base + r * kSortThreads + tid;
        if (p < np) atomicAdd(&h[(keys[p] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    counts[(size_t)tid * ntile + blockIdx.x] = h[tid];
}

// Exclusive scan of a digit's row over the tiles (blockIdx.x = digit), in
// chunks of 4 entries per thread; the row's total goes to totals[digit].
__global__ __launch_bounds__(kSortThreads)
void k_sort_scan(unsigned* __restrict__ counts, unsigned* __restrict__ totals, int ntile) {
    __shared__ unsigned wsum[kSortWaves];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    unsigned* row = counts + (size_t)blockIdx.x * ntile;
    unsigned carry = 0u;
    for (int c0 = 0; c0 < ntile; c0 += 4 * kSortThreads) {
        unsigned x[4];
        unsigned s = 0u;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = c0 + 4 * tid + q;
            x[q] = e < ntile ? row[e] : 0u;
            s += x[q];
        }
        unsigned incl = s;   // inclusive scan of the thread sums over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        unsigned before = carry, all = 0u;
#pragma unroll
        for (int v = 0; v < kSortWaves; v++) {
            if (v < w) before += wsum[v];
            all += wsum[v];
        }
        before += incl - s;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = c0 + 4 * tid + q;
            if (e < ntile) row[e] = before;
            before += x[q];
        }
        carry += all;
        __syncthreads();   // wsum is reused by the next chunk
    }
    if (tid == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kSortThreads)
void k_sort_scatter(const unsigned* __restrict__ kin, const int* __restrict__ vin, unsigned* __restrict__ kout,
                    int* __restrict__ vout, const unsigned* __restrict__ counts, const unsigned* __restrict__ totals,
                    long long np, int shift, int ntile) {
    // running digit counts of each wave, then the wave's offsets in the tile
    __shared__ unsigned wcnt[kSortWaves][kSortDigits];
    __shared__ unsigned toff[kSortDigits];   // output offset of the tile's run of each digit
    __shared__ unsigned wsum[kSortWaves];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int v = 0; v < kSortWaves; v++) wcnt[v][tid] = 0u;
    {   // keys of smaller digits: exclusive scan of the totals (thread = digit)
        const unsigned t = totals[tid];
        unsigned incl = t;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        unsigned before = incl - t;
        for (int v = 0; v < w; v++) before += wsum[v];
        toff[tid] = before + counts[(size_t)tid * ntile + blockIdx.x];
    }
    __syncthreads();
    const long long base = (long long)blockIdx.x * kSortTile + (long long)w * (kSortTile / kSortWaves);
    unsigned key[kSortRows], rank[kSortRows];
    int val[kSortRows];
#pragma unroll
    for (int r = 0; r < kSortRows; r++) {
        const long long p = base + r * 64 + lane;
        key[r] = p < np ? kin[p] : 0u;
        val[r] = p < np ? vin[p] : 0;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    // the leader's update must reach the next row's reads in program order
    volatile unsigned* wc = wcnt[w];
#pragma unroll
    for (int r = 0; r < kSortRows; r++) {
        const bool ok = base + r * 64 + lane < np;
        const unsigned d = (key[r] >> shift) & 0xffu;
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const bool set = (d >> bit) & 1u;
            const unsigned long long m = __ballot(set);
            peers &= set ? m : ~m;
        }
        const unsigned before = wc[d];
        rank[r] = before + (unsigned)__popcll(peers & below);
        if (ok && (peers & below) == 0ull) wc[d] = before + (unsigned)__popcll(peers);
    }
    __syncthreads();
    {   // thread tid = digit: exclusive sum over the waves
        unsigned run = 0u;
#pragma unroll
        for (int v = 0; v < kSortWaves; v++) {
            const unsigned t = wcnt[v][tid];
            wcnt[v][tid] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRows; r++) {
        if (base + r * 64 + lane >= np) continue;
        const unsigned d = (key[r] >> shift) & 0xffu;
        const unsigned pos = toff[d] + wcnt[w][d] + rank[r];
        kout[pos] = key[r];
        vout[pos] = val[r];
    }
}

void free_sort_buffers(mhit_ctx* c) {
    void* ptrs[] = {c->skey[0], c->skey[1], c->sidx[0], c->sidx[1], c->sort_tmp,
                    c->alt[0], c->alt[1], c->alt[2], c->alt[3], c->alt[4], c->alt[5], c->alt[6]};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    c->skey[0] = c->skey[1] = nullptr;
    c->sidx[0] = c->sidx[1] = nullptr;
    c->sort_tmp = nullptr;
    for (double*& a : c->alt) a = nullptr;
    c->sort_tmp_bytes = 0;
    c->sort_cap = 0;
}

int sort_tracers(mhit_ctx* c) {
    const long long np = c->np;
    if (np <= 1) return MHIT_OK;
    if (np > (long long)INT32_MAX) { set_error("tracer sort: more than 2^31 - 1 tracers"); return MHIT_ERR_ARG; }
    const int ntile = div_up(np, kSortTile);
    if (np > c->sort_cap) {
        free_sort_buffers(c);
        for (int t = 0; t < 2; t++) {
            MHIT_HIP_TRY(hipMalloc((void**)&c->skey[t], np * sizeof(unsigned)));
            MHIT_HIP_TRY(hipMalloc((void**)&c->sidx[t], np * sizeof(int)));
        }
        for (double*& a : c->alt) MHIT_HIP_TRY(hipMalloc((void**)&a, np * sizeof(double)));
        c->sort_tmp_bytes = (size_t)kSortDigits * (ntile + 1) * sizeof(unsigned);
        MHIT_HIP_TRY(hipMalloc(&c->sort_tmp, c->sort_tmp_bytes));
        c->sort_cap = np;
    }
    unsigned* counts = (unsigned*)c->sort_tmp;               // [digit][tile]
    unsigned* totals = counts + (size_t)kSortDigits * ntile;   // [digit]
    const int blocks = div_up(np, 256);
    k_tracer_keys<<<blocks, 256, 0, c->stream>>>(c->px, c->py, c->pz, c->skey[0], c->sidx[0], np, c->g);
    MHIT_HIP_TRY(hipGetLastError());
    // key bits in use: n^3 < 2^(3 axis_bits)
    int axis_bits = 0;
    while ((1 << axis_bits) < c->n) axis_bits++;
    int cur = 0;
    for (int shift = 0; shift < 3 * axis_bits; shift += 8) {
        k_sort_hist<<<ntile, kSortThreads, 0, c->stream>>>(c->skey[cur], counts, np, shift, ntile);
        k_sort_scan<<<kSortDigits, kSortThreads, 0, c->stream>>>(counts, totals, ntile);
        k_sort_scatter<<<ntile, kSortThreads, 0, c->stream>>>(c->skey[cur], c->sidx[cur], c->skey[cur ^ 1],
                                                               c->sidx[cur ^ 1], counts, totals, np, shift, ntile);
        MHIT_HIP_TRY(hipGetLastError());
        cur ^= 1;
    }
    double** cols[7] = {&c->ptag, &c->px, &c->py, &c->pz, &c->pu1, &c->pv1, &c->pw1};
    Cols7 src, dst;
    for (int t = 0; t < 7; t++) { src.c[t] = *cols[t]; dst.c[t] = c->alt[t]; }
    k_tracer_permute<<<blocks, 256, 0, c->stream>>>(c->sidx[cur], src, dst, np);
    MHIT_HIP_TRY(hipGetLastError());
    for (int t = 0; t < 7; t++) std::swap(*cols[t], c->alt[t]);
    prof_mark(c, ST_TRACER_SORT);
    return MHIT_OK;
}

int launch_tracers(mhit_ctx* c, hipStream_t st) {
    if (c->np <= 0) return MHIT_OK;
    const int blocks = div_up(c->np, 256);
    if (c->P.interp_mode == MHIT_INTERP_TRILINEAR)
        k_tracers<MHIT_INTERP_TRILINEAR><<<blocks, 256, 0, st>>>(
            c->u, c->v, c->w, c->px, c->py, c->pz, c->pu1, c->pv1, c->pw1, c->np, c->g, c->S);
    else
        k_tracers<MHIT_INTERP_REFERENCE><<<blocks, 256, 0, st>>>(
            c->u, c->v, c->w, c->px, c->py, c->pz, c->pu1, c->pv1, c->pw1, c->np, c->g, c->S);
    prof_mark(c, ST_TRACERS);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

}  // namespace mhit
