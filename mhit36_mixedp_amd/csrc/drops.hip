// drops.hip -- drop statistics from the phase field on the GPU.
//
// The reference does this after the run, from the phi/u/v/w dumps
// (post/unionfind_periodic/main.f90: mask phi >= 0.5, 6- or 26-point
// connectivity across the periodic faces, iterated min-label propagation;
// per drop: voxel count, centre as the circular mean of the cell centres
// x(i) = (i - 1/2) dx, mean velocity, volume = count dx^3) and
// post/DropCountSize/calculate_deq.f90 (d_eq = (6 V / pi)^(1/3)).
//
// Here it runs on the live state between steps:
//   1. init:    L[q] = q for masked cells, -1 otherwise
//   2. merge:   for every masked cell and each "backward" neighbour (half the
//               stencil, periodic wrap), union of the two trees with an
//               atomic min link (a root is always linked to a smaller
//               index, so every chain strictly decreases: no cycles, and the
//               final root of a component is its smallest cell index)
//   3. flatten: L[q] = root(q)
//   4. roots:   cells with L[q] == q, sorted on the host -- the drop number
//               is the rank of the smallest cell index, the reference's
//               ordering of its final (minimum) labels
//   5. sums:    per drop, cell count, sum of sin/cos of the cell centres on
//               each axis and the velocity sums, with FP64 atomics
//               pre-reduced across the lanes of a wave that share a drop
// The label array reuses the spectrum buffer (scratch between steps).
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mhit {
int sync_scalars_public(mhit_ctx* c);
}

using namespace mhit;

namespace {

constexpr int kStats = 10;   // count, sin x, cos x, sin y, cos y, sin z, cos z, u, v, w

__device__ __forceinline__ int load_relaxed(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int find_root(const int* L, int x) {
    int y;
    while ((y = load_relaxed(&L[x])) != x) x = y;   // strictly decreasing chain
    return x;
}

__device__ void unite(int* L, int a, int b) {
    for (;;) {
        a = find_root(L, a);
        b = find_root(L, b);
        if (a == b) return;
        if (a > b) { const int t = a; a = b; b = t; }   // link the larger root to the smaller
        const int old = atomicMin(&L[b], a);
        if (old == b) return;                           // b was still a root: linked
        b = old;                                        // b had been linked meanwhile: retry lower
    }
}

__global__ void k_cc_init(const double* __restrict__ phi, int* __restrict__ L, double thr, long long n3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n3) L[q] = (phi[q] >= thr) ? (int)q : -1;
}

// neighbour offsets of the backward half of the stencil
__constant__ int c_nb[13][3] = {
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},                                  // faces
    {-1, -1, 0}, {1, -1, 0}, {-1, 0, -1}, {1, 0, -1}, {0, -1, -1}, {0, 1, -1},  // edges
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},                  // corners
};

__global__ void k_cc_merge(int* __restrict__ L, int n, int nnb) {
    const long long n3 = (long long)n * n * n;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3 || L[q] < 0) return;
    const int i = (int)(q % n), j = (int)((q / n) % n), k = (int)(q / ((long long)n * n));
    for (int t = 0; t < nnb; t++) {
        int a = i + c_nb[t][0], b = j + c_nb[t][1], cz = k + c_nb[t][2];
        a = a < 0 ? a + n : (a >= n ? a - n : a);
        b = b < 0 ? b + n : (b >= n ? b - n : b);
        cz = cz < 0 ? cz + n : cz;
        const long long r = ((long long)cz * n + b) * n + a;
        if (load_relaxed(&L[r]) >= 0) unite(L, (int)q, (int)r);
    }
}

__global__ void k_cc_flatten(int* __restrict__ L, long long n3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n3 && L[q] >= 0) L[q] = find_root(L, (int)q);
}

__global__ void k_cc_roots(const int* __restrict__ L, long long n3, int* __restrict__ roots, int cap,
                           int* __restrict__ count) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n3 && L[q] == (int)q) {
        const int s = atomicAdd(count, 1);
        if (s < cap) roots[s] = (int)q;
    }
}

__global__ void k_cc_sums(const int* __restrict__ L, const int* __restrict__ roots, int nroots,
                          const double* __restrict__ u, const double* __restrict__ v,
                          const double* __restrict__ w, const double* __restrict__ ctab,
                          const DevScalars* __restrict__ S, double* __restrict__ stats, int n) {
    const long long n3 = (long long)n * n * n;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int d = -1;
    double val[kStats] = {};
    if (q < n3) {
        const int r = L[q];
        if (r >= 0) {
            int lo = 0, hi = nroots - 1;          // rank of the root (roots sorted)
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (roots[mid] < r) lo = mid + 1; else hi = mid;
            }
            d = lo;
            const int i = (int)(q % n), j = (int)((q / n) % n), k = (int)(q / ((long long)n * n));
            // ctab: sin(x_m), cos(x_m) of the cell centres x_m = (m + 1/2) dx
            val[0] = 1.0;
            val[1] = ctab[2 * i]; val[2] = ctab[2 * i + 1];
            val[3] = ctab[2 * j]; val[4] = ctab[2 * j + 1];
            val[5] = ctab[2 * k]; val[6] = ctab[2 * k + 1];
            val[7] = u[q] - S->mean[0];
            val[8] = v[q] - S->mean[1];
            val[9] = w[q] - S->mean[2];
        }
    }
    // lanes of a wave that all belong to the same drop add once
    const int d0 = __shfl(d, 0, 64);
    const bool uniform = __all(d == d0);
    if (uniform) {
        if (d0 < 0) return;
#pragma unroll
        for (int f = 0; f < kStats; f++) {
            double x = val[f];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            val[f] = x;
        }
        if ((threadIdx.x & 63) == 0)
            for (int f = 0; f < kStats; f++) atomicAdd(&stats[(size_t)d0 * kStats + f], val[f]);
    } else if (d >= 0) {
        for (int f = 0; f < kStats; f++) atomicAdd(&stats[(size_t)d * kStats + f], val[f]);
    }
}

}  // namespace

extern "C" {

int mhit_drop_stats(mhit_ctx* c, double threshold, int connectivity, int64_t max_drops,
                    mhit_drop* out, int64_t* ndrops) {
    if (!c || !ndrops || max_drops < 0 || (max_drops > 0 && !out) ||
        (connectivity != 6 && connectivity != 26)) {
        set_error("bad arguments (connectivity is 6 or 26)");
        return MHIT_ERR_ARG;
    }
    if (!c->P.phiflag) { set_error("no phase field (phiflag = 0)"); return MHIT_ERR_STATE; }
    int rc;
    if ((rc = sync_scalars_public(c))) return rc;
    const int n = c->n;
    const long long n3 = (long long)n * n * n;
    int* L = (int*)c->spec;   // the half spectrum holds >= n^3 ints; scratch between steps
    const int nb = (int)((n3 + 255) / 256);
    const double L2PI = 2.0 * M_PI, dx = L2PI / n;

    std::vector<double> tab(2 * (size_t)n);
    for (int m = 0; m < n; m++) {
        tab[2 * m] = std::sin(L2PI * ((m + 0.5) * dx) / L2PI);
        tab[2 * m + 1] = std::cos(L2PI * ((m + 0.5) * dx) / L2PI);
    }
    double* dtab = nullptr;
    int* dcount = nullptr;
    int* roots = nullptr;
    double* stats = nullptr;
    auto cleanup = [&]() {
        if (dtab) (void)hipFree(dtab);
        if (dcount) (void)hipFree(dcount);
        if (roots) (void)hipFree(roots);
        if (stats) (void)hipFree(stats);
    };
#define DROPS_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); \
        set_error(std::string("HIP: ") + hipGetErrorString(e_)); return MHIT_ERR_HIP; } } while (0)
    DROPS_TRY(hipMalloc((void**)&dtab, tab.size() * sizeof(double)));
    DROPS_TRY(hipMalloc((void**)&dcount, sizeof(int)));
    DROPS_TRY(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    k_cc_init<<<nb, 256, 0, c->stream>>>(c->phi, L, threshold, n3);
    k_cc_merge<<<nb, 256, 0, c->stream>>>(L, n, connectivity == 6 ? 3 : 13);
    k_cc_flatten<<<nb, 256, 0, c->stream>>>(L, n3);
    // roots: first pass counts (capacity 0), second collects
    int count = 0, cap = 0;
    for (int pass = 0; pass < 2; pass++) {
        DROPS_TRY(hipMemsetAsync(dcount, 0, sizeof(int), c->stream));
        k_cc_roots<<<nb, 256, 0, c->stream>>>(L, n3, roots, cap, dcount);
        DROPS_TRY(hipGetLastError());
        DROPS_TRY(hipMemcpyAsync(&count, dcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        DROPS_TRY(hipStreamSynchronize(c->stream));
        if (pass == 0) {
            if (count == 0) break;
            cap = count;
            DROPS_TRY(hipMalloc((void**)&roots, (size_t)cap * sizeof(int)));
        }
    }
    *ndrops = count;
    if (count > 0) {
        std::vector<int> hr(count);
        DROPS_TRY(copy_on_stream(c, hr.data(), roots, (size_t)count * sizeof(int), hipMemcpyDeviceToHost));
        std::sort(hr.begin(), hr.end());
        DROPS_TRY(copy_on_stream(c, roots, hr.data(), (size_t)count * sizeof(int), hipMemcpyHostToDevice));
        DROPS_TRY(hipMalloc((void**)&stats, (size_t)count * kStats * sizeof(double)));
        DROPS_TRY(hipMemsetAsync(stats, 0, (size_t)count * kStats * sizeof(double), c->stream));
        k_cc_sums<<<nb, 256, 0, c->stream>>>(L, roots, count, c->u, c->v, c->w, dtab, c->S, stats, n);
        DROPS_TRY(hipGetLastError());
        std::vector<double> hs((size_t)count * kStats);
        DROPS_TRY(hipMemcpyAsync(hs.data(), stats, hs.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        DROPS_TRY(hipStreamSynchronize(c->stream));
        const int64_t nout = std::min<int64_t>(count, max_drops);
        for (int64_t d = 0; d < nout; d++) {
            const double* s = &hs[(size_t)d * kStats];
            auto centre = [&](double sn, double cs) {
                double x = L2PI * std::atan2(sn, cs) / L2PI;
                return x < 0.0 ? x + L2PI : x;
            };
            mhit_drop& o = out[d];
            o.cells = (int64_t)std::llround(s[0]);
            o.x = centre(s[1], s[2]);
            o.y = centre(s[3], s[4]);
            o.z = centre(s[5], s[6]);
            o.u = s[7] / s[0];
            o.v = s[8] / s[0];
            o.w = s[9] / s[0];
            o.volume = s[0] * dx * dx * dx;
            o.d_eq = std::cbrt(6.0 * o.volume / M_PI);
            o.first_cell = hr[d];
        }
    }
    cleanup();
#undef DROPS_TRY
    return MHIT_OK;
}

}  // extern "C"
