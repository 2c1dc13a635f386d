// init.hip -- initial conditions computed on the device.
#include "internal.hpp"

namespace mhit {

// Taylor-Green, main.f90:324-337.  x(jg), x(kg) of the time-stepper frame
// are xext[j+1], xext[k+1] (xext = x_ext of main.f90:216-217).
__global__ __launch_bounds__(256)
void k_init_tgv(double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                const double* __restrict__ xe, Grid g) {
    const int n = g.n;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= n || j >= n) return;
    const size_t q = idx3(i, j, k, n);
    const double xi = xe[i], yj = xe[j + 1], zk = xe[k + 1];
    u[q] = sin(xi - g.dx / 2) * cos(yj) * cos(zk);
    v[q] = -cos(xi) * sin(yj - g.dx / 2) * cos(zk);
    w[q] = 0.0;
}

// Centred drop, main.f90:370-382.
__global__ __launch_bounds__(256)
void k_init_drop(double* __restrict__ phi, const double* __restrict__ xe, double radius, Grid g) {
    const int n = g.n;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= n || j >= n) return;
    const double h = g.lx / 2;
    const double a = xe[i] - h, b = xe[j + 1] - h, c = xe[k + 1] - h;
    const double pos = a * a + b * b + c * c;
    phi[idx3(i, j, k, n)] = 0.5 * (1.0 - tanh((sqrt(pos) - radius) / 2 / g.eps));
}

// Sum of Fourier modes in index space:
//   u_d(i,j,k) = sum_m Re(A_md exp(2 pi i (mx i + my j + mz k)/n)),
// modes as 9 doubles (mx, my, mz, Re/Im A_x, A_y, A_z), E[t] = exp(2 pi i t/n).
__global__ __launch_bounds__(256)
void k_init_modes(double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                  const double* __restrict__ modes, int nmodes, const double2* __restrict__ E, Grid g) {
    __shared__ double sm[64 * 9];
    const int n = g.n;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    double su = 0.0, sv = 0.0, sw = 0.0;
    for (int m0 = 0; m0 < nmodes; m0 += 64) {
        const int cnt = min(64, nmodes - m0);
        __syncthreads();
        for (int e = tid; e < cnt * 9; e += 256) sm[e] = modes[(size_t)m0 * 9 + e];
        __syncthreads();
        if (i < n && j < n) {
            for (int mm = 0; mm < cnt; mm++) {
                const double* M = sm + mm * 9;
                const int mx = (int)M[0], my = (int)M[1], mz = (int)M[2];
                const double2 ex = E[(((long long)mx * i) % n + n) % n];
                const double2 ey = E[(((long long)my * j) % n + n) % n];
                const double2 ez = E[(((long long)mz * k) % n + n) % n];
                const double exy_r = ex.x * ey.x - ex.y * ey.y, exy_i = ex.x * ey.y + ex.y * ey.x;
                const double er = exy_r * ez.x - exy_i * ez.y, ei = exy_r * ez.y + exy_i * ez.x;
                su += M[3] * er - M[4] * ei;
                sv += M[5] * er - M[6] * ei;
                sw += M[7] * er - M[8] * ei;
            }
        }
    }
    if (i < n && j < n) {
        const size_t q = idx3(i, j, k, n);
        u[q] = su; v[q] = sv; w[q] = sw;
    }
}

// Drops: phi = max over drops of 1/2(1 - tanh((r - R)/(2 eps))), r the
// periodic minimum-image distance from the cell centre (time-stepper frame).
__global__ __launch_bounds__(256)
void k_init_drops(double* __restrict__ phi, const double* __restrict__ xe,
                  const double* __restrict__ centers, int ndrops, double radius, Grid g) {
    __shared__ double sc[256 * 3];
    const int n = g.n;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const double L = g.lx;
    const double x = (i < n) ? xe[i] : 0.0, y = (j < n) ? xe[j + 1] : 0.0, z = xe[k + 1];
    double r2min = INFINITY;
    for (int d0 = 0; d0 < ndrops; d0 += 256) {
        const int cnt = min(256, ndrops - d0);
        __syncthreads();
        for (int e = tid; e < cnt * 3; e += 256) sc[e] = centers[(size_t)d0 * 3 + e];
        __syncthreads();
        for (int d = 0; d < cnt; d++) {
            double a = fabs(x - sc[3 * d]), b = fabs(y - sc[3 * d + 1]), c = fabs(z - sc[3 * d + 2]);
            a = fmin(a, L - a); b = fmin(b, L - b); c = fmin(c, L - c);
            r2min = fmin(r2min, a * a + b * b + c * c);
        }
    }
    if (i < n && j < n)
        phi[idx3(i, j, k, n)] = 0.5 * (1.0 - tanh((sqrt(r2min) - radius) / 2 / g.eps));
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform tracers: coordinate d of tracer p from splitmix64(seed, 3p+d).
__global__ __launch_bounds__(256)
void k_init_tracers(double* __restrict__ tag, double* __restrict__ px, double* __restrict__ py,
                    double* __restrict__ pz, double* __restrict__ pu1, double* __restrict__ pv1,
                    double* __restrict__ pw1, long long np, unsigned long long seed, double L) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const unsigned long long base = splitmix64(seed);
    double r[3];
    for (int d = 0; d < 3; d++) {
        const unsigned long long h = splitmix64(base + 3ull * (unsigned long long)p + d);
        r[d] = (double)(h >> 11) * 0x1.0p-53;
    }
    tag[p] = (double)(p + 1);
    px[p] = r[0] * L; py[p] = r[1] * L; pz[p] = r[2] * L;
    pu1[p] = 0.0; pv1[p] = 0.0; pw1[p] = 0.0;
}

int launch_init_tgv(mhit_ctx* c) {
    CellLaunch L(c->n);
    k_init_tgv<<<L.grid, L.block, 0, c->stream>>>(c->u, c->v, c->w, c->xext, c->g);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

int launch_init_drop(mhit_ctx* c) {
    CellLaunch L(c->n);
    k_init_drop<<<L.grid, L.block, 0, c->stream>>>(c->phi, c->xext, c->P.radius, c->g);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

int launch_init_modes(mhit_ctx* c, const double* modes, int nmodes, const double2* E) {
    CellLaunch L(c->n);
    k_init_modes<<<L.grid, L.block, 0, c->stream>>>(c->u, c->v, c->w, modes, nmodes, E, c->g);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

int launch_init_drops(mhit_ctx* c, const double* centers, int ndrops, double radius) {
    CellLaunch L(c->n);
    k_init_drops<<<L.grid, L.block, 0, c->stream>>>(c->phi, c->xext, centers, ndrops,
                                                    radius, c->g);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

int launch_init_tracers(mhit_ctx* c, int64_t n, uint64_t seed) {
    const int blocks = div_up(n, 256);
    k_init_tracers<<<blocks, 256, 0, c->stream>>>(c->ptag, c->px, c->py, c->pz, c->pu1, c->pv1, c->pw1,
                                                  (long long)n, (unsigned long long)seed, c->g.lx);
    MHIT_HIP_TRY(hipGetLastError());
    return MHIT_OK;
}

}  // namespace mhit
