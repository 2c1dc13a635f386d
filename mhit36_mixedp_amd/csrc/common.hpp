// common.hpp -- shared definitions of the MHIT36 HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>

#include "../../include/mhit36.h"

namespace mhit {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);

#define MHIT_HIP_TRY(expr)                                                   \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            ::mhit::set_error(std::string(#expr) + ": " +                    \
                              hipGetErrorString(e_));                        \
            return MHIT_ERR_HIP;                                             \
        }                                                                    \
    } while (0)

// ----------------------------------------------------------- device state
// Scalars that the reference keeps on the host between kernels
// (module.f90:59-61, main.f90:440-444, 801-802, 977-1003).  Here they live
// in device memory so that a whole step can run without a host round trip
// (and be captured in a graph).
struct DevScalars {
    double alpha, beta;   // AB2 coefficients of the step about to run
    double gumax;         // signed max of u,v,w after the last step (gamma)
    double cou;           // gumax*dt*dxi of the last step
    double mean[3];       // means removed by the last correction: the stored
                          // velocity is u_stored = u + mean (subtracted on load)
    long long step;       // completed steps
    int halted;           // set when cou > 7 (cfl_abort); kernels then no-op
    int pad_;
};

// Per-run constants passed by value to kernels.
struct Grid {
    int n;                // points per direction
    double dx, dxi, ddxi, rhoi, rho, mu, dt, lx;
    double eps, epsi, enm, sigma;
    double f1, f2, f3;
};

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// periodic neighbours
__device__ __forceinline__ int wrap_p(int i, int n) { return i + 1 == n ? 0 : i + 1; }
__device__ __forceinline__ int wrap_m(int i, int n) { return i == 0 ? n - 1 : i - 1; }

__device__ __forceinline__ size_t idx3(int i, int j, int k, int n) {
    return ((size_t)k * n + (size_t)j) * n + (size_t)i;
}

// Standard launch shape for cell kernels: 64 x 4 threads over (x, y), one
// block row per z plane.
struct CellLaunch {
    dim3 grid, block;
    explicit CellLaunch(int n) : grid(div_up(n, 64), div_up(n, 4), n), block(64, 4, 1) {}
};

// Block-level reduction helpers (wavefront size 64 on CDNA).
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace mhit
