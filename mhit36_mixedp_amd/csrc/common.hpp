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
    int skipped;          // steps run while halted since the host last looked
};

// Launch-shape choices (z-chunk lengths and tile rows of the marching
// passes, XCD-aware tile order), which do not change results, and the form
// of the column transforms, which changes them only by rounding (the same
// DFT with a different operation order).  Defaults are set by grid size at
// create; MHIT_TUNE="ns_mz=64,ac_mz=64,ns_my=16,ac_my=8,xcd=1,sort=10,fft=1,
// ycols=16,zcols=8,xrows=8" overrides them.
struct Tuning {
    int ns_mz = 64;       // planes per workgroup, NS pass (set by grid size at create)
    int ac_mz = 64;       // planes per workgroup, ACDI pass (likewise)
    int xcd = 0;          // XCD-aware tile order (mhit_create: 1 up to 256^3
                          // and for boxes batched in an ensemble, 0 above)
    int sort_every = 10;  // steps between tracer re-orderings (0: never)
    int ns_my = 16;       // tile rows, NS pass: 16 or 8 (64 columns), 32 (32
                          // columns) or 10 (96 columns); mhit_create: 8 for
                          // n <= 64, 10 where n is a multiple of 96 but not
                          // of 64 (96)
    int ac_my = 16;       // tile rows, ACDI pass (the same choices)
    int fft_wave = 1;     // column transforms of eight columns in the wavefront
                          // form (lane/register transposes by permlane and DPP)
                          // where n = 64, 128, 256, 512; 0: register-direct form
    // FFT launch widths (0: the default for the size and precision; n >= 64
    // has compiled alternatives, fft_widths in poisson.hip)
    int fft_ycols = 0;    // columns per workgroup, y passes
    int fft_zcols = 0;    // columns per workgroup, z pass
    int fft_xrows = 0;    // row pairs per workgroup, x passes
    int fft_xy = 0;       // x and y transforms of a z plane in one launch where
                          // compiled (n = 32 ... 128 FP32, 32 ... 96 FP64);
                          // mhit_create: 1 there
    int fft_xplan = 0;    // radix plan of the separate x passes: 0 radix 8 first
                          // (512 = 8.8.8), 1 radix 4 until 64 points remain
                          // (512 = 4.4.4.8); the plans differ at n = 128,
                          // 256, 384, 512
};

// Per-run constants passed by value to kernels.
struct Grid {
    int n;                // points per direction
    double dx, dxi, ddxi, rhoi, rho, mu, dt, lx;
    double eps, epsi, enm, sigma;
    double f1, f2, f3;
};

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// Columns of a fused pass's tile for its Tuning rows (64 x 16, 64 x 8,
// 32 x 32, 96 x 10).
inline int tile_cols(int rows) { return rows == 32 ? 32 : (rows == 10 ? 96 : 64); }

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

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs of the MI355X (linear id mod 8), each with its own L2.  This remap
// gives XCD x the contiguous range [x T/8, (x+1) T/8) of tiles in x-fastest
// order, so the workgroups resident on one XCD at a time cover an x-y patch
// of neighbouring tiles whose halo rows and columns then come from that
// XCD's L2.  Identity when the tile count is not a multiple of 8.
constexpr int kXcds = 8;

// Workgroup coordinates within one box.  A launch may cover several boxes
// of the same size (an ensemble): their grids are stacked along z, so
// gridDim.z = nz * nbox and box = blockIdx.z / nz.  A single box is nbox = 1.
// Because each box's grid is a whole multiple of the per-box workgroup
// count, the XCD that a workgroup lands on (linear id mod 8) is the same as
// in a launch of that box alone whenever that count is a multiple of 8.
struct Blk {
    int x, y, z;      // workgroup coordinates in the box's grid
    int nx, ny, nz;   // the box's grid
    int box;          // index of the box in the launch
};
__device__ __forceinline__ Blk launch_blk(int nbox) {
    Blk b;
    b.x = blockIdx.x;
    b.y = blockIdx.y;
    b.nx = gridDim.x;
    b.ny = gridDim.y;
    b.nz = gridDim.z / nbox;
    b.box = blockIdx.z / b.nz;
    b.z = blockIdx.z - b.box * b.nz;
    return b;
}

// Device pointers taken from an argument struct (possibly read from a table
// in memory) are generic to the compiler, which then emits flat memory
// instructions; gmem() states that they point to global memory.
#define MHIT_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ MHIT_GLOBAL T* gmem(T* p) { return (MHIT_GLOBAL T*)p; }

// Load / store of a two-component HIP vector (float2, double2) through a
// global pointer: the vector type's operators take generic references, so
// the pair moves as a builtin two-element vector.
template <typename V>
__device__ __forceinline__ V gload(const MHIT_GLOBAL V* p) {
    typedef typename V::value_type E __attribute__((ext_vector_type(2)));
    const E t = *(const MHIT_GLOBAL E*)p;
    V r;
    r.x = t.x;
    r.y = t.y;
    return r;
}
template <typename V>
__device__ __forceinline__ void gstore(MHIT_GLOBAL V* p, const V& v) {
    typedef typename V::value_type E __attribute__((ext_vector_type(2)));
    E t;
    t.x = v.x;
    t.y = v.y;
    *(MHIT_GLOBAL E*)p = t;
}

// Accesses of data that a pass touches exactly once (its outputs, the
// correction's and the inverse x transform's inputs): the nontemporal hint
// (`nt`) streams them through the caches instead of letting them displace
// lines that neighbouring workgroups read again (halo rows and columns).
// Measured at 512^3: nontemporal stores in every pass 8.56-8.71 -> 8.34 ms
// per step; nontemporal loads only where they paid (the NS, ACDI and y
// passes read their once-used inputs faster with plain loads).
template <typename T>
__device__ __forceinline__ T ld_once(const MHIT_GLOBAL T* p) {
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void st_once(MHIT_GLOBAL T* p, T v) {
    __builtin_nontemporal_store(v, p);
}
template <typename V>
__device__ __forceinline__ V gload_once(const MHIT_GLOBAL V* p) {
    typedef typename V::value_type E __attribute__((ext_vector_type(2)));
    const E t = ld_once((const MHIT_GLOBAL E*)p);
    V r;
    r.x = t.x;
    r.y = t.y;
    return r;
}
template <typename V>
__device__ __forceinline__ void gstore_once(MHIT_GLOBAL V* p, const V& v) {
    typedef typename V::value_type E __attribute__((ext_vector_type(2)));
    E t;
    t.x = v.x;
    t.y = v.y;
    st_once((MHIT_GLOBAL E*)p, t);
}

__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz, const Blk& b, int enable = 1) {
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    const int total = nx * ny * nz;
    int L = b.x + nx * (b.y + ny * b.z);
    if (enable && total % kXcds == 0) L = (L % kXcds) * (total / kXcds) + L / kXcds;
    bx = L % nx;
    by = (L / nx) % ny;
    bz = L / (nx * ny);
}

// 1/x for a normal x > 0: the hardware estimate (v_rcp_f64) refined by one
// Newton step, r' = r + r (1 - x r) in two FMAs: within 1 ulp of 1/x, in
// three operations instead of the ten of an IEEE division.
__device__ __forceinline__ double rcp_refined(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// 1/sqrt(x) for a normal x > 0: the hardware estimate (v_rsq_f64) refined
// by one Newton step, r' = r + r/2 (1 - x r^2), in four operations.
__device__ __forceinline__ double rsq_refined(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    return fma(0.5 * r, fma(-x * r, r, 1.0), r);
}

// ln(a/b) for a, b > 0 without rounding the quotient first, in ~40 VALU
// operations (one reciprocal) instead of a division plus a library log.
// Method: a = ma 2^ea, b = mb 2^eb (frexp); move a factor 2 between ma and
// the exponent so that q = ma/mb lies in [sqrt(1/2), sqrt(2)); then
// ln(a/b) = (ea - eb) ln 2 + 2 atanh(s), s = (ma - mb)/(ma + mb), with the
// series 2 atanh(s) = 2 sum_k s^(2k+1)/(2k+1) (Abramowitz & Stegun 4.1.27)
// truncated after s^23 (|s| <= 3 - 2 sqrt 2, remainder < 1e-19 relative).
// ma - mb is exact (Sterbenz); ln 2 is split so that (ea - eb) ln2_hi is
// exact.  Error <= 2 ulp of the exact ln(a/b) (checked with every FMA
// emulated exactly and the hardware reciprocal estimate anywhere within
// 2 ulp of 1/(ma + mb)).  a <= 0 gives -inf (a = 0)
// or NaN, as log(a/b) would.
// fma(a, b, c) in FP64 as the three-address VOP3 instruction (the same
// IEEE fused multiply-add).  Left to itself the compiler emits the
// two-address v_fmac_f64 and, when the addend must survive (a polynomial
// coefficient held across a loop), first copies it into the destination.
__device__ __forceinline__ double fma3(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ double ln_ratio(double a, double b) {
    constexpr double LN2_HI = 0x1.62e42fee00000p-1;   // ln 2 to 32 bits
    constexpr double LN2_LO = 0x1.a39ef35793c76p-33;  // ln 2 - LN2_HI
    int ea, eb;
    double ma = frexp(a, &ea);
    const double mb = frexp(b, &eb);
    // q = ma/mb into [sqrt(1/2), sqrt(2)): a power-of-two rescale of ma
    // (exact), selected without a branch
    const int d = (ma < mb * 0.70710678118654752440) ? 1 : ((ma >= mb * 1.41421356237309504880) ? -1 : 0);
    ma = ldexp(ma, d);
    ea -= d;
    const double s = (ma - mb) * rcp_refined(ma + mb);
    const double t = s * s;
    double p = 1.0 / 23.0;
    p = fma3(p, t, 1.0 / 21.0);
    p = fma3(p, t, 1.0 / 19.0);
    p = fma3(p, t, 1.0 / 17.0);
    p = fma3(p, t, 1.0 / 15.0);
    p = fma3(p, t, 1.0 / 13.0);
    p = fma3(p, t, 1.0 / 11.0);
    p = fma3(p, t, 1.0 / 9.0);
    p = fma3(p, t, 1.0 / 7.0);
    p = fma3(p, t, 1.0 / 5.0);
    p = fma3(p, t, 1.0 / 3.0);
    const double s2 = 2.0 * s;
    const double E = (double)(ea - eb);
    double r = fma(E, LN2_HI, fma(E, LN2_LO, fma(s2 * t, p, s2)));
    if (!(a > 0.0)) r = (a == 0.0) ? -INFINITY : __builtin_nan("");
    return r;
}

// Workgroup barrier for LDS hand-off only.  __syncthreads() carries a
// workgroup release fence, which on CDNA also waits for every outstanding
// global load (vmcnt(0)); the marching passes keep the next plane's global
// loads in flight across barriers, so they synchronise on LDS traffic
// alone: wait for this wave's LDS operations, then s_barrier.  The memory
// clobber stops the compiler from moving memory accesses across it; loads
// into registers are still waited for before their first use.
__device__ __forceinline__ void lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

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
