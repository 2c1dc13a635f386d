// spectra.hip -- kinetic-energy spectrum of the live velocity field.
//
// The reference leaves spectra to post-processing on the dumps
// (python_utilities/plot_fields.py:116-151, a 1-D FFT of one profile) and
// the dissipation to post/compute_diss.  Here the 3-D shell spectrum
//   E(k) = 1/2 sum_{k - 1/2 <= |m| < k + 1/2} (|u^(m)|^2 + |v^(m)|^2 + |w^(m)|^2) / N^2,
// N = n^3, is computed on the device with the context's own R2C FFT (in the
// Poisson precision), so that sum_k E(k) is the kinetic energy
// 1/2 <u^2 + v^2 + w^2> of mhit_diag.  Each half-spectrum mode with
// 0 < kx < n/2 stands for itself and its Hermitian partner (weight 2); the
// kx = 0 and kx = n/2 (Nyquist) modes count once.  The spectrum buffer is
// scratch between steps.
#include "fft.hpp"
#include "internal.hpp"

#include <cmath>
#include <mutex>
#include <string>

namespace mhit {
int sync_scalars_public(mhit_ctx* c);
}

using namespace mhit;

namespace {

__device__ __forceinline__ int signed_wavenumber(int m, int n) { return m <= n / 2 ? m : m - n; }

// One block per (ky, kz) row of the half spectrum (plus its Nyquist mode);
// shells accumulated in LDS, then added to the global bins.
template <typename V>
__global__ void k_shell_bins(const V* __restrict__ spec, const V* __restrict__ nyq, int n,
                             double scale, double* __restrict__ bins, int nbins) {
    extern __shared__ double sb[];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) sb[b] = 0.0;
    __syncthreads();
    const int ky = blockIdx.x, kz = blockIdx.y, h = n / 2;
    const int sy = signed_wavenumber(ky, n), sz = signed_wavenumber(kz, n);
    const size_t row = ((size_t)kz * n + ky);
    for (int kx = threadIdx.x; kx <= h; kx += blockDim.x) {
        const V f = (kx < h) ? spec[row * h + kx] : nyq[row];
        const double re = (double)f.x, im = (double)f.y;
        const double w = (kx == 0 || kx == h) ? 1.0 : 2.0;
        const double k = sqrt((double)kx * kx + (double)sy * sy + (double)sz * sz);
        const int b = (int)floor(k + 0.5);
        if (b < nbins) atomicAdd(&sb[b], w * scale * (re * re + im * im));
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (sb[b] != 0.0) atomicAdd(&bins[b], sb[b]);
}

// bins[0, nb) = 0 (a fill kernel rather than a memset node, so the same
// launch sequence can sit in a captured graph)
__global__ void k_zero_bins(double* __restrict__ bins, int nb) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) bins[b] = 0.0;
}

// Number of shells the transform of an n^3 box fills.
int shells_used(int n) { return (int)std::floor(std::sqrt(3.0) * (n / 2) + 0.5) + 1; }

// The spectrum with its shell sums in bins[0, min(nbins, shells)) (device
// memory the caller provides); E is written directly, no host staging.
int spectrum_into(mhit_ctx* c, int32_t nbins, double* E, int32_t* nshells, double* bins) {
    const int n = c->n;
    const int used = shells_used(n);
    if (nshells) *nshells = used;
    const int nb = nbins < used ? nbins : used;
    int rc = MHIT_OK;
    k_zero_bins<<<div_up(nb, 256), 256, 0, c->stream>>>(bins, nb);
    hipError_t err = hipGetLastError();
    const double N = (double)n * n * n;
    const double scale = 0.5 / (N * N);
    double e0 = 0.0;   // the k = 0 shell, from the true mean
    double* fields[3] = {c->u, c->v, c->w};
    for (int f = 0; f < 3 && err == hipSuccess && rc == MHIT_OK; f++) {
        if ((rc = fft_r2c_dev(c, fields[f]))) break;
        dim3 grid(n, n);
        if (c->fp32)
            k_shell_bins<float2><<<grid, 256, nb * sizeof(double), c->stream>>>(
                (const float2*)c->spec, (const float2*)c->nyq, n, scale, bins, nb);
        else
            k_shell_bins<double2><<<grid, 256, nb * sizeof(double), c->stream>>>(
                (const double2*)c->spec, (const double2*)c->nyq, n, scale, bins, nb);
        err = hipGetLastError();
        // the device keeps u + mean (deferred mean removal): the k = 0 mode
        // is rebuilt from the stored sum minus the mean
        double f0[2] = {0.0, 0.0};
        if (err == hipSuccess) {
            if (c->fp32) {
                float t[2];
                err = hipMemcpyAsync(t, c->spec, sizeof(t), hipMemcpyDeviceToHost, c->stream);
                if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
                f0[0] = t[0];
            } else {
                err = hipMemcpyAsync(f0, c->spec, sizeof(f0), hipMemcpyDeviceToHost, c->stream);
                if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
            }
        }
        const double m = f0[0] / N - c->S_host.mean[f];
        e0 += 0.5 * m * m;
    }
    if (err == hipSuccess && rc == MHIT_OK)
        err = hipMemcpyAsync(E, bins, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (err == hipSuccess && rc == MHIT_OK) err = hipStreamSynchronize(c->stream);
    if (rc) return rc;
    MHIT_HIP_TRY(err);
    E[0] = e0;
    for (int b = nb; b < nbins; b++) E[b] = 0.0;
    return MHIT_OK;
}

int spectrum_check(mhit_ctx* c, int32_t nbins, double* E) {
    if (!c || nbins <= 0 || !E) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    int rc;
    if ((rc = sync_scalars_public(c))) return rc;
    if (c->S_host.halted) { set_error("stopped: cou > 7"); return MHIT_ERR_CFL; }
    return MHIT_OK;
}

// Shell sums of mhit_energy_spectrum_noalloc: one array per device, held
// until the process ends, shared by every call on that device (calls that
// overlap on one device race on it).  Host code never uses the array's
// name as an address: each device's copy is looked up once with
// hipGetSymbolAddress (with that device current) and kept.
constexpr int kScratchBins = MHIT_SPECTRUM_NOALLOC_MAX_BINS;
static __device__ __align__(256) double scratch[kScratchBins];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_on(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) { set_error("device ordinal out of range"); return MHIT_ERR_ARG; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        int cur = -1;
        MHIT_HIP_TRY(hipGetDevice(&cur));
        if (cur != device) MHIT_HIP_TRY(hipSetDevice(device));
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (cur != device) MHIT_HIP_TRY(hipSetDevice(cur));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return MHIT_ERR_HIP;
        }
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return MHIT_OK;
}

}  // namespace

extern "C" {

int mhit_energy_spectrum(mhit_ctx* c, int32_t nbins, double* E, int32_t* nshells) {
    int rc;
    if ((rc = spectrum_check(c, nbins, E))) return rc;
    const int used = shells_used(c->n);
    const int nb = nbins < used ? nbins : used;
    double* bins = nullptr;
    MHIT_HIP_TRY(hipMalloc((void**)&bins, (size_t)nb * sizeof(double)));
    rc = spectrum_into(c, nbins, E, nshells, bins);
    (void)hipFree(bins);
    return rc;
}

int mhit_energy_spectrum_noalloc(mhit_ctx* c, int32_t nbins, double* E, int32_t* nshells) {
    int rc;
    if ((rc = spectrum_check(c, nbins, E))) return rc;
    const int used = shells_used(c->n);
    const int nb = nbins < used ? nbins : used;
    if (nb > kScratchBins) {
        set_error("mhit_energy_spectrum_noalloc: " + std::to_string(nb) + " shells exceed the " +
                  std::to_string(kScratchBins) + "-shell scratch (ask for fewer bins)");
        return MHIT_ERR_ARG;
    }
    double* bins = nullptr;
    if ((rc = scratch_on(c->device, &bins))) return rc;
    return spectrum_into(c, nbins, E, nshells, bins);
}

}  // extern "C"
