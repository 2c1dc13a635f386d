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
#include <vector>

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

}  // namespace

extern "C" {

int mhit_energy_spectrum(mhit_ctx* c, int32_t nbins, double* E, int32_t* nshells) {
    if (!c || nbins <= 0 || !E) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    int rc;
    if ((rc = sync_scalars_public(c))) return rc;
    if (c->S_host.halted) { set_error("stopped: cou > 7"); return MHIT_ERR_CFL; }
    const int n = c->n;
    const int used = (int)std::floor(std::sqrt(3.0) * (n / 2) + 0.5) + 1;
    if (nshells) *nshells = used;
    const int nb = nbins < used ? nbins : used;
    double* bins = nullptr;
    MHIT_HIP_TRY(hipMalloc((void**)&bins, (size_t)nb * sizeof(double)));
    hipError_t err = hipMemsetAsync(bins, 0, (size_t)nb * sizeof(double), c->stream);
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
    std::vector<double> h(nb);
    if (err == hipSuccess && rc == MHIT_OK)
        err = hipMemcpyAsync(h.data(), bins, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (err == hipSuccess && rc == MHIT_OK) err = hipStreamSynchronize(c->stream);
    (void)hipFree(bins);
    if (rc) return rc;
    MHIT_HIP_TRY(err);
    h[0] = e0;
    for (int b = 0; b < nbins; b++) E[b] = b < nb ? h[b] : 0.0;
    return MHIT_OK;
}

}  // extern "C"
