"""Write the FFT golden vectors used by the CPU and GPU tests.

The reference's own FFT is cuFFT (multi/main.f90:188-205, 853-912), which is
closed and absent here; its published algorithm is the unnormalised DFT
X[k] = sum_j x[j] exp(-2 pi i jk/N) (forward) and exp(+2 pi i jk/N)
(inverse).  numpy.fft (pocketfft) computes the same DFT, so it writes the
vectors.  The validation curves val/cudec_128_ab.txt, val/single_128_ab.txt
and val/single_256_ab.txt (TGV dissipation vs t at Re 1600, 128^3 and
256^3, AB2) are copied from the reference as data
(tests/test_gpu_validation.py).

Run:  python tests/golden/make_golden.py   (writes fft_golden.npz)
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20241202)
    out = {}
    for n in (8, 12, 16, 24, 48, 96):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"fft1d_in_{n}"] = x
        out[f"fft1d_fwd_{n}"] = np.fft.fft(x)
        out[f"fft1d_inv_{n}"] = np.fft.ifft(x) * n
    for n in (8, 12):
        x = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
        out[f"fft3d_in_{n}"] = x
        out[f"fft3d_fwd_{n}"] = np.fft.fftn(x)
    for n in (16, 24):
        x = rng.standard_normal((n, n, n))
        out[f"rfft3d_in_{n}"] = x
        out[f"rfft3d_fwd_{n}"] = np.fft.rfftn(x)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "fft_golden.npz"))


if __name__ == "__main__":
    main()
