"""Writes the golden fixtures that pin the FP64 oracle. Run in the build container (numpy/scipy only):

    python tests/golden/make_golden.py

Nothing from the reference is executed: the reference's own tests hold no vectors for the focusing path, so these
fixtures come from independent library implementations of the same mathematics (numpy.fft, numpy.polyfit,
numpy.blackman, scipy.signal.windows.hamming) and from closed forms.
"""
import os

import numpy as np
import scipy.signal

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_vectors():
    rng = np.random.default_rng(20261202)
    lengths = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 20, 21, 25, 27, 28, 32, 35, 42, 49, 56, 60, 64, 84,
               96, 105, 112, 120, 128, 147, 168, 180, 210, 243, 256, 343, 360, 384, 625, 1000, 1024, 2401]
    out = {}
    for n in lengths:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x_{n}"] = x
        out[f"fwd_{n}"] = np.fft.fft(x)
        out[f"inv_{n}"] = np.fft.ifft(x) * n  # unnormalised inverse
    out["lengths"] = np.array(lengths)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


def polyfit_vectors():
    rng = np.random.default_rng(7)
    out = {}
    # Doppler-centroid shaped: 33 block centres of a 5616-sample row, degree 2
    x = np.array([(xb + xb + 175) // 2 for xb in range(0, 5616, 175)], dtype=np.float64)
    y = 150.0 + 0.01 * x - 2e-6 * x * x + rng.normal(0, 0.5, x.size)
    out["dc_x"], out["dc_y"], out["dc_c"] = x, y, np.polyfit(x, y, 2)
    # SRGR shaped: 100 ground distances up to 100 km, degree 4
    g = np.linspace(0.0, 1.0e5, 100) + rng.uniform(0, 10, 100)
    s = 0.39 * g + 1.1e-7 * g * g - 2e-13 * g ** 3 + 1e-19 * g ** 4
    out["sr_x"], out["sr_y"], out["sr_c"] = g, s, np.polyfit(g, s, 4)
    # small random fits
    for d in (0, 1, 3, 5):
        xr = np.sort(rng.uniform(-3, 3, 20))
        yr = rng.standard_normal(20)
        out[f"r{d}_x"], out[f"r{d}_y"], out[f"r{d}_c"] = xr, yr, np.polyfit(xr, yr, d)
    np.savez_compressed(os.path.join(HERE, "polyfit_vectors.npz"), **out)


def windows():
    out = {
        "blackman16": np.blackman(16),
        "hamming_periodic_704": scipy.signal.windows.hamming(704, sym=False),
        "hamming_periodic_521": scipy.signal.windows.hamming(521, sym=False),
    }
    np.savez_compressed(os.path.join(HERE, "windows.npz"), **out)


def chirp_vectors():
    """The ERS and ASAR range chirps from the closed form, evaluated independently in numpy (FP64)."""
    out = {}
    for name, fs, T, B in (("ers", 18.96e6, 37.12e-6, 15.55e6), ("asar", 19.208e6, 27.1e-6, 16e6)):
        n = int(np.floor(T * fs + 0.5))
        kr = B / T
        i = np.arange(n)
        t = (i - n // 2) / fs
        c = np.exp(1j * np.pi * kr * t * t)
        w = 0.54 - 0.46 * np.cos(2 * np.pi * i / n)
        cw = c * w
        out[f"{name}_plain"] = c / np.sqrt(n)
        out[f"{name}_windowed"] = cw / np.sqrt(np.sum(np.abs(cw) ** 2))
        out[f"{name}_params"] = np.array([fs, T, B, n, kr])
    np.savez_compressed(os.path.join(HERE, "chirp_vectors.npz"), **out)


if __name__ == "__main__":
    fft_vectors()
    polyfit_vectors()
    windows()
    chirp_vectors()
    print("golden fixtures written to", HERE)
