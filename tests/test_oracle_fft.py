"""Pins the oracle's FP64 mixed-radix FFT (stand-in for cuFFT, range_compression.cu:80-100) against numpy.fft
golden vectors, a brute-force DFT, Parseval and the round trip."""
import numpy as np
import pytest

import pyoracle as po
from conftest import golden


@pytest.fixture(scope="module")
def fftv():
    return golden("fft_vectors.npz")


def test_fft_matches_numpy_golden(fftv):
    for n in fftv["lengths"]:
        x = fftv[f"x_{n}"].copy().reshape(1, -1)
        ref = fftv[f"fwd_{n}"]
        po.fft_rows(x)
        err = np.abs(x[0] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err < 1e-13, (n, err)


def test_inverse_matches_numpy_golden(fftv):
    for n in fftv["lengths"]:
        x = fftv[f"x_{n}"].copy().reshape(1, -1)
        ref = fftv[f"inv_{n}"]
        po.fft_rows(x, inverse=True)
        err = np.abs(x[0] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err < 1e-13, (n, err)


@pytest.mark.parametrize("n", [11, 13, 22, 33, 77, 121])
def test_fft_other_primes_vs_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = po.dft(x)
    # the brute-force DFT itself against the closed form of a single tone
    tone = np.exp(2j * np.pi * 3 * np.arange(n) / n)
    spec = po.dft(tone)
    assert abs(spec[3] - n) < 1e-9 and np.abs(np.delete(spec, 3)).max() < 1e-9
    y = x.copy().reshape(1, -1)
    po.fft_rows(y)
    assert np.abs(y[0] - ref).max() < 1e-10 * n


@pytest.mark.parametrize("n", [4116, 6144, 6400, 16464])
def test_parseval_and_round_trip(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).reshape(1, -1)
    x0 = x.copy()
    po.fft_rows(x)
    assert abs(np.sum(np.abs(x) ** 2) / n - np.sum(np.abs(x0) ** 2)) < 1e-9 * np.sum(np.abs(x0) ** 2)
    po.fft_rows(x, inverse=True)
    assert np.abs(x / n - x0).max() < 1e-12


def test_columns_equal_rows_of_transpose():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((84, 20)) + 1j * rng.standard_normal((84, 20))
    b = np.ascontiguousarray(a.T)
    po.fft_cols(a)
    po.fft_rows(b)
    assert np.abs(a - b.T).max() < 1e-12
