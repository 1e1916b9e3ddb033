"""Pins of the oracle's FP64 FFT (the stand-in for MKL's FFTW3 API used at
WFT/WFT2_CPU.cpp:745,841-843) against numpy.fft (pocketfft) and against the
committed golden vectors tests/golden/fft_vectors.npz (made by
tests/golden/make_fft_vectors.py with numpy)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fft_vectors.npz")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 16, 20, 27, 30, 49, 64, 98,
                               120, 125, 128, 320, 343, 576, 840, 1120, 1152, 11, 13, 97])
def test_fft1d_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = max(1.0, np.sqrt(n))
    assert np.abs(oracle.fft1d(x, -1) - np.fft.fft(x)).max() < 1e-13 * scale * np.log2(n + 1) * 10
    assert np.abs(oracle.fft1d(x, +1) - np.fft.ifft(x) * n).max() < 1e-13 * scale * np.log2(n + 1) * 10


@pytest.mark.parametrize("r,c", [(1, 1), (2, 3), (12, 20), (30, 28), (64, 80)])
def test_fft2d_matches_numpy(oracle, r, c):
    rng = np.random.default_rng(r * 100 + c)
    x = rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))
    assert np.abs(oracle.fft2d(x, -1) - np.fft.fft2(x)).max() < 1e-11
    assert np.abs(oracle.fft2d(x, +1) - np.fft.ifft2(x) * r * c).max() < 1e-11


def test_fft_golden_vectors(oracle):
    g = np.load(GOLD)
    for name in g.files:
        if not name.startswith("x_"):
            continue
        n = name[2:]
        x = g[name]
        assert np.abs(oracle.fft1d(x, -1) - g["X_" + n]).max() < 1e-10
        assert np.abs(oracle.fft1d(x, +1) - g["Y_" + n]).max() < 1e-10


def test_fft_impulse_and_linearity(oracle):
    # closed forms: DFT of a unit impulse at p is exp(-2 pi i k p / n); linearity
    n = 56
    p = 5
    d = np.zeros(n, complex)
    d[p] = 1
    k = np.arange(n)
    assert np.abs(oracle.fft1d(d) - np.exp(-2j * np.pi * k * p / n)).max() < 1e-14
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(n) + 0j, rng.standard_normal(n) * 1j
    assert np.abs(oracle.fft1d(2 * a + 3 * b) - (2 * oracle.fft1d(a) + 3 * oracle.fft1d(b))).max() < 1e-12
