"""Writes tests/golden/fft_vectors.npz: inputs and numpy.fft (pocketfft)
outputs, forward and unnormalised backward, on FFT sizes the oracle's padding
rule produces (2^a 3^b 5^c 7^d) plus a prime.  Run: python tests/golden/make_fft_vectors.py"""
import os

import numpy as np

out = {}
rng = np.random.default_rng(20240607)
for n in (32, 60, 98, 320, 1120, 17):
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out[f"x_{n}"] = x
    out[f"X_{n}"] = np.fft.fft(x)
    out[f"Y_{n}"] = np.fft.ifft(x) * n
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_vectors.npz"), **out)
