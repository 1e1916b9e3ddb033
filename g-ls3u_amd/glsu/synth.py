"""Synthetic fringe-sequence definition (host side).

phi(x, y, t) = (amp0 + amp1 t) * peaks(X, Y) + 2 pi carrier x, X/Y = linspace(-3, 3)
along columns/rows; I = 128 + 100 cos(phi) + noise * N(0, 1).  The frames
themselves are generated on the GPU (glsu.gen_frames); this module gives the
exact initial phase phi(., ., 0) the DPRA constructor needs.
"""
import numpy as np


def peaks(X, Y):
    return (3.0 * (1.0 - X) ** 2 * np.exp(-X * X - (Y + 1.0) ** 2)
            - 10.0 * (X / 5.0 - X ** 3 - Y ** 5) * np.exp(-X * X - Y * Y)
            - np.exp(-(X + 1.0) ** 2 - Y * Y) / 3.0)


def phase(width, height, amp, carrier):
    X = np.linspace(-3.0, 3.0, width) if width > 1 else np.array([3.0])
    Y = np.linspace(-3.0, 3.0, height) if height > 1 else np.array([3.0])
    XX, YY = np.meshgrid(X, Y)
    return amp * peaks(XX, YY) + 2.0 * np.pi * carrier * np.arange(width)[None, :]
