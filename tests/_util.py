"""Shared helpers for the tests (error metrics, case printing)."""
import sys

import numpy as np


def wrap(a):
    return np.angle(np.exp(1j * np.asarray(a, np.float64)))


def wrapped_err(a, b):
    e = wrap(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return float(np.sqrt(np.mean(e * e))), float(np.max(np.abs(e)))


def say(*args):
    """Name the case before each GPU step (flushed) so a crash log identifies it."""
    print(*args, flush=True)
    sys.stdout.flush()


def aia_frames(M, H, W, seed, noise=5.0, amp=3.0, carrier=0.05, dtype=np.float64):
    """Randomly phase-shifted fringe frames for AIA: I = 128 + 100 cos(phi +
    delta_k) + noise, phi = amp * peaks + 2 pi carrier x (the synthetic phase of
    the DPRA workloads), delta_k uniform in (-pi, pi].  uint8 frames are rounded
    and clipped.  Returns (frames M x H x W, phi, delta)."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(-3, 3, W), np.linspace(-3, 3, H))
    peaks = (3 * (1 - X) ** 2 * np.exp(-X ** 2 - (Y + 1) ** 2)
             - 10 * (X / 5 - X ** 3 - Y ** 5) * np.exp(-X ** 2 - Y ** 2)
             - np.exp(-(X + 1) ** 2 - Y ** 2) / 3)
    phi = amp * peaks + 2 * np.pi * carrier * np.arange(W)[None, :]
    delta = rng.uniform(-np.pi, np.pi, M)
    f = 128 + 100 * np.cos(phi[None] + delta[:, None, None]) + rng.normal(0, noise, (M, H, W))
    if np.dtype(dtype) == np.uint8:
        f = np.clip(np.rint(f), 0, 255)
    return f.astype(dtype), wrap(phi), delta
