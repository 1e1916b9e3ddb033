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
