"""The reference's matrix files, through libglsu's readers and writers.

Binary matrix (Utils/matrixIO.inl:33-104, header Utils/matrixIO.h:18-35):
int32 rows, int32 cols, row-major elements (the reference's phase maps,
Experiments/imgs/*/*_phi.bin, are float32).  Complex text
(Utils/Utils.cpp:80-193): "rows,cols" then "re,im" per element.
"""
import ctypes
import os

import numpy as np

from ._lib import check, lib


def read_bin(path, dtype=np.float32):
    """(rows, cols) array of `dtype` from a binary matrix file."""
    dt = np.dtype(dtype)
    r, c = ctypes.c_int(), ctypes.c_int()
    p = os.fsencode(path)
    check(lib().glsu_read_matrix_bin(p, dt.itemsize, ctypes.byref(r), ctypes.byref(c), None, 0))
    a = np.empty((r.value, c.value), dt)
    check(lib().glsu_read_matrix_bin(p, dt.itemsize, ctypes.byref(r), ctypes.byref(c),
                                     a.ctypes.data, a.nbytes))
    return a


def write_bin(path, a):
    a = np.ascontiguousarray(a)
    if a.ndim != 2:
        raise ValueError("a must be 2-D")
    check(lib().glsu_write_matrix_bin(os.fsencode(path), a.dtype.itemsize, a.shape[0], a.shape[1],
                                      a.ctypes.data))


def read_fp(path):
    """(rows, cols) complex128 array from a complex text file."""
    r, c = ctypes.c_int(), ctypes.c_int()
    p = os.fsencode(path)
    check(lib().glsu_read_complex_text(p, ctypes.byref(r), ctypes.byref(c), None, 0))
    a = np.empty((r.value, c.value), np.complex128)
    check(lib().glsu_read_complex_text(p, ctypes.byref(r), ctypes.byref(c), a.ctypes.data,
                                       2 * a.size))
    return a


def write_fp(path, a, precision=0):
    """precision <= 0 keeps the reference's stream default (6 significant digits)."""
    a = np.ascontiguousarray(a, np.complex128)
    if a.ndim != 2:
        raise ValueError("a must be 2-D")
    check(lib().glsu_write_complex_text(os.fsencode(path), a.shape[0], a.shape[1], a.ctypes.data,
                                        int(precision)))
