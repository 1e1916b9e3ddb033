"""ctypes wrapper around oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The FP64 CPU oracle restates the reference's CPU path (DPRA/dpra_cpu.cpp +
WFT/WFT2_CPU.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker / CPU baseline.
The product (libglsu) never routes through it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile liboracle.so from oracle/glsu_oracle.c (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        D, I, U64 = ctypes.c_double, ctypes.c_int, ctypes.c_uint64
        V = ctypes.c_void_p
        sig = {
            "orc_opt_table": (I, [_ip, I]),
            "orc_first_greater": (I, [I]),
            "orc_opt_size": (I, [I]),
            "orc_pair_count": (I, [D, D, D]),
            "orc_fft1d": (I, [I, _dp, I]),
            "orc_fft2d": (I, [I, I, _dp, I]),
            "orc_hash3": (U64, [U64, U64, U64]),
            "orc_normal": (D, [U64, U64, U64]),
            "orc_peaks": (D, [D, D]),
            "orc_gen_phase": (None, [I, I, D, D, _dp]),
            "orc_gen_frame": (None, [I, I, I, U64, D, D, D, D, _dp, _dp]),
            "orc_wft2_create": (V, [I, I, D, D, D, D, D, D, D, D, D, I, I]),
            "orc_wft2_destroy": (None, [V]),
            "orc_wft2_padded": (I, [V, _ip, _ip]),
            "orc_wft2_halfwin": (I, [V, _ip, _ip]),
            "orc_wft2_pairs": (I, [V, _ip, _ip]),
            "orc_wft2_thr": (D, [V]),
            "orc_wft2_compute": (I, [V, _dp, _dp]),
            "orc_wfr2_compute": (I, [V, _dp, _dp]),
            "orc_dpra_create": (V, [_dp, I, I, I, D, D, D, D, D, D, D, D, D, I, I]),
            "orc_dpra_destroy": (None, [V]),
            "orc_dpra_ls": (I, [V, _dp, _dp]),
            "orc_dpra_frame": (I, [V, _dp, _dp, _dp]),
            "orc_dpra_update_ref": (None, [V]),
            "orc_dpra_set_state": (None, [V, _dp, _dp]),
            "orc_dpra_run": (I, [V, _dp, I, _dp]),
            "orc_solve3_sym": (I, [_dp, _dp]),
            "orc_dpra_thr": (D, [V]),
            "orc_wrapped_err": (None, [_dp, _dp, ctypes.c_int64, _dp, _dp]),
            "orc_aia_chol3_solve": (I, [_dp, _dp]),
            "orc_aia_phi": (I, [_dp, I, ctypes.c_int64, _dp, _dp, _dp, I]),
            "orc_aia_delta": (I, [_dp, I, ctypes.c_int64, _dp, _dp, I]),
            "orc_aia": (I, [_dp, I, ctypes.c_int64, _dp, I, D, I, _dp, _ip, _dp, I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
def opt_table():
    buf = (ctypes.c_int * 512)()
    n = lib().orc_opt_table(buf, 512)
    return [buf[i] for i in range(n)]


def first_greater(x):
    return lib().orc_first_greater(int(x))


def opt_size(x):
    return lib().orc_opt_size(int(x))


def pair_count(wl, wh, wi):
    return lib().orc_pair_count(float(wl), float(wh), float(wi))


def fft1d(x, sign=-1):
    a = _c128(x).copy()
    lib().orc_fft1d(a.size, a.view(np.float64).ctypes.data_as(_dp), int(sign))
    return a


def fft2d(x, sign=-1):
    a = _c128(x).copy()
    r, c = a.shape
    lib().orc_fft2d(r, c, a.view(np.float64).ctypes.data_as(_dp), int(sign))
    return a


def normal(seed, t, p):
    return lib().orc_normal(seed, t, p)


def gen_phase(W, H, amp, carrier):
    phi = np.empty((H, W), np.float64)
    lib().orc_gen_phase(W, H, float(amp), float(carrier), _ptr(phi))
    return phi


def gen_frame(W, H, t, seed, amp0=3.0, amp1=0.1, carrier=0.05, noise=10.0):
    """Return (I, phi) float64 arrays of shape (H, W) for frame t."""
    I = np.empty((H, W), np.float64)
    phi = np.empty((H, W), np.float64)
    lib().orc_gen_frame(W, H, int(t), int(seed), float(amp0), float(amp1),
                        float(carrier), float(noise), _ptr(I), _ptr(phi))
    return I, phi


def gen_sequence(W, H, T, seed, amp0=3.0, amp1=0.1, carrier=0.05, noise=10.0,
                 dtype=np.float64):
    """Frames (T, H, W) in `dtype` plus phi(.,.,0) in float64."""
    frames = np.empty((T, H, W), dtype)
    phi0 = None
    for t in range(T):
        I, phi = gen_frame(W, H, t, seed, amp0, amp1, carrier, noise)
        frames[t] = I.astype(dtype)
        if t == 0:
            phi0 = phi
    return frames, phi0


class WFF2:
    """FP64 WFF2 (WFT2_cpu in WFF mode).  pair_parallel: threads take (wy, wx)
    pairs rather than wy rows (same arithmetic per pair)."""

    def __init__(self, W, H, sigx=10.0, wxl=-0.2, wxh=0.2, wxi=0.1,
                 sigy=None, wyl=None, wyh=None, wyi=None, thr=-1.0,
                 nthreads=None, separable=True, pair_parallel=False):
        sigy = sigx if sigy is None else sigy
        wyl = wxl if wyl is None else wyl
        wyh = wxh if wyh is None else wyh
        wyi = wxi if wyi is None else wyi
        nthreads = nthreads or min(os.cpu_count() or 1, 16)
        self.W, self.H = W, H
        flags = (1 if separable else 0) | (2 if pair_parallel else 0)
        self._h = lib().orc_wft2_create(W, H, sigx, wxl, wxh, wxi, sigy, wyl, wyh,
                                        wyi, thr, nthreads, flags)
        if not self._h:
            raise ValueError("FFT padding is out of range [4096]")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_wft2_destroy(self._h)
            self._h = None

    @property
    def padded(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().orc_wft2_padded(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    @property
    def halfwin(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().orc_wft2_halfwin(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    @property
    def pairs(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().orc_wft2_pairs(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    @property
    def thr(self):
        return lib().orc_wft2_thr(self._h)

    def __call__(self, f):
        f = _c128(f)
        assert f.shape == (self.H, self.W)
        out = np.empty_like(f)
        lib().orc_wft2_compute(self._h, f.view(np.float64).ctypes.data_as(_dp),
                               out.view(np.float64).ctypes.data_as(_dp))
        return out


WFR_FIELDS = ("wx", "wy", "phase", "phase_comp", "b", "r", "cxx", "cyy")


class WFR2(WFF2):
    """FP64 WFR2 (WFT2_cpu in WFR mode): returns a dict of the eight result
    maps (WFT/WFT.h:52-60)."""

    def __call__(self, f):
        f = _c128(f)
        assert f.shape == (self.H, self.W)
        out = np.empty((8, self.H, self.W))
        rc = lib().orc_wfr2_compute(self._h, f.view(np.float64).ctypes.data_as(_dp), _ptr(out))
        if rc:
            raise ValueError("curvature padding is out of range [4096]")
        return dict(zip(WFR_FIELDS, out))


class DPRA:
    """FP64 DPRA_CPU chain (LS + WFF2 + update).  pair_parallel: WFF2 threads
    take (wy, wx) pairs rather than wy rows (same arithmetic; for checking
    long sequences -- the CPU baseline keeps the reference's wy loop)."""

    def __init__(self, phi0, rr=1, sigx=10.0, wxl=-0.2, wxh=0.2, wxi=0.1,
                 sigy=None, wyl=None, wyh=None, wyi=None, thr=-1.0,
                 nthreads=None, float_cast=True, pair_parallel=False):
        phi0 = _f64(phi0)
        H, W = phi0.shape
        sigy = sigx if sigy is None else sigy
        wyl = wxl if wyl is None else wyl
        wyh = wxh if wyh is None else wyh
        wyi = wxi if wyi is None else wyi
        nthreads = nthreads or min(os.cpu_count() or 1, 16)
        self.W, self.H = W, H
        self._h = lib().orc_dpra_create(_ptr(phi0), W, H, rr, sigx, wxl, wxh, wxi,
                                        sigy, wyl, wyh, wyi, thr, nthreads,
                                        (0 if float_cast else 1) | (2 if pair_parallel else 0))
        if not self._h:
            raise ValueError("invalid DPRA parameters or FFT padding out of range")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_dpra_destroy(self._h)
            self._h = None

    @property
    def thr(self):
        return lib().orc_dpra_thr(self._h)

    def ls(self, frame):
        f = _f64(frame)
        z = np.empty(f.shape, np.complex128)
        lib().orc_dpra_ls(self._h, _ptr(f), z.view(np.float64).ctypes.data_as(_dp))
        return z

    def dpra_per_frame(self, frame):
        f = _f64(frame)
        dphi = np.empty(f.shape, np.float64)
        phic = np.empty(f.shape, np.float64)
        lib().orc_dpra_frame(self._h, _ptr(f), _ptr(dphi), _ptr(phic))
        return dphi, phic

    def update_ref_phi(self):
        lib().orc_dpra_update_ref(self._h)

    def set_state(self, phiref=None, dphiref=None):
        """Overwrite the reference phase and/or the cumulative phase change."""
        a = None if phiref is None else _f64(phiref)
        b = None if dphiref is None else _f64(dphiref)
        for v in (a, b):
            assert v is None or v.shape == (self.H, self.W)
        lib().orc_dpra_set_state(self._h, _ptr(a) if a is not None else None,
                                 _ptr(b) if b is not None else None)

    def run(self, frames):
        fr = _f64(frames)
        T = fr.shape[0]
        out = np.empty(fr.shape, np.float64)
        lib().orc_dpra_run(self._h, _ptr(fr), T, _ptr(out))
        return out


def solve3_sym(A, b):
    """A: 3x3 (upper triangle used), b: 3 -> (x, singular)."""
    Acm = np.ascontiguousarray(np.asarray(A, np.float64).T).ravel()  # col-major
    bb = _f64(b).copy()
    s = lib().orc_solve3_sym(_ptr(Acm), _ptr(bb))
    return bb, bool(s)


def wrapped_err(a, b):
    a = _f64(a).ravel()
    b = _f64(b).ravel()
    r, m = ctypes.c_double(), ctypes.c_double()
    lib().orc_wrapped_err(_ptr(a), _ptr(b), a.size, ctypes.byref(r), ctypes.byref(m))
    return r.value, m.value


# ---------------------------------------------------------------- AIA
def _frames2d(f):
    f = _f64(f)
    return np.ascontiguousarray(f.reshape(f.shape[0], -1)), f.shape[1:]


def aia_phi(f, delta, nthreads=0, return_x=False):
    """computePhi (AIA/aia_cpuf.cpp:112-183): phase of every pixel for the
    given shifts; with return_x also the per-pixel solutions (N x 3)."""
    F, shape = _frames2d(f)
    d = _f64(delta)
    phi = np.empty(F.shape[1])
    X = np.empty((F.shape[1], 3)) if return_x else None
    info = lib().orc_aia_phi(_ptr(F), F.shape[0], F.shape[1], _ptr(d), _ptr(phi),
                             _ptr(X) if return_x else None, nthreads or os.cpu_count())
    if info:
        raise ArithmeticError(f"leading minor of order {info} is not positive definite")
    return (phi.reshape(shape), X) if return_x else phi.reshape(shape)


def aia_delta(f, phi, nthreads=0):
    """computeDelta (AIA/aia_cpuf.cpp:185-245): shifts of every frame for the
    given phase."""
    F, _ = _frames2d(f)
    p = _f64(phi).ravel()
    d = np.empty(F.shape[0])
    info = lib().orc_aia_delta(_ptr(F), F.shape[0], F.shape[1], _ptr(p), _ptr(d),
                               nthreads or os.cpu_count())
    if info:
        raise ArithmeticError(f"leading minor of order {info} is not positive definite")
    return d


def aia(f, delta, max_iter=20, max_err=1e-4, err_mode=0, nthreads=0):
    """AIA operator() (AIA/aia_cpuf.cpp:16-107, AIA/aia_cudaf.cu:290-372):
    returns (phi, deltas, iters, err) with phi + delta_0 and delta - delta_0
    wrapped."""
    F, shape = _frames2d(f)
    d = _f64(delta).copy()
    phi = np.empty(F.shape[1])
    it, err = ctypes.c_int(), ctypes.c_double()
    info = lib().orc_aia(_ptr(F), F.shape[0], F.shape[1], _ptr(d), int(max_iter),
                         float(max_err), int(err_mode), _ptr(phi), ctypes.byref(it),
                         ctypes.byref(err), nthreads or os.cpu_count())
    if info:
        raise ArithmeticError(f"leading minor of order {info} is not positive definite")
    return phi.reshape(shape), d, it.value, err.value
