"""AIA: the initial phase Phi0 and the phase shifts of M >= 3 randomly
phase-shifted frames, on libglsu's HIP kernels.

Mirrors AIA::AIA_CUDAF (AIA/aia_cudaf.h:15-72): construction with the frame
count and image size (AIA_CUDAF(iM, icols, irows), :23-25) and a call that
returns (phi, deltas, runningtime, iters, err) as operator() fills its output
arguments (:28-38, AIA/aia_cudaf.cu:290-372).  Frames go in as a
(M, height, width) uint8 / float32 / float64 tensor on the GPU (device path)
or a numpy array / list of 2-D arrays (host path, the reference's
std::vector<cv::Mat> form).
"""
import ctypes

import numpy as np
import torch

from ._lib import F32, F64, U8, check, check_out, lib

ERR_ABS, ERR_REL = 0, 1
_DT = {torch.uint8: U8, torch.float32: F32, torch.float64: F64}
_NDT = {np.dtype(np.uint8): U8, np.dtype(np.float32): F32, np.dtype(np.float64): F64}


class AIA:
    """AIA(nframes, width, height).  Call with the frames; optional initial
    shifts `deltas` (length nframes; None draws delta_0 = 0 and N(0,1) shifts
    from `seed`, as the reference does from a clock seed, AIA/aia_cpuf.cpp:31-45),
    `max_iter` / `max_err` (reference defaults 20 / 1e-4) and `err_mode`:
    ERR_ABS = max |delta - delta_old| (AIA_CUDAF, AIA_CPU_Dn, fp_aia.m),
    ERR_REL = the same on shifts relative to delta_0 (AIA_CPU_DnF, AIA_CUDAF_YC2).
    Returns (phi, deltas, ms, iters, err): phi + delta_0 wrapped (float32, on
    the frames' device), deltas - delta_0 wrapped (numpy float32)."""

    def __init__(self, nframes, width, height, stream=None):
        self.M, self.W, self.H = int(nframes), int(width), int(height)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        check(lib().glsu_aia_create(ctypes.byref(h), self.M, self.W, self.H,
                                    ctypes.c_void_p(self.stream.cuda_stream)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            try:
                lib().glsu_aia_destroy(h)
            except Exception:  # interpreter shutdown
                pass

    __del__ = close

    def info(self):
        b, c, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().glsu_aia_info(self._h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(k)))
        return dict(blocks=b.value, chunks=c.value, frames_per_chunk=k.value)

    def time_iter(self, frames, deltas, reps=20, stage=0):
        """Average ms over `reps` repetitions with the convergence test off of
        one iteration (stage 0: pass + solve launches) or of the pass kernel
        alone (stage 1)."""
        if (frames.dtype not in _DT or tuple(frames.shape) != (self.M, self.H, self.W)
                or not frames.is_cuda or not frames.is_contiguous()):
            raise ValueError("frames must be a contiguous (M, H, W) uint8/float32/float64 CUDA tensor")
        d = self._deltas(deltas)
        ms = ctypes.c_double()
        check(lib().glsu_aia_time_iter(self._h, ctypes.c_void_p(frames.data_ptr()),
                                       _DT[frames.dtype], d.ctypes.data, int(stage), int(reps),
                                       ctypes.byref(ms)))
        return ms.value

    def _deltas(self, deltas):
        d = np.zeros(self.M, np.float32)
        if deltas is not None:
            deltas = np.asarray(deltas, np.float32).ravel()
            if deltas.size != self.M:
                raise ValueError(f"expected {self.M} initial shifts, got {deltas.size}")
            d[:] = deltas
        return d

    def __call__(self, frames, deltas=None, max_iter=20, max_err=1e-4, err_mode=ERR_ABS, seed=0,
                 phi=None):
        if isinstance(frames, torch.Tensor) and frames.is_cuda:
            return self._run_device(frames, deltas, max_iter, max_err, err_mode, seed, phi)
        return self._run_host(frames, deltas, max_iter, max_err, err_mode, seed)

    def _run_device(self, frames, deltas, max_iter, max_err, err_mode, seed, phi):
        if frames.dtype not in _DT:
            raise TypeError(f"frames must be uint8, float32 or float64, not {frames.dtype}")
        if tuple(frames.shape) != (self.M, self.H, self.W) or not frames.is_contiguous():
            raise ValueError(f"frames must be a contiguous ({self.M}, {self.H}, {self.W}) tensor")
        if phi is None:
            phi = torch.empty((self.H, self.W), dtype=torch.float32, device=frames.device)
        check_out(phi, (self.H, self.W), torch.float32, "phi")
        d = self._deltas(deltas)
        it, err, ms = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().glsu_aia_run(self._h, ctypes.c_void_p(frames.data_ptr()), _DT[frames.dtype],
                                 d.ctypes.data, int(deltas is not None), int(seed), int(max_iter),
                                 float(max_err), int(err_mode), ctypes.c_void_p(phi.data_ptr()),
                                 ctypes.byref(it), ctypes.byref(err), ctypes.byref(ms)))
        return phi, d, ms.value, it.value, err.value

    def _run_host(self, frames, deltas, max_iter, max_err, err_mode, seed):
        fr = [np.ascontiguousarray(f) for f in frames]
        if len(fr) != self.M:
            raise ValueError(f"expected {self.M} frames, got {len(fr)}")
        dt = fr[0].dtype
        if dt not in _NDT or any(f.dtype != dt or f.shape != (self.H, self.W) for f in fr):
            raise ValueError(f"frames must be {self.H}x{self.W} arrays of one of uint8, float32, "
                             "float64")
        ptrs = (ctypes.c_void_p * self.M)(*[f.ctypes.data for f in fr])
        phi = np.empty((self.H, self.W), np.float32)
        d = self._deltas(deltas)
        it, err, ms = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().glsu_aia_run_host(self._h, ptrs, _NDT[dt], d.ctypes.data,
                                      int(deltas is not None), int(seed), int(max_iter),
                                      float(max_err), int(err_mode), phi.ctypes.data,
                                      ctypes.byref(it), ctypes.byref(err), ctypes.byref(ms)))
        return phi, d, ms.value, it.value, err.value
