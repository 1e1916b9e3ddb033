"""DPRA on the GPU -- mirror of DPRA::DPRA_CUDAF / DPRA::DPRA_CPU.

Reference interface: DPRA/dpra_cudaf.h:21-45 and DPRA/dpra_cpu.h:15-64:
  ctor(Phi0, W, H, rr), dpra_per_frame(frame, dPhi, time), update_ref_phi(),
  operator()(frames, dPhi_Sum, time) with `if (i % rr == 0) update_ref_phi()`.
Frames and outputs are CUDA tensors; the work runs in libglsu's HIP kernels.
"""
import ctypes

import numpy as np
import torch

from ._lib import F32, F64, U8, Tuning, check, check_out, lib, tuning_arg, tuning_dict
from .params import wff_params

_DT = {torch.uint8: U8, torch.float32: F32, torch.float64: F64}


class DPRA:
    def __init__(self, phi0, rr=1, params=None, stream=None, planx=-1, plany=-1, tuning=None):
        """phi0: host array (H, W) of the initial phase (float64 recommended).
        params: WffParams; None -> sigma 10, [-0.2, 0.2] step 0.1, thr < 0
        (the configs' settings; glsu.dpra_reference_params() gives the values
        the reference DPRA hard-codes).  tuning: glsu.Tuning or a dict of its
        fields (plan ids, launch shapes), e.g. from glsu.autotune."""
        phi = np.ascontiguousarray(np.asarray(phi0, dtype=np.float64))
        self.height, self.width = phi.shape
        self.rr = int(rr)
        self.params = params if params is not None else wff_params()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        t = tuning_arg(tuning, planx, plany)
        check(lib().glsu_dpra_create_tuned(
            ctypes.byref(h), phi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.width,
            self.height, self.rr, ctypes.byref(self.params), ctypes.c_void_p(self.stream.cuda_stream),
            ctypes.byref(t)))
        self._h = h
        self.last_ms = None

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            try:
                self._destroy(h)
            except Exception:  # interpreter shutdown
                pass

    @staticmethod
    def _destroy(h):
        lib().glsu_dpra_destroy(h)

    __del__ = close

    def tuning(self):
        """The plan ids and launch shapes in use."""
        t = Tuning.make()
        check(lib().glsu_dpra_get_tuning(self._h, ctypes.byref(t)))
        return tuning_dict(t)

    def _check_frame(self, frame, nd=2):
        if not frame.is_cuda or not frame.is_contiguous() or frame.dtype not in _DT:
            raise ValueError("frame must be a contiguous uint8/float32/float64 CUDA tensor")
        if tuple(frame.shape[-2:]) != (self.height, self.width) or frame.dim() != nd:
            raise ValueError(f"frame shape {tuple(frame.shape)} does not match "
                             f"{(self.height, self.width)}")
        return _DT[frame.dtype]

    def dpra_per_frame(self, frame, dphi=None, phicurr=None, timed=False):
        """One frame: LS + WFF2 + phase update.  Returns dphi (float32 CUDA)."""
        dt = self._check_frame(frame)
        hw = (self.height, self.width)
        if dphi is None:
            dphi = torch.empty(hw, dtype=torch.float32, device=frame.device)
        check_out(dphi, hw, torch.float32, "dphi")
        if phicurr is not None:
            check_out(phicurr, hw, torch.float32, "phicurr")
        ms = ctypes.c_double()
        check(lib().glsu_dpra_frame(self._h, ctypes.c_void_p(frame.data_ptr()), dt,
                                    ctypes.c_void_p(dphi.data_ptr()),
                                    ctypes.c_void_p(phicurr.data_ptr()) if phicurr is not None else None,
                                    ctypes.byref(ms) if timed else None))
        self.last_ms = ms.value if timed else None
        return dphi

    def update_ref_phi(self):
        check(lib().glsu_dpra_update_ref(self._h))

    def __call__(self, frames, out=None, graph=False, timed=False):
        """frames: (T, H, W) CUDA tensor.  Returns dPhi_Sum as (T, H, W) float32."""
        dt = self._check_frame(frames, nd=3)
        T = frames.shape[0]
        shape = (T, self.height, self.width)
        if out is None:
            out = torch.empty(shape, dtype=torch.float32, device=frames.device)
        check_out(out, shape, torch.float32)
        ms = ctypes.c_double()
        fn = lib().glsu_dpra_run_graph if graph else lib().glsu_dpra_run
        check(fn(self._h, ctypes.c_void_p(frames.data_ptr()), int(T), dt,
                 ctypes.c_void_p(out.data_ptr()), ctypes.byref(ms) if timed else None))
        self.last_ms = ms.value if timed else None
        return out

    def profile(self, frames, out=None):
        """Run the frame loop with events between kernels; returns the summed
        device milliseconds of (least squares, rows fwd, columns, rows inv)."""
        dt = self._check_frame(frames, nd=3)
        T = frames.shape[0]
        shape = (T, self.height, self.width)
        if out is None:
            out = torch.empty(shape, dtype=torch.float32, device=frames.device)
        check_out(out, shape, torch.float32)
        ms = (ctypes.c_double * 4)()
        check(lib().glsu_dpra_profile(self._h, ctypes.c_void_p(frames.data_ptr()), int(T), dt,
                                      ctypes.c_void_p(out.data_ptr()), ms))
        return list(ms)

    def profile_skips(self, frames, out=None):
        """profile() that also returns the column transforms skipped because a
        column's thresholded signal was zero: (ms[4], skipped, transforms)."""
        dt = self._check_frame(frames, nd=3)
        T = frames.shape[0]
        shape = (T, self.height, self.width)
        if out is None:
            out = torch.empty(shape, dtype=torch.float32, device=frames.device)
        check_out(out, shape, torch.float32)
        ms = (ctypes.c_double * 4)()
        sk, tot = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(lib().glsu_dpra_profile_skips(self._h, ctypes.c_void_p(frames.data_ptr()), int(T),
                                            dt, ctypes.c_void_p(out.data_ptr()), ms,
                                            ctypes.byref(sk), ctypes.byref(tot)))
        return list(ms), sk.value, tot.value

    def time_kernel(self, stage, reps=20):
        """Average ms per launch of one WFF2 kernel (1 rows fwd, 2 columns,
        3 rows inv + update), timed with HIP events on the handle's stream."""
        ms = ctypes.c_double()
        check(lib().glsu_dpra_time_kernel(self._h, int(stage), int(reps), ctypes.byref(ms)))
        return ms.value

    def frame_learned(self, frame, net, dphi=None, tol=0.03, tile=(0, 0), sync=True):
        """dpra_per_frame with the U-Net `net` (glsu.learned.UNet) in place of
        WFF2 and a WFF2 spot check on the 64x64 tile at `tile` (x, y): if the
        network's phase there differs from WFF2's by more than `tol` rad RMS,
        the frame is filtered by WFF2 instead.  Returns (residual, fallback);
        with sync=False the call does not wait for the decision (it is taken
        on the device) and returns None (see learned_stats)."""
        dt = self._check_frame(frame)
        hw = (self.height, self.width)
        if dphi is None:
            dphi = torch.empty(hw, dtype=torch.float32, device=frame.device)
        check_out(dphi, hw, torch.float32, "dphi")
        res = ctypes.c_double()
        fb = ctypes.c_int()
        check(lib().glsu_dpra_frame_learned(
            self._h, net._h, ctypes.c_void_p(frame.data_ptr()), dt, ctypes.c_void_p(dphi.data_ptr()),
            float(tol), int(tile[0]), int(tile[1]), ctypes.byref(res) if sync else None,
            ctypes.byref(fb) if sync else None))
        self.last_dphi = dphi
        return (res.value, bool(fb.value)) if sync else None

    def learned_stats(self, reset=False):
        """Spot-check statistics of the learned frames since the last reset."""
        f, b = ctypes.c_int(), ctypes.c_int()
        m, x = ctypes.c_double(), ctypes.c_double()
        check(lib().glsu_dpra_learned_stats(self._h, ctypes.byref(f), ctypes.byref(b),
                                            ctypes.byref(m), ctypes.byref(x), int(bool(reset))))
        return dict(frames=f.value, fallbacks=b.value, mean_rms=m.value, max_rms=x.value)

    def last_net(self):
        z = torch.empty((self.height, self.width), dtype=torch.complex64, device="cuda")
        check(lib().glsu_dpra_copy_net(self._h, ctypes.c_void_p(z.data_ptr())))
        return z

    def get_state(self, dphiref=True):
        """(phiref, dphiref) as host float32 arrays; dphiref=False skips the
        second (returns None in its place)."""
        n = self.height * self.width
        phr = np.empty(n, np.float32)
        dpr = np.empty(n, np.float32) if dphiref else None
        check(lib().glsu_dpra_get_state(self._h, phr.ctypes.data,
                                        dpr.ctypes.data if dphiref else None))
        return (phr.reshape(self.height, self.width),
                dpr.reshape(self.height, self.width) if dphiref else None)

    def set_state(self, phiref=None, dphiref=None):
        a = None if phiref is None else np.ascontiguousarray(phiref, np.float32)
        b = None if dphiref is None else np.ascontiguousarray(dphiref, np.float32)
        check(lib().glsu_dpra_set_state(self._h, a.ctypes.data if a is not None else None,
                                        b.ctypes.data if b is not None else None))

    def last_z(self):
        z = torch.empty((self.height, self.width), dtype=torch.complex64, device="cuda")
        check(lib().glsu_dpra_copy_z(self._h, ctypes.c_void_p(z.data_ptr())))
        return z

    @property
    def thr(self):
        t = ctypes.c_double()
        check(lib().glsu_dpra_thr(self._h, ctypes.byref(t)))
        return t.value


def gen_frames(width, height, nframes, seed, amp0=3.0, amp1=0.1, carrier=0.05, noise=10.0, t0=0,
               out=None, stream=None):
    """Synthetic float32 fringe frames generated on the GPU (same definition as
    the oracle's generator: SplitMix64 + FP64 Box-Muller)."""
    if out is None:
        out = torch.empty((nframes, height, width), dtype=torch.float32, device="cuda")
    check_out(out, (nframes, height, width), torch.float32)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().glsu_gen_frames(ctypes.c_void_p(out.data_ptr()), int(width), int(height), int(t0),
                                int(nframes), int(seed), float(amp0), float(amp1), float(carrier),
                                float(noise), ctypes.c_void_p(s.cuda_stream)))
    return out
