"""Many DPRA sequences of uneven sizes in one batched, graph-replayed run.

Each sequence behaves as its own DPRA_CPU / DPRA_CUDAF object
(DPRA/dpra_cpu.h:30-48, DPRA/dpra_cudaf.h:28-45): constructed from its Phi0
and rr, then driven by operator()(frames, dPhi_Sum, time).  The batch runs all
of them together on the device (one least-squares launch over all images and
one launch per WFF2 stage per FFT-plan group, per frame) and matches
glsu.DPRA on each sequence bit for bit when both use the same tuning.
"""
import ctypes

import numpy as np
import torch

from ._lib import SeqDesc, Tuning, check, check_out, lib, tuning_arg, tuning_dict
from .params import wff_params

_DT = {torch.uint8: 0, torch.float32: 1, torch.float64: 2}


class DPRABatch:
    def __init__(self, phi0s, rrs=1, params=None, stream=None, tunings=None):
        """phi0s: list of host arrays (H_i, W_i).  rrs: int or list.  tunings:
        None, or one tuning (dict / glsu.Tuning) per sequence, e.g. from
        glsu.autotune.tune for that sequence's size."""
        self._phi = [np.ascontiguousarray(np.asarray(p, dtype=np.float64)) for p in phi0s]
        n = len(self._phi)
        if n < 1:
            raise ValueError("empty batch")
        rrs = [int(rrs)] * n if np.isscalar(rrs) else [int(r) for r in rrs]
        self.shapes = [p.shape for p in self._phi]
        self.params = params if params is not None else wff_params()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        descs = (SeqDesc * n)()
        for i, p in enumerate(self._phi):
            descs[i] = SeqDesc(p.shape[1], p.shape[0], rrs[i],
                               p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        tarr = None
        if tunings is not None:
            tarr = (Tuning * n)()
            for i, t in enumerate(tunings):
                tarr[i] = tuning_arg(t)
        h = ctypes.c_void_p()
        check(lib().glsu_dpra_batch_create(ctypes.byref(h), descs, n, ctypes.byref(self.params),
                                           ctypes.c_void_p(self.stream.cuda_stream), tarr))
        self._h = h
        self.n = n
        self.last_ms = None

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            try:
                lib().glsu_dpra_batch_destroy(h)
            except Exception:  # interpreter shutdown
                pass

    __del__ = close

    def info(self):
        c, t = ctypes.c_int(), ctypes.c_int()
        check(lib().glsu_dpra_batch_info(self._h, ctypes.byref(c), ctypes.byref(t)))
        return dict(classes=c.value, tiles=t.value)

    def tuning(self, i):
        t = Tuning.make()
        check(lib().glsu_dpra_batch_get_tuning(self._h, int(i), ctypes.byref(t)))
        return tuning_dict(t)

    def __call__(self, frames, outs=None, graph=True, timed=False):
        """frames: list of (T, H_i, W_i) CUDA tensors (same T and dtype).
        Returns the list of dPhi_Sum tensors (T, H_i, W_i) float32."""
        if len(frames) != self.n:
            raise ValueError(f"expected {self.n} frame tensors")
        T, dt = frames[0].shape[0], frames[0].dtype
        for f, (H, W) in zip(frames, self.shapes):
            if (not f.is_cuda or not f.is_contiguous() or f.dtype != dt or f.dtype not in _DT
                    or tuple(f.shape) != (T, H, W)):
                raise ValueError("frames must be contiguous CUDA tensors (T, H_i, W_i) of one dtype")
        if outs is None:
            outs = [torch.empty((T, H, W), dtype=torch.float32, device=frames[0].device)
                    for (H, W) in self.shapes]
        if len(outs) != self.n:
            raise ValueError(f"expected {self.n} output tensors")
        for i, (o, (H, W)) in enumerate(zip(outs, self.shapes)):
            check_out(o, (T, H, W), torch.float32, f"outs[{i}]")
        fp = (ctypes.c_void_p * self.n)(*[f.data_ptr() for f in frames])
        op = (ctypes.c_void_p * self.n)(*[o.data_ptr() for o in outs])
        ms = ctypes.c_double()
        check(lib().glsu_dpra_batch_run(self._h, fp, int(T), _DT[dt], op, int(bool(graph)),
                                        ctypes.byref(ms) if timed else None))
        self.last_ms = ms.value if timed else None
        return outs

    def get_state(self, i):
        H, W = self.shapes[i]
        phr = np.empty((H, W), np.float32)
        dpr = np.empty((H, W), np.float32)
        check(lib().glsu_dpra_batch_get_state(self._h, int(i), phr.ctypes.data, dpr.ctypes.data))
        return phr, dpr
