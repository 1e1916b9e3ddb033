"""WFT2 on the GPU -- mirror of WFT_FPA::WFT::WFT2_CUDAF, WFF and WFR modes.

Reference interface: WFT/WFT2_CUDAf.h:21-40 (ctor(W, H, type, sigma_x, wxl,
wxh, wxi, sigma_y, wyl, wyh, wyi, thr, z, nStreams); operator()(d_f, d_z, time)).
Results live in device memory, as WFT2_DeviceResultsF's members do
(m_d_filtered in WFF mode; m_d_wx ... m_d_cyy in WFR mode, WFT/WFT.h:114-135).
"""
import ctypes

import torch

from . import _lib
from ._lib import WFF, WFR, Tuning, WfrResults, check, check_out, lib, tuning_arg, tuning_dict
from .params import wff_params


class WFT2:
    def __init__(self, width, height, type=WFF, params=None, stream=None, planx=-1, plany=-1, tuning=None):
        """params: WffParams (see glsu.wff_params); None -> wff_params() defaults
        (sigma 10, [-0.2, 0.2] step 0.1, thr < 0).  tuning: glsu.Tuning or a
        dict of its fields (plan ids, launch shapes), e.g. from glsu.autotune."""
        self.width, self.height = int(width), int(height)
        self.type = int(type)
        self.params = params if params is not None else wff_params()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        t = tuning_arg(tuning, planx, plany)
        check(lib().glsu_wft2_create_tuned(ctypes.byref(h), self.width, self.height, int(type),
                                           ctypes.byref(self.params),
                                           ctypes.c_void_p(self.stream.cuda_stream),
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
        lib().glsu_wft2_destroy(h)

    __del__ = close

    def info(self):
        v = [ctypes.c_int() for _ in range(6)]
        thr = ctypes.c_double()
        check(lib().glsu_wft2_info(self._h, *[ctypes.byref(x) for x in v], ctypes.byref(thr)))
        nx, ny, sx, sy, nwx, nwy = (x.value for x in v)
        return dict(nx=nx, ny=ny, sx=sx, sy=sy, nwx=nwx, nwy=nwy, thr=thr.value)

    def tuning(self):
        """The plan ids and launch shapes in use."""
        t = Tuning.make()
        check(lib().glsu_wft2_get_tuning(self._h, ctypes.byref(t)))
        return tuning_dict(t)

    def time_kernel(self, f, stage, reps=10):
        """Average ms of one kernel stage (1 rows, 2 columns, 3 the rest) over
        reps launches on this handle's stream, after one transform of f."""
        self._check_f(f)
        ms = ctypes.c_double()
        check(lib().glsu_wft2_time_kernel(self._h, ctypes.c_void_p(f.data_ptr()), int(stage),
                                          int(reps), ctypes.byref(ms)))
        return ms.value

    WFR_FIELDS = ("wx", "wy", "phase", "phase_comp", "b", "r", "cxx", "cyy")

    def wfr(self, f, out=None, timed=False):
        """WFR mode: f complex64 CUDA tensor (height, width).  Returns a dict of
        the eight float32 maps (wx, wy, phase, phase_comp, b, r, cxx, cyy); out
        may supply a (8, height, width) float32 tensor to fill."""
        self._check_f(f)
        if out is None:
            out = torch.empty((8, self.height, self.width), dtype=torch.float32, device=f.device)
        check_out(out, (8, self.height, self.width), torch.float32)
        z = WfrResults(*[out[i].data_ptr() for i in range(8)])
        ms = ctypes.c_double()
        check(lib().glsu_wft2_compute_wfr(self._h, ctypes.c_void_p(f.data_ptr()), ctypes.byref(z),
                                          ctypes.byref(ms) if timed else None))
        self.last_ms = ms.value if timed else None
        return dict(zip(self.WFR_FIELDS, out))

    def _check_f(self, f):
        if f.dtype != torch.complex64 or not f.is_cuda or not f.is_contiguous():
            raise ValueError("f must be a contiguous complex64 CUDA tensor")
        if tuple(f.shape) != (self.height, self.width):
            raise ValueError(f"f has shape {tuple(f.shape)}, expected {(self.height, self.width)}")

    def __call__(self, f, out=None, timed=False):
        """f: complex64 CUDA tensor (height, width).  WFF mode: returns the
        filtered field; WFR mode: the dict of wfr()."""
        if self.type == WFR:
            return self.wfr(f, out, timed)
        self._check_f(f)
        if out is None:
            out = torch.empty_like(f)
        check_out(out, (self.height, self.width), torch.complex64)
        ms = ctypes.c_double()
        check(lib().glsu_wft2_compute(self._h, ctypes.c_void_p(f.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()),
                                      ctypes.byref(ms) if timed else None))
        self.last_ms = ms.value if timed else None
        return out
