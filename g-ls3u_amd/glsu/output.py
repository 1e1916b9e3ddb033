"""Output stage: phase maps to 8-bit video frames on the GPU, as the
reference demo writes them (App_DPRA/dprawidget.cpp:541-549)."""
import ctypes

import torch

from ._lib import check, check_out, lib


def phase_to_u8(phi, out=None, stream=None):
    """uchar((atan2(sin p, cos p) + pi) / 2 / pi * 255) of a float32 CUDA tensor
    of any shape (e.g. (frames, H, W) of cumulative dphi); returns uint8."""
    if phi.dtype != torch.float32 or not phi.is_cuda or not phi.is_contiguous():
        raise ValueError("phi must be a contiguous float32 CUDA tensor")
    if out is None:
        out = torch.empty(phi.shape, dtype=torch.uint8, device=phi.device)
    check_out(out, phi.shape, torch.uint8)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().glsu_phase_to_u8(ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                 phi.numel(), ctypes.c_void_p(s.cuda_stream)))
    return out
