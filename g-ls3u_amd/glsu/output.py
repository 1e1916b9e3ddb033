"""Output stage: phase maps to 8-bit video frames on the GPU, as the
reference demo writes them (App_DPRA/dprawidget.cpp:541-549)."""
import ctypes

import torch

from ._lib import check, lib


def phase_to_u8(phi, out=None, stream=None):
    """uchar((atan2(sin p, cos p) + pi) / 2 / pi * 255) of a float32 CUDA tensor
    of any shape (e.g. (frames, H, W) of cumulative dphi); returns uint8."""
    if phi.dtype != torch.float32 or not phi.is_cuda or not phi.is_contiguous():
        raise ValueError("phi must be a contiguous float32 CUDA tensor")
    if out is None:
        out = torch.empty(phi.shape, dtype=torch.uint8, device=phi.device)
    if out.dtype != torch.uint8 or out.shape != phi.shape or not out.is_contiguous():
        raise ValueError("out must be a contiguous uint8 tensor of phi's shape")
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().glsu_phase_to_u8(ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                 phi.numel(), ctypes.c_void_p(s.cuda_stream)))
    return out
