"""glsu -- MI355X (gfx950) G-LS3U hot path: DPRA + WFF2 + reference update.

Host-side mirror of the reference's class interfaces (DPRA_CUDAF / DPRA_CPU,
WFT2_CUDAF / WFT2_cpu and AIA_CUDAF) over the C ABI of libglsu.so.  Device memory is handed
over as torch tensors (plumbing only); all computation runs in the library's
HIP kernels.
"""
from ._lib import (GlsuError, PaddingError, NotPositiveDefiniteError, WffParams, Tuning, build, check, lib, fft_plans,  # noqa: F401
                   WFF, WFR, U8, F32, F64, LIB_PATH)
from .params import wff_params, dpra_reference_params, wft2_default_params  # noqa: F401
from .wft2 import WFT2  # noqa: F401
from .dpra import DPRA, gen_frames  # noqa: F401
from .batch import DPRABatch  # noqa: F401
from .aia import AIA  # noqa: F401
from .output import phase_to_u8  # noqa: F401
from . import synth, autotune, learned, io  # noqa: F401
from .learned import UNet  # noqa: F401
