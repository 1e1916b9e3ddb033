"""Output stage (glsu_phase_to_u8) against the reference demo's formula
(App_DPRA/dprawidget.cpp:547) evaluated with numpy in float32: identical bytes
except where the float32 value lies within rounding of an integer boundary
(transcendental rounding differs between libraries by an ulp)."""
import numpy as np
import pytest

from _util import say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402


def reference_bytes(p):
    p = p.astype(np.float32)
    w = np.arctan2(np.sin(p), np.cos(p)).astype(np.float32)
    v = (w + np.float32(np.pi)) / np.float32(2.0) / np.float32(np.pi) * np.float32(255.0)
    return v.astype(np.uint8), v


@pytest.mark.parametrize("shape", [(3, 257, 131), (1, 1, 5), (4096,), (2, 64, 64)])
def test_phase_to_u8_matches_formula(shape):
    rng = np.random.default_rng(sum(shape))
    p = rng.uniform(-40, 40, shape).astype(np.float32)
    p.flat[:4] = [0.0, np.float32(np.pi), -np.float32(np.pi), 1e-8][: min(4, p.size)]
    ref, v = reference_bytes(p)
    t = torch.from_numpy(p).cuda()
    say(f"[phase_to_u8] shape {shape}")
    out = glsu.phase_to_u8(t).cpu().numpy()
    diff = out.astype(int) - ref.astype(int)
    bad = diff != 0
    say(f"   mismatches {bad.sum()} of {p.size}")
    assert np.abs(diff).max() <= 1
    # a mismatch only where the exact value sits on an integer boundary (FP64)
    p64 = p.astype(np.float64)
    v64 = (np.arctan2(np.sin(p64), np.cos(p64)) + np.pi) / 2 / np.pi * 255
    if bad.any():
        assert np.abs(v64[bad] - np.rint(v64[bad])).max() < 1e-4
    assert bad.mean() < 1e-3
    # unaligned pointers take the one-pixel path: same bytes
    if p.size > 8:
        import ctypes
        flat = torch.from_numpy(p.reshape(-1)).cuda()
        o2 = torch.zeros(p.size + 1, dtype=torch.uint8, device="cuda")
        glsu.check(glsu.lib().glsu_phase_to_u8(ctypes.c_void_p(flat.data_ptr() + 4),
                                               ctypes.c_void_p(o2.data_ptr() + 1), p.size - 1,
                                               None))
        torch.cuda.synchronize()
        assert np.array_equal(o2.cpu().numpy()[1:p.size], out.reshape(-1)[1:])
