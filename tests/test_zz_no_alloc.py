"""glsu_wft2_compute_noalloc (the WFF2 entry point that allocates nothing)
against glsu_wft2_compute on a fresh handle with the same parameters, through
the C ABI: two no-allocation calls in a row, the second with a smaller field
(what the first left in the scratch array must not leak into it), then
fields the entry point must refuse with a clear error before any launch.
Bar: relative field error and phase error <= 1e-6 -- both paths run the same
kernels; only the window spectra and twiddles are computed on the device (FP64)
instead of the host, which can move a table entry by one float ulp at most."""
import ctypes

import numpy as np
import pytest

from _util import say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402


def _field(W, H, seed):
    rng = np.random.default_rng(seed)
    ph = 0.45 * np.arange(W)[None, :] - 0.2 * np.arange(H)[:, None] + rng.normal(0, 0.5, (H, W))
    amp = 1.0 + 0.3 * rng.normal(size=(H, W))
    return torch.from_numpy((amp * np.exp(1j * ph)).astype(np.complex64)).cuda()


def _noalloc(f, params, out):
    H, W = f.shape
    s = torch.cuda.current_stream()
    rc = glsu.lib().glsu_wft2_compute_noalloc(W, H, ctypes.byref(params),
                                              ctypes.c_void_p(f.data_ptr()),
                                              ctypes.c_void_p(out.data_ptr()),
                                              ctypes.c_void_p(s.cuda_stream))
    return rc


@pytest.mark.parametrize("thr", [-1.0, 2.5])
def test_noalloc_matches_handle_path_two_calls(thr):
    params = glsu.wff_params(10.0, -0.2, 0.2, 0.1, thr=thr)
    cases = [(1024, 1024, 1), (333, 257, 2)]  # the second call is smaller
    fields = [_field(W, H, seed) for (W, H, seed) in cases]
    refs = [glsu.WFT2(W, H, params=params)(f) for (W, H, _), f in zip(cases, fields)]
    outs = []
    for (W, H, _), f in zip(cases, fields):
        say(f"[noalloc] {W}x{H} thr={thr}")
        out = torch.full_like(f, complex(float("nan"), float("nan")))
        assert _noalloc(f, params, out) == glsu._lib.OK, glsu.lib().glsu_last_error()
        outs.append(out)
    torch.cuda.synchronize()
    for (W, H, _), out, ref in zip(cases, outs, refs):
        a, b = out.cpu().numpy(), ref.cpu().numpy()
        rel = float(np.abs(a - b).max() / np.abs(b).max())
        e = np.angle(a * np.conj(b))
        rms = float(np.sqrt(np.mean(e * e)))
        say(f"   {W}x{H}: bitwise equal {np.array_equal(a, b)}, rel {rel:.2e}, "
            f"phase rms {rms:.2e}")
        assert np.isfinite(a).all()
        assert rel <= 1e-6 and rms <= 1e-6


def test_noalloc_rejects_fields_beyond_the_scratch():
    params = glsu.wff_params(10.0, -0.2, 0.2, 0.1)
    f = torch.zeros((1, 1), dtype=torch.complex64, device="cuda")  # never read
    rc = glsu.lib().glsu_wft2_compute_noalloc(2048, 2048, ctypes.byref(params),
                                              ctypes.c_void_p(f.data_ptr()),
                                              ctypes.c_void_p(f.data_ptr()), None)
    msg = glsu.lib().glsu_last_error().decode()
    say(f"[noalloc] 2048x2048 x 5 wx: rc {rc}, {msg}")
    assert rc == glsu._lib.ERR_UNSUPPORTED and "scratch" in msg
    rc = glsu.lib().glsu_wft2_compute_noalloc(4050, 64, ctypes.byref(params),
                                              ctypes.c_void_p(f.data_ptr()),
                                              ctypes.c_void_p(f.data_ptr()), None)
    assert rc == glsu._lib.ERR_PADDING
