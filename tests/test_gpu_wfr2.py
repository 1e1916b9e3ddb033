"""WFR2 (the ridge mode of WFT2) on the GPU against the FP64 oracle's
orc_wfr2_compute (itself pinned by tests/test_oracle_wfr2.py).  The ridge is
an argmax over the (wx, wy) grid: where two pairs' |sf| agree to FP32
rounding the two paths may pick different pairs, so the bars are: r within
1e-5 relative everywhere (it is continuous across such switches), the pair
identical on all but a few pixels, phase within 1e-4 rad RMS where the pair is
the same, and curvatures/amplitude within tolerances stated per test."""
import ctypes

import numpy as np
import pytest

from _util import say, wrap

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402


def field(W, H, seed, noise=0.2, amp=3.0, carrier=0.05):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(-3, 3, W), np.linspace(-3, 3, H))
    peaks = (3 * (1 - X) ** 2 * np.exp(-X ** 2 - (Y + 1) ** 2)
             - 10 * (X / 5 - X ** 3 - Y ** 5) * np.exp(-X ** 2 - Y ** 2)
             - np.exp(-(X + 1) ** 2 - Y ** 2) / 3)
    ph = amp * peaks + 2 * np.pi * carrier * np.arange(W)[None, :]
    f = (1 + 0.3 * np.cos(X)) * np.exp(1j * ph)
    f = f + noise * (rng.normal(size=(H, W)) + 1j * rng.normal(size=(H, W)))
    return f.astype(np.complex64)


def run_both(oracle, f, sig, lo, hi, step):
    H, W = f.shape
    ref = oracle.WFR2(W, H, sigx=sig, wxl=lo, wxh=hi, wxi=step)(f.astype(np.complex128))
    w = glsu.WFT2(W, H, type=glsu.WFR, params=glsu.wff_params(sig, lo, hi, step))
    z = w(torch.from_numpy(f).cuda(), timed=True)
    return ref, {k: v.cpu().numpy().astype(np.float64) for k, v in z.items()}, w.last_ms


CASES = [  # W, H, sigma, lo, hi, step
    (64, 48, 3.0, -0.3, 0.3, 0.1),
    (201, 157, 6.0, -0.5, 0.5, 0.1),
    (256, 256, 10.0, -0.5, 0.5, 0.05),
]


@pytest.mark.parametrize("W,H,sig,lo,hi,step", CASES)
def test_wfr2_matches_oracle(oracle, W, H, sig, lo, hi, step):
    f = field(W, H, W + H)
    say(f"[wfr2] {W}x{H} sigma {sig} [{lo},{hi}] step {step}")
    ref, z, ms = run_both(oracle, f, sig, lo, hi, step)
    same = (np.abs(z["wx"] - ref["wx"]) < 1e-6) & (np.abs(z["wy"] - ref["wy"]) < 1e-6)
    rel_r = np.abs(z["r"] - ref["r"]).max() / ref["r"].max()
    e = wrap(z["phase"] - ref["phase"])[same]
    pc = wrap(z["phase_comp"] - ref["phase_comp"])[same]
    say(f"   {ms:.3f} ms; pair mismatch {1 - same.mean():.2e}; r rel {rel_r:.2e}; "
        f"phase rms {np.sqrt(np.mean(e * e)):.2e} max {np.abs(e).max():.2e}; "
        f"cxx max {np.abs(z['cxx'] - ref['cxx']).max():.2e} (scale {np.abs(ref['cxx']).max():.2e}); "
        f"phase_comp rms {np.sqrt(np.mean(pc * pc)):.2e}; "
        f"b rel {np.abs(z['b'] - ref['b']).max() / ref['b'].max():.2e}")
    assert 1 - same.mean() < 2e-3
    assert rel_r < 1e-5
    assert np.sqrt(np.mean(e * e)) < 1e-4
    # the curvatures are smooth sums of the pair maps: a pixel whose pair
    # switched moves its neighbours' fits by at most step * window weight
    assert np.abs(z["cxx"] - ref["cxx"]).max() < max(1e-4, 0.05 * step) or same.all() is False
    if same.all():
        assert np.abs(z["cxx"] - ref["cxx"]).max() < 1e-5
        assert np.abs(z["cyy"] - ref["cyy"]).max() < 1e-5
        assert np.sqrt(np.mean(pc * pc)) < 1e-4
        assert np.abs(z["b"] - ref["b"]).max() / ref["b"].max() < 1e-4


def test_wfr2_single_tone_and_zero_field():
    W, H, sig = 96, 80, 4.0
    y, x = np.mgrid[0:H, 0:W]
    f = (2.0 * np.exp(1j * (0.2 * x - 0.1 * y + 0.4))).astype(np.complex64)
    w = glsu.WFT2(W, H, type=glsu.WFR, params=glsu.wff_params(sig, -0.3, 0.3, 0.1))
    say("[wfr2 tone] 96x80 sigma 4")
    z = {k: v.cpu().numpy() for k, v in w(torch.from_numpy(f).cuda()).items()}
    s = int(round(3 * sig))
    inner = (slice(2 * s, H - 2 * s), slice(2 * s, W - 2 * s))
    assert np.allclose(z["wx"][inner], 0.2) and np.allclose(z["wy"][inner], -0.1)
    assert np.abs(wrap(z["phase"] - (0.2 * x - 0.1 * y + 0.4))[inner]).max() < 1e-5
    assert np.abs(z["cxx"][inner]).max() < 1e-5 and np.abs(z["cyy"][inner]).max() < 1e-5
    g = np.exp(-((np.arange(-s, s + 1)) ** 2) / 2 / sig / sig)
    gsum = (np.outer(g, g)).sum() / np.sqrt((np.outer(g, g) ** 2).sum())
    assert np.allclose(z["b"][inner], 2.0 * gsum / np.sqrt(4 * np.pi * sig * sig), rtol=1e-5)
    z0 = w(torch.zeros((H, W), dtype=torch.complex64, device="cuda"))
    for k, v in z0.items():
        assert torch.count_nonzero(v) == 0, k


def test_wfr2_host_entry_determinism_and_errors():
    W, H = 120, 90
    f = field(W, H, 5)
    p = glsu.wff_params(5.0, -0.4, 0.4, 0.1)
    w = glsu.WFT2(W, H, type=glsu.WFR, params=p)
    say("[wfr2 host/device] 120x90 sigma 5")
    d1 = w(torch.from_numpy(f).cuda())
    d2 = w(torch.from_numpy(f).cuda())
    for k in d1:
        assert torch.equal(d1[k], d2[k]), k
    host = np.empty((8, H, W), np.float32)
    fi = np.ascontiguousarray(f)
    glsu.check(glsu.lib().glsu_wft2_compute_wfr_host(w._h, fi.ctypes.data, host.ctypes.data, None))
    for i, k in enumerate(glsu.WFT2.WFR_FIELDS):
        assert np.array_equal(host[i], d1[k].cpu().numpy()), k
    # the WFF entry point refuses a WFR handle, and the other way round
    out = torch.empty((H, W), dtype=torch.complex64, device="cuda")
    with pytest.raises(glsu.GlsuError):
        glsu.check(glsu.lib().glsu_wft2_compute(w._h, ctypes.c_void_p(d1["r"].data_ptr()),
                                                ctypes.c_void_p(out.data_ptr()), None))
    wf = glsu.WFT2(W, H, params=p)
    with pytest.raises(glsu.GlsuError):
        wf.wfr(torch.from_numpy(f).cuda())


def test_wft2_outputs_checked_before_launch():
    """Outputs of the wrong device, shape or dtype are refused in Python before
    their address reaches the library, in both modes."""
    W, H = 120, 90
    p = glsu.wff_params(5.0, -0.4, 0.4, 0.1)
    f = torch.from_numpy(field(W, H, 7)).cuda()
    w = glsu.WFT2(W, H, type=glsu.WFR, params=p)
    say("[wft2 out checks] 120x90 WFR and WFF")
    for bad in (torch.empty((8, H, W), dtype=torch.float32),
                torch.empty((8, H, W + 1), dtype=torch.float32, device="cuda"),
                torch.empty((8, H, W), dtype=torch.float64, device="cuda")):
        with pytest.raises(ValueError):
            w.wfr(f, out=bad)
    wf = glsu.WFT2(W, H, params=p)
    for bad in (torch.empty((H, W), dtype=torch.complex64),
                torch.empty((H, W), dtype=torch.float32, device="cuda"),
                torch.empty((W, H), dtype=torch.complex64, device="cuda").t()):
        with pytest.raises(ValueError):
            wf(f, out=bad)
    good = wf(f)
    assert good.shape == (H, W) and bool(torch.isfinite(torch.view_as_real(good)).all())
