"""Heavy WFF2 settings (sigma 12, wx, wy in [-0.5, 0.5] step 0.05: 441 pairs,
noise 20, carrier 0.08) against the FP64 oracle.

Full-size (2048^2) results are checked through the path's locality: every
output pixel of WFF2 depends only on its input within +-2s, and the DPRA least
squares only within +-1, so the oracle run on a window of (tile + 2(2s+1))
pixels around a tile reproduces the full-image result on that tile exactly
(same global threshold).  That pins the 2048^2 run tile by tile at a cost the
CPU oracle finishes in seconds.
"""
import numpy as np
import pytest

from _util import say, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402

SIG, LO, HI, ST = 12.0, -0.5, 0.5, 0.05
GEN = dict(amp0=3.0, amp1=0.1, carrier=0.08, noise=20.0)


def _params():
    return glsu.wff_params(SIG, LO, HI, ST)


def test_heavy_pair_grid():
    w = glsu.WFT2(300, 200, params=_params())
    i = w.info()
    assert (i["nwx"], i["nwy"], i["sx"], i["sy"]) == (21, 21, 36, 36)
    assert i["nx"] >= 300 + 36 and i["ny"] >= 200 + 36


def test_heavy_wft2_matches_oracle(oracle):
    W, H = 200, 176
    frames, phi0 = oracle.gen_sequence(W, H, 2, 31, dtype=np.float32, **GEN)
    od = oracle.DPRA(phi0, sigx=SIG, wxl=LO, wxh=HI, wxi=ST)
    f = od.ls(frames[1].astype(np.float64)).astype(np.complex64)
    ref = oracle.WFF2(W, H, SIG, LO, HI, ST)(f.astype(np.complex128))
    out = glsu.WFT2(W, H, params=_params())(torch.from_numpy(f).cuda()).cpu().numpy()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    rms, mx = wrapped_err(np.angle(out), np.angle(ref))
    say(f"[heavy wft2] {W}x{H} 441 pairs: rel={rel:.2e} phase rms={rms:.2e} max={mx:.2e}")
    assert rel < 1e-4 and rms < 1e-4


def test_heavy_chain_matches_oracle(oracle):
    W, H, T = 256, 256, 4
    frames, phi0 = oracle.gen_sequence(W, H, T, 32, dtype=np.float32, **GEN)
    ref = oracle.DPRA(phi0, sigx=SIG, wxl=LO, wxh=HI, wxi=ST).run(frames.astype(np.float64))
    out = glsu.DPRA(phi0, params=_params())(torch.from_numpy(frames).cuda()).cpu().numpy()
    for t in range(T):
        rms, mx = wrapped_err(out[t], ref[t])
        say(f"[heavy chain] frame {t}: rms={rms:.2e} max={mx:.2e}")
        assert rms < 1e-3


def _tiles(W, H, n=64):
    return [(0, 0), (W - n, H - n), (W // 2 - 100, H // 3), (W // 3, H - n - 300)]


def test_heavy_full_size_tiles_match_oracle(oracle):
    """2048^2, frame 0 of the heavy sequence: the GPU chain against the FP64
    oracle on (64 + 2(2s+1))^2 windows around four tiles (two at corners)."""
    W = H = 2048
    n, s = 64, 36
    m = 2 * s + 1
    seed = 33
    frames = glsu.gen_frames(W, H, 1, seed, **GEN)
    phi0 = glsu.synth.phase(W, H, GEN["amp0"], GEN["carrier"])
    d = glsu.DPRA(phi0, params=_params())
    out = d(frames).cpu().numpy()[0]
    host = frames.cpu().numpy()[0].astype(np.float64)
    for (x0, y0) in _tiles(W, H, n):
        wx0, wy0 = max(x0 - m, 0), max(y0 - m, 0)
        wx1, wy1 = min(x0 + n + m, W), min(y0 + n + m, H)
        od = oracle.DPRA(np.ascontiguousarray(phi0[wy0:wy1, wx0:wx1]),
                         sigx=SIG, wxl=LO, wxh=HI, wxi=ST)
        ref, _ = od.dpra_per_frame(np.ascontiguousarray(host[wy0:wy1, wx0:wx1]))
        r = ref[y0 - wy0:y0 - wy0 + n, x0 - wx0:x0 - wx0 + n]
        g = out[y0:y0 + n, x0:x0 + n]
        rms, mx = wrapped_err(g, r)
        say(f"[heavy 2048^2] tile ({x0},{y0}) window {wx1 - wx0}x{wy1 - wy0}: "
            f"rms={rms:.2e} max={mx:.2e}")
        assert rms < 1e-3


def test_near_singular_pixels_use_fp64_solve(oracle):
    """A reference phase with a stationary point (local fringe frequency -> 0):
    the 3x3 systems there are near-singular and are solved in FP64 from the
    compensated sums.  The GPU least squares must follow the FP64 oracle given
    the same float32 inputs."""
    W, H = 160, 128
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    phr = (2e-3 * ((x - 80) ** 2 + (y - 64) ** 2) + 0.3).astype(np.float32)
    delta = 0.4 + 0.05 * np.sin(x / 17.0)
    rng = np.random.default_rng(7)
    I = (128 + 100 * np.cos(phr.astype(np.float64) + delta) + rng.normal(0, 2, (H, W))).astype(np.float32)
    od = oracle.DPRA(phr.astype(np.float64))
    zref = od.ls(I.astype(np.float64))
    d = glsu.DPRA(phr.astype(np.float64))
    d.dpra_per_frame(torch.from_numpy(I).cuda())
    z = d.last_z().cpu().numpy()
    near = (np.abs(x - 80) <= 6) & (np.abs(y - 64) <= 6)
    rms_n, mx_n = wrapped_err(np.angle(z[near]), np.angle(zref[near]))
    rms_a, _ = wrapped_err(np.angle(z), np.angle(zref))
    say(f"[near-singular] {near.sum()} px near the stationary point: rms={rms_n:.2e} "
        f"max={mx_n:.2e}; whole image rms={rms_a:.2e}")
    assert rms_n < 2e-3
    assert rms_a < 1e-3
