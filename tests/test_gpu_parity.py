"""GPU parity: libglsu (HIP, FP32) against the FP64 CPU oracle on the same inputs.

Bars: wrapped phase error of the DPRA chain <= 1e-3 rad RMS per
frame.  Isolated WFF2 on identical inputs: relative field error <= 1e-4 and
phase error <= 1e-4 rad RMS.  DPRA least squares alone: <= 1e-4 rad RMS.
"""
import numpy as np
import pytest

from _util import say, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402


def _seq(O, W, H, T, seed, **kw):
    frames, phi0 = O.gen_sequence(W, H, T, seed, dtype=np.float32, **kw)
    return frames, phi0


@pytest.mark.parametrize("W,H,sig,lo,hi,st", [
    (256, 256, 10.0, -0.2, 0.2, 0.1),
    (200, 180, 10.0, -0.2, 0.2, 0.1),
    (64, 48, 2.0, -0.6, 0.6, 0.3),
    (371, 366, 10.0, -0.2, 0.2, 0.1),
])
def test_wft2_matches_oracle(oracle, W, H, sig, lo, hi, st):
    say(f"[wft2] {W}x{H} sigma={sig} grid=({lo},{hi},{st})")
    frames, phi0 = _seq(oracle, W, H, 2, seed=11)
    od = oracle.DPRA(phi0, sigx=sig, wxl=lo, wxh=hi, wxi=st)
    f = od.ls(frames[1].astype(np.float64)).astype(np.complex64)
    ref = oracle.WFF2(W, H, sig, lo, hi, st)(f.astype(np.complex128))
    w = glsu.WFT2(W, H, params=glsu.wff_params(sig, lo, hi, st))
    out = w(torch.from_numpy(f).cuda()).cpu().numpy()
    torch.cuda.synchronize()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    rms, mx = wrapped_err(np.angle(out), np.angle(ref))
    say(f"   rel={rel:.3e} phase rms={rms:.3e} max={mx:.3e} thr={w.info()['thr']:.6f}")
    assert rel < 1e-4
    assert rms < 1e-4


@pytest.mark.parametrize("W,H,x,y", [
    (300, 190, (6.0, -0.4, 0.4, 0.2), (12.0, -0.1, 0.2, 0.1)),
    (150, 260, (14.0, -0.15, 0.15, 0.05), (4.0, -0.6, 0.3, 0.3)),
])
def test_wft2_asymmetric_windows_match_oracle(oracle, W, H, x, y):
    """Different window width, frequency range and step along x and y: the
    row and column plans, padded lengths and window spectra differ per axis
    (the oracle is pinned for this case by the brute-force convolution)."""
    say(f"[wft2 asym] {W}x{H} x={x} y={y}")
    frames, phi0 = _seq(oracle, W, H, 2, seed=17)
    od = oracle.DPRA(phi0)
    f = od.ls(frames[1].astype(np.float64)).astype(np.complex64)
    ref = oracle.WFF2(W, H, x[0], x[1], x[2], x[3], sigy=y[0], wyl=y[1], wyh=y[2],
                      wyi=y[3])(f.astype(np.complex128))
    w = glsu.WFT2(W, H, params=glsu.wff_params(x[0], x[1], x[2], x[3], sigma_y=y[0], wyl=y[1],
                                               wyh=y[2], wyi=y[3]))
    out = w(torch.from_numpy(f).cuda()).cpu().numpy()
    torch.cuda.synchronize()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    rms, mx = wrapped_err(np.angle(out), np.angle(ref))
    say(f"   rel={rel:.3e} phase rms={rms:.3e} max={mx:.3e} thr={w.info()['thr']:.6f}")
    assert rel < 1e-4
    assert rms < 1e-4


@pytest.mark.parametrize("W,H", [(256, 256), (1024, 1024), (77, 53)])
def test_dpra_ls_matches_oracle(oracle, W, H):
    say(f"[dpra-ls] {W}x{H}")
    frames, phi0 = _seq(oracle, W, H, 2, seed=5)
    od = oracle.DPRA(phi0)
    zref = od.ls(frames[1].astype(np.float64))
    d = glsu.DPRA(phi0)
    d.dpra_per_frame(torch.from_numpy(frames[1]).cuda())
    z = d.last_z().cpu().numpy()
    rms, mx = wrapped_err(np.angle(z), np.angle(zref))
    say(f"   z phase rms={rms:.3e} max={mx:.3e}  |z|-1 max={np.abs(np.abs(z) - 1).max():.2e}")
    assert rms < 1e-4
    assert mx < 1e-2


@pytest.mark.parametrize("W,H,T,amp0,amp1,carrier", [
    (1024, 1024, 3, 3.0, 0.1, 0.05),      # the C2 settings, first frames
    (256, 256, 20, 1.0, 0.02, 0.15),      # carrier dominates the peaks gradient
    (512, 384, 8, 3.0, 0.1, 0.05),
])
def test_dpra_chain_matches_oracle(oracle, W, H, T, amp0, amp1, carrier):
    """Well-conditioned sequences (the local fringe frequency stays away from
    zero): strict bar, <= 1e-3 rad RMS on every frame."""
    say(f"[chain] {W}x{H} T={T} amp=({amp0},{amp1}) carrier={carrier}")
    frames, phi0 = _seq(oracle, W, H, T, seed=7, amp0=amp0, amp1=amp1, carrier=carrier)
    ref = oracle.DPRA(phi0).run(frames.astype(np.float64))
    out = glsu.DPRA(phi0)(torch.from_numpy(frames).cuda()).cpu().numpy()
    for t in range(T):
        rms, mx = wrapped_err(out[t], ref[t])
        say(f"   frame {t}: rms={rms:.3e} max={mx:.3e}")
        assert rms < 1e-3


def test_graph_equals_eager(oracle):
    """hipGraph replays of the frame chain equal the eager chain bit for bit;
    the same handle replays its captured graphs (calls 3 and 4), with eager
    work and synchronisation between the replays."""
    W, H, T = 192, 160, 6
    say(f"[graph] {W}x{H} T={T}")
    frames, phi0 = _seq(oracle, W, H, T, seed=3)
    fr = torch.from_numpy(frames).cuda()
    e = glsu.DPRA(phi0)
    g = glsu.DPRA(phi0)
    for rep in range(4):
        say(f"   call {rep}")
        a = e(fr).cpu()
        b = g(fr, graph=True).cpu()
        torch.cuda.synchronize()
        assert torch.equal(a, b), f"call {rep}: max diff {(a - b).abs().max().item()}"


def test_every_fft_plan_matches_oracle(oracle):
    """Each compiled FFT plan (length N = T*E), forced on both axes, gives the
    same WFF2 result as the FP64 oracle on an image it covers."""
    plans = glsu.fft_plans()
    sig, lo, hi, st = 2.0, -0.6, 0.6, 0.3
    for p in plans:
        W = min(p["N"] - 6, 300)
        H = max(min(p["N"] - 6, 120), 16)
        say(f"[plan {p['id']}] T={p['T']} E={p['E']} N={p['N']} image {W}x{H} "
            f"exchange={p['exchange']}")
        rng = np.random.default_rng(p["id"])
        ph = 0.6 * np.arange(W)[None, :] + 0.25 * np.arange(H)[:, None] + rng.normal(0, 0.3, (H, W))
        f = np.exp(1j * ph).astype(np.complex64)
        ref = oracle.WFF2(W, H, sig, lo, hi, st)(f.astype(np.complex128))
        # rows use one wavefront per transform: narrow plans run on the column axis only
        px = p["id"] if p["T"] == 64 else -1
        w = glsu.WFT2(W, H, params=glsu.wff_params(sig, lo, hi, st), planx=px, plany=p["id"])
        out = w(torch.from_numpy(f).cuda()).cpu().numpy()
        rel = np.abs(out - ref).max() / np.abs(ref).max()
        say(f"   rel={rel:.2e}")
        assert rel < 1e-4, p
