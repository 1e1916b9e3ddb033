"""Pins the FP64 oracle's WFR2 (orc_wfr2_compute) -- CPU only.

  brute force  MatLab/fp_wft2f.m:103-147 written literally with numpy/scipy on
               tiny inputs: sf = conv2(f, gwave) cropped, strict-> ridge update
               in (wy, wx) order, cxx = -conv2(padarray(wx, 'replicate'), x.*g,
               'same')/sum(x.*x.*g) cropped, the phase and amplitude
               compensation -- to 1e-11;
  known answers  a single tone on the grid: ridge at its frequency, phase equal
               to the tone's, zero curvature and unit amplitude b in the
               interior; a linear chirp: cxx near its rate; a zero field: r = 0
               and wx = wy = phase = 0 (strict >, no update);
  rules        the literal 2-D window and the separable form agree; the
               curvature padding's acceptance rule (OPT(H + 3 sy)).
"""
import numpy as np
import pytest
from scipy.signal import convolve2d


def window(sig_x, sig_y, sx, sy):
    yy, xx = np.mgrid[-sy:sy + 1, -sx:sx + 1].astype(np.float64)
    g = np.exp(-xx * xx / 2 / sig_x / sig_x - yy * yy / 2 / sig_y / sig_y)
    return g / np.sqrt((g * g).sum()), xx, yy


def wfr2_literal(f, sig, lo, hi, step):
    """fp_wft2f.m's WFR branch, with the C++ class's (row = y) geometry."""
    H, W = f.shape
    s = int(round(3 * sig))
    g, xx, yy = window(sig, sig, s, s)
    n = int((hi - lo) * (1 / step)) + 1
    r = np.zeros((H, W)); wx = np.zeros((H, W)); wy = np.zeros((H, W)); p = np.zeros((H, W))
    for iy in range(n):
        wyt = lo + iy * step
        for ix in range(n):
            wxt = lo + ix * step
            gw = g * np.exp(1j * (wxt * xx + wyt * yy))
            sf = convolve2d(f, gw, mode="full")[s:s + H, s:s + W]
            t = np.abs(sf) > r
            r = np.where(t, np.abs(sf), r)
            wx = np.where(t, wxt, wx)
            wy = np.where(t, wyt, wy)
            p = np.where(t, np.angle(sf), p)
    pad = lambda a: np.pad(a, s, mode="edge")
    cxx = -convolve2d(pad(wx), xx * g, mode="same")[s:s + H, s:s + W] / (xx * xx * g).sum()
    cyy = -convolve2d(pad(wy), yy * g, mode="same")[s:s + H, s:s + W] / (yy * yy * g).sum()
    pc = np.angle(np.exp(1j * (p - 0.5 * np.arctan(sig ** 2 * cxx) - 0.5 * np.arctan(sig ** 2 * cyy))))
    b = r * ((1 + sig ** 4 * cxx ** 2) / 4 / np.pi / sig ** 2) ** 0.25 \
          * ((1 + sig ** 4 * cyy ** 2) / 4 / np.pi / sig ** 2) ** 0.25
    return dict(wx=wx, wy=wy, phase=p, phase_comp=pc, b=b, r=r, cxx=cxx, cyy=cyy)


@pytest.mark.parametrize("W,H,sep", [(20, 18, True), (17, 23, False), (26, 20, True)])
def test_brute_force_tiny(oracle, W, H, sep):
    rng = np.random.default_rng(W * H)
    y, x = np.mgrid[0:H, 0:W]
    ph = 0.01 * (x - W / 2) ** 2 + 0.3 * x - 0.2 * y + rng.normal(0, 0.3, (H, W))
    f = (1 + 0.2 * rng.normal(size=(H, W))) * np.exp(1j * ph)
    sig, lo, hi, step = 2.0, -0.4, 0.4, 0.2
    ref = wfr2_literal(f, sig, lo, hi, step)
    z = oracle.WFR2(W, H, sigx=sig, wxl=lo, wxh=hi, wxi=step, separable=sep)(f)
    for k in ("r", "wx", "wy", "cxx", "cyy", "b"):
        assert np.abs(z[k] - ref[k]).max() < 1e-11, k
    for k in ("phase", "phase_comp"):
        assert np.abs(np.angle(np.exp(1j * (z[k] - ref[k])))).max() < 1e-11, k


def test_single_tone_known_answer(oracle):
    W, H, sig = 64, 56, 3.0
    y, x = np.mgrid[0:H, 0:W]
    wx0, wy0, p0 = 0.2, -0.1, 0.3
    f = 2.5 * np.exp(1j * (wx0 * x + wy0 * y + p0))
    z = oracle.WFR2(W, H, sigx=sig, wxl=-0.3, wxh=0.3, wxi=0.1)(f)
    s = int(round(3 * sig))
    inner = (slice(2 * s, H - 2 * s), slice(2 * s, W - 2 * s))
    assert np.allclose(z["wx"][inner], wx0) and np.allclose(z["wy"][inner], wy0)
    e = np.angle(np.exp(1j * (z["phase"] - (wx0 * x + wy0 * y + p0))))
    assert np.abs(e[inner]).max() < 1e-12
    assert np.abs(z["cxx"][inner]).max() < 1e-12 and np.abs(z["cyy"][inner]).max() < 1e-12
    # r = amplitude * sum(g) (g normalised to unit L2 norm); b recovers the amplitude
    g, _, _ = window(sig, sig, s, s)
    assert np.allclose(z["r"][inner], 2.5 * g.sum(), rtol=1e-12)
    assert np.allclose(z["b"][inner], 2.5 * g.sum() / np.sqrt(4 * np.pi * sig * sig), rtol=1e-12)
    # a sign slip in the phase (angle(conj sf)) would fail this
    assert np.abs(np.angle(np.exp(1j * (-z["phase"] - (wx0 * x + wy0 * y + p0))))[inner]).max() > 0.1


def test_chirp_curvature(oracle):
    W, H, sig, a = 160, 72, 4.0, 0.004
    y, x = np.mgrid[0:H, 0:W]
    f = np.exp(1j * (0.5 * a * (x - W / 2) ** 2))
    z = oracle.WFR2(W, H, sigx=sig, wxl=-0.4, wxh=0.4, wxi=0.02)(f)
    s = int(round(3 * sig))
    inner = (slice(2 * s, H - 2 * s), slice(3 * s, W - 3 * s))
    # local frequency a (x - W/2) on a 0.02 grid, its fitted slope ~ a
    assert np.abs(z["wx"][inner] - a * (x - W / 2)[inner]).max() <= 0.011
    assert abs(np.median(z["cxx"][inner]) - a) < 0.2 * a
    assert np.abs(z["cyy"][inner]).max() < 1e-12


def test_zero_field_and_padding_rule(oracle):
    z = oracle.WFR2(30, 20, sigx=2.0, wxl=-0.2, wxh=0.2, wxi=0.1)(np.zeros((20, 30)))
    for k in ("r", "wx", "wy", "phase", "cxx", "cyy", "b"):
        assert np.all(z[k] == 0), k
    # WFF accepts OPT(4020 + 60) = 4096, but WFR needs OPT(4020 + 90): none
    w = oracle.WFR2(4020, 8, sigx=10.0, wxl=-0.1, wxh=0.1, wxi=0.1)
    with pytest.raises(ValueError):
        w(np.zeros((8, 4020), np.complex128))
