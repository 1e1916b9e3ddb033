"""Pins of the oracle's WFF2 (WFT/WFT2_CPU.cpp:190-390, init :697-845,
threshold :977-1022): brute-force direct convolution, closed forms and
invariants.  Also pins the exactness claims the GPU design rests on
(centred circular convolution of any length N >= size + s equals the linear
one; separable 1-D spectra equal FFT2 of the 2-D window; the centred window's
spectra are real)."""
import numpy as np
import pytest

import _bruteforce as bf


def _field(H, W, seed=0):
    rng = np.random.default_rng(seed)
    ph = 0.7 * np.arange(W)[None, :] + 0.35 * np.arange(H)[:, None] + rng.normal(0, 0.4, (H, W))
    return np.exp(1j * ph)


@pytest.mark.parametrize("H,W,sig,lo,hi,st,thr", [
    (24, 20, 2.0, -0.6, 0.6, 0.3, -1.0),
    (17, 23, 1.5, -1.0, 1.0, 0.5, -1.0),
    (20, 20, 2.0, -0.6, 0.6, 0.3, 2.5),
    (12, 30, 1.0, -0.3, 0.9, 0.6, 0.0),
])
@pytest.mark.parametrize("separable", [True, False])
def test_wff2_bruteforce(oracle, H, W, sig, lo, hi, st, thr, separable):
    f = _field(H, W, seed=H * W)
    ref, t = bf.wff2(f, sig, lo, hi, st, thr=thr)
    w = oracle.WFF2(W, H, sig, lo, hi, st, thr=thr, separable=separable, nthreads=3)
    out = w(f)
    assert abs(w.thr - t) < 1e-14
    assert np.abs(out - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_wff2_nonsquare_sigma(oracle):
    f = _field(18, 26, 3)
    ref, _ = bf.wff2(f, 2.0, -0.5, 0.5, 0.5, sig_y=1.5, wyl=-0.4, wyh=0.4, wyi=0.4)
    out = oracle.WFF2(26, 18, 2.0, -0.5, 0.5, 0.5, 1.5, -0.4, 0.4, 0.4, nthreads=2)(f)
    assert np.abs(out - ref).max() < 1e-12


def test_wff2_infinite_threshold_is_zero(oracle):
    f = _field(32, 40, 4)
    out = oracle.WFF2(40, 32, 3.0, -0.3, 0.3, 0.3, thr=np.inf)(f)
    assert np.abs(out).max() == 0.0


def test_wff2_conjugation_symmetry(oracle):
    # symmetric grid: WFF(conj f) = conj WFF(f)
    f = _field(40, 48, 5)
    w1 = oracle.WFF2(48, 40, 3.0, -0.4, 0.4, 0.2, nthreads=2)
    w2 = oracle.WFF2(48, 40, 3.0, -0.4, 0.4, 0.2, nthreads=2)
    a = w1(f)
    b = w2(np.conj(f))
    assert np.abs(b - np.conj(a)).max() < 1e-12


def test_wff2_single_tone_known_answer(oracle):
    # f = exp(j(wx x + wy y)) with (wx, wy) on the grid, no noise: the phase of
    # the filtered field equals the phase of f away from the border (>= 2s)
    H, W, sig = 96, 100, 3.0
    wx, wy = 0.2, -0.2
    x = np.arange(W)[None, :]
    y = np.arange(H)[:, None]
    f = np.exp(1j * (wx * x + wy * y))
    out = oracle.WFF2(W, H, sig, -0.4, 0.4, 0.2)(f)
    s = bf.c_round(3 * sig)
    inner = (slice(2 * s, H - 2 * s), slice(2 * s, W - 2 * s))
    e = np.angle(out[inner] * np.conj(f[inner]))
    assert np.abs(e).max() < 1e-10


def test_wff2_threshold_closed_form_and_cached(oracle):
    # thr = 6 sqrt(mean |f|^2 / 3) on the first call, then kept (m_rThr >= 0)
    f = 0.5 * _field(20, 24, 6)
    w = oracle.WFF2(24, 20, 2.0, -0.6, 0.6, 0.3)
    w(f)
    t1 = w.thr
    assert abs(t1 - 6 * np.sqrt(np.mean(np.abs(f) ** 2) / 3)) < 1e-14
    w(3.0 * f)
    assert w.thr == t1
    # unit-modulus fields give 2 sqrt(3) for every frame (the DPRA case)
    w2 = oracle.WFF2(24, 20, 2.0, -0.6, 0.6, 0.3)
    w2(_field(20, 24, 7))
    assert abs(w2.thr - 2 * np.sqrt(3)) < 1e-14


def test_wff2_separable_equals_literal_fft2_gwave(oracle):
    f = _field(50, 44, 8)
    a = oracle.WFF2(44, 50, 4.0, -0.3, 0.3, 0.15, separable=True)(f)
    b = oracle.WFF2(44, 50, 4.0, -0.3, 0.3, 0.15, separable=False)(f)
    assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()


def _centred_spectrum(sig, w, N, shift=0):
    """(unscaled) DFT of the centred, modulated window g(d) e^{jwd}, d in
    [-s, s], placed at d + shift (mod N)."""
    s = bf.c_round(3 * sig)
    d = np.arange(-s, s + 1)
    g = np.exp(-(d * d) / 2 / sig / sig)
    g /= np.sqrt((g * g).sum())
    k = np.arange(N)[:, None]
    return (g[None, :] * np.exp(1j * (w * d[None, :] - 2 * np.pi * ((k * (d[None, :] + shift)) % N) / N))).sum(1)


def _centred_separable_wff2(f, sig, lo, hi, st, Nx, Ny, thr, real=False):
    """The GPU formulation in numpy: centred circular convolutions of lengths
    Nx (rows) and Ny (columns), pair loop regrouped as rows->columns->rows.
    real: keep only the real part of the window spectra, as the GPU does."""
    H, W = f.shape
    s = bf.c_round(3 * sig)
    n = int((hi - lo) * (1 / st)) + 1
    d = np.arange(-s, s + 1)
    g = np.exp(-(d * d) / 2 / sig / sig)
    g /= np.sqrt((g * g).sum())

    def spec(w, N):
        k = np.arange(N)[:, None]
        return (g[None, :] * np.exp(1j * (w * d[None, :] - 2 * np.pi * ((k * d[None, :]) % N) / N))).sum(1)

    ws = [lo + i * st for i in range(n)]
    Hx = [spec(w, Nx) for w in ws]
    Hy = [spec(w, Ny) for w in ws]
    if real:
        Hx = [h.real.astype(complex) for h in Hx]
        Hy = [h.real.astype(complex) for h in Hy]
    fp = np.zeros((H, Nx), complex)
    fp[:, :W] = f
    Z = np.fft.fft(fp, axis=1)
    F = np.zeros((H, Nx), complex)
    for kx in range(n):
        A = np.fft.ifft(Z * Hx[kx][None, :], axis=1)[:, :W]
        Ap = np.zeros((Ny, W), complex)
        Ap[:H] = A
        Ah = np.fft.fft(Ap, axis=0)
        acc = np.zeros_like(Ah)
        for ky in range(n):
            S = np.fft.ifft(Ah * Hy[ky][:, None], axis=0)
            S[H:] = 0
            S = S * (np.abs(S) >= thr)
            acc += np.fft.fft(S, axis=0) * Hy[ky][:, None]
        D = np.fft.ifft(acc, axis=0)[:H]
        Dp = np.zeros((H, Nx), complex)
        Dp[:, :W] = D
        F += np.fft.fft(Dp, axis=1) * Hx[kx][None, :]
    return np.fft.ifft(F, axis=1)[:, :W] * st * st / 4 / np.pi ** 2


@pytest.mark.parametrize("Nx,Ny", [(46, 38), (48, 40), (64, 64), (97, 53)])
def test_centred_circular_any_length_is_exact(oracle, Nx, Ny):
    # any N >= size + s reproduces the reference's OPT(size + 2s) top-left result
    H, W, sig = 32, 40, 2.0
    f = _field(H, W, 9)
    ref = oracle.WFF2(W, H, sig, -0.6, 0.6, 0.3)(f)
    out = _centred_separable_wff2(f, sig, -0.6, 0.6, 0.3, Nx, Ny, 2 * np.sqrt(3))
    assert np.abs(out - ref).max() < 1e-12


def test_centred_circular_too_short_is_not_exact(oracle):
    # N = size + s - 1 wraps one sample: the check above would catch a wrong bound
    H, W, sig = 32, 40, 2.0
    f = _field(H, W, 9)
    ref = oracle.WFF2(W, H, sig, -0.6, 0.6, 0.3)(f)
    out = _centred_separable_wff2(f, sig, -0.6, 0.6, 0.3, W + 5, H + 5, 2 * np.sqrt(3))
    assert np.abs(out - ref).max() > 1e-6


def test_padding_error(oracle):
    # OPT(W + 2s) > 4096 -> the reference throws int -1; the oracle returns NULL
    with pytest.raises(ValueError):
        oracle.WFF2(4096 - 60 + 2, 64, 10.0)
    oracle.WFF2(4096 - 60, 16, 10.0, -0.1, 0.1, 0.1)  # boundary: fits


@pytest.mark.parametrize("sig,w,N", [(10.0, -0.2, 1280), (10.0, 0.1, 1280), (12.0, 0.45, 2304),
                                     (10.0, -2.0, 1280), (2.0, 0.3, 46)])
def test_centred_window_spectrum_is_real(sig, w, N):
    # the window is even, so the sine terms of d and -d cancel for every bin
    h = _centred_spectrum(sig, w, N)
    assert np.abs(h.imag).max() < 1e-13 * np.abs(h).max()
    # a window placed one sample off centre carries a linear phase: not real
    h1 = _centred_spectrum(sig, w, N, shift=1)
    assert np.abs(h1.imag).max() > 1e-3 * np.abs(h1).max()


@pytest.mark.parametrize("Nx,Ny", [(46, 38), (97, 53)])
def test_real_window_spectra_reproduce_the_oracle(oracle, Nx, Ny):
    # the GPU stores only the real part of the spectra; the result is unchanged
    H, W, sig = 32, 40, 2.0
    f = _field(H, W, 9)
    ref = oracle.WFF2(W, H, sig, -0.6, 0.6, 0.3)(f)
    out = _centred_separable_wff2(f, sig, -0.6, 0.6, 0.3, Nx, Ny, 2 * np.sqrt(3), real=True)
    assert np.abs(out - ref).max() < 1e-12
