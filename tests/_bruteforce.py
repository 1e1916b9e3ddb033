"""Independent brute-force restatements used to pin the oracle (pure numpy,
small inputs only).  Written from the formulas, not from the oracle's code:

WFF2 (WFT/WFT2_CPU.cpp:190-390, MatLab/fp_wft2f.m:159-195):
  F = wxi*wyi/(4 pi^2) * sum_{wy,wx} w * T(w * f),  w(y,x) = g(y,x) e^{j(wx x + wy y)},
  g = exp(-x^2/2sx^2 - y^2/2sy^2) on [-3s,3s]^2 normalised to unit L2 norm,
  (a * w)(p) = sum_k a(k) w(p - k) (linear, 'same' crop), T(u) = u if |u| >= thr.
DPRA LS (DPRA/dpra_cpu.cpp:153-232, MatLab/fp_ls3u.m:71-90): per pixel, clipped
  3x3 neighbourhood, A = [[9, Sc, Ss], [Sc, Scc, Scs], [Ss, Scs, Sss]],
  b = [SI, SIc, SIs], x = A^-1 b (numpy.linalg.solve), Delta = atan2(-x2, x1).
"""
import numpy as np


def c_round(v):
    """C round(): halves away from zero (WFT/WFT2_CPU.cpp:701-702), unlike Python's round."""
    return int(np.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def window(sig_x, sig_y):
    sx, sy = c_round(3 * sig_x), c_round(3 * sig_y)
    x = np.arange(-sx, sx + 1)[None, :]
    y = np.arange(-sy, sy + 1)[:, None]
    g = np.exp(-x * x / 2 / sig_x ** 2 - y * y / 2 / sig_y ** 2)
    return g / np.sqrt((g * g).sum()), sx, sy


def conv_same(a, w, sx, sy):
    """r(p) = sum_k a(k) w(p - k) with w indexed by offsets (dy, dx) in [-sy,sy]x[-sx,sx]."""
    H, W = a.shape
    r = np.zeros_like(a)
    for iy, dy in enumerate(range(-sy, sy + 1)):
        for ix, dx in enumerate(range(-sx, sx + 1)):
            sh = np.zeros_like(a)
            ys0, ys1 = max(0, dy), min(H, H + dy)
            xs0, xs1 = max(0, dx), min(W, W + dx)
            if ys0 < ys1 and xs0 < xs1:
                sh[ys0:ys1, xs0:xs1] = a[ys0 - dy:ys1 - dy, xs0 - dx:xs1 - dx]
            r += sh * w[iy, ix]
    return r


def wff2(f, sig_x, wxl, wxh, wxi, sig_y=None, wyl=None, wyh=None, wyi=None, thr=-1.0):
    sig_y = sig_x if sig_y is None else sig_y
    wyl = wxl if wyl is None else wyl
    wyh = wxh if wyh is None else wyh
    wyi = wxi if wyi is None else wyi
    g, sx, sy = window(sig_x, sig_y)
    if thr < 0:
        thr = 6 * np.sqrt(np.mean(np.abs(f) ** 2) / 3)
    nx = int((wxh - wxl) * (1 / wxi)) + 1
    ny = int((wyh - wyl) * (1 / wyi)) + 1
    x = np.arange(-sx, sx + 1)[None, :]
    y = np.arange(-sy, sy + 1)[:, None]
    out = np.zeros(f.shape, complex)
    for iy in range(ny):
        wy = wyl + iy * wyi
        for ix in range(nx):
            wx = wxl + ix * wxi
            w = g * np.exp(1j * (wx * x + wy * y))
            s = conv_same(f.astype(complex), w, sx, sy)
            s = s * (np.abs(s) >= thr)
            out += conv_same(s, w, sx, sy)
    return out * wxi * wyi / 4 / np.pi ** 2, thr


def dpra_ls(I, phr):
    H, W = I.shape
    z = np.zeros((H, W), complex)
    for i in range(H):
        for j in range(W):
            ii = slice(max(i - 1, 0), min(i + 1, H - 1) + 1)
            jj = slice(max(j - 1, 0), min(j + 1, W - 1) + 1)
            c = np.cos(phr[ii, jj]).ravel()
            s = np.sin(phr[ii, jj]).ravel()
            f = I[ii, jj].ravel()
            A = np.array([[9, c.sum(), s.sum()], [c.sum(), (c * c).sum(), (c * s).sum()],
                          [s.sum(), (c * s).sum(), (s * s).sum()]])
            b = np.array([f.sum(), (f * c).sum(), (f * s).sum()])
            xs = np.linalg.solve(A, b)
            z[i, j] = np.exp(1j * np.arctan2(-xs[2], xs[1]))
    return z
