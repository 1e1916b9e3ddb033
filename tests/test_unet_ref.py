"""Pins of the FP64 CPU U-Net reference (oracle/unet_ref.py): every parameter
gradient against central finite differences, and the pieces against
independent closed forms (identity at init, conv against direct sums)."""
import numpy as np
import pytest

from oracle import unet_ref as U


def _setup(levels=4, base=4, n=2, H=16, W=16, seed=0):
    p = U.init_params(levels, base, seed)
    rng = np.random.default_rng(seed + 1)
    p["out.w"] = rng.normal(0, 0.3, p["out.w"].shape)   # non-zero so gradients flow
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.normal(0, 0.05, p[k].shape)
    t = rng.uniform(-np.pi, np.pi, (n, H, W))
    x = np.stack([np.cos(t), np.sin(t)], -1)
    tf = t + rng.normal(0, 0.3, t.shape)
    u = np.stack([np.cos(tf), np.sin(tf)], -1)
    return p, x, u


def test_layer_specs_match_the_described_network():
    s = U.layer_specs()
    assert [x[0] for x in s] == ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b",
                                 "d2a", "d2b", "d1a", "d1b", "d0a", "d0b", "out"]
    assert s[0][2:] == (2, 32) and s[7][2:] == (256, 256)
    assert s[8][2:] == (128 + 256, 128) and s[12][2:] == (32 + 64, 32)
    assert s[-1] == ("out", "c1", 32, 2)


def test_untrained_net_is_identity():
    p = U.init_params(4, 4, 0)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 16, 16, 2))
    assert np.array_equal(U.forward(p, x), x)


def test_conv3_against_direct_sum():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, 5, 6, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    out = U.conv3(x, w, b)
    ref = np.zeros_like(out)
    for n in range(2):
        for y in range(5):
            for xx in range(6):
                for o in range(4):
                    s = b[o]
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            yy, x2 = y + dy, xx + dx
                            if 0 <= yy < 5 and 0 <= x2 < 6:
                                s += (w[o, dy + 1, dx + 1] * x[n, yy, x2]).sum()
                    ref[n, y, xx, o] = s
    assert np.abs(out - ref).max() < 1e-12


def test_every_gradient_matches_finite_differences():
    p, x, u = _setup()
    L, g = U.loss_and_grads(p, x, u)
    rng = np.random.default_rng(9)
    eps = 1e-6
    worst = 0.0
    for name in p:
        flat = p[name].reshape(-1)
        idx = rng.choice(flat.size, size=min(4, flat.size), replace=False)
        for i in idx:
            old = flat[i]
            flat[i] = old + eps
            lp, _ = U.loss_and_grads(p, x, u)
            flat[i] = old - eps
            lm, _ = U.loss_and_grads(p, x, u)
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            an = g[name].reshape(-1)[i]
            err = abs(fd - an) / max(1e-6, abs(fd) + abs(an))
            worst = max(worst, err)
            assert err < 1e-4, (name, i, fd, an)
    assert worst < 1e-4


def test_maxpool_and_up_adjoints():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, 8, 6, 3))
    g = rng.normal(size=(2, 4, 3, 3))
    # <maxpool_bwd(x, g), dx> == <g, d maxpool(x)[dx]> for a small perturbation
    dx = rng.normal(size=x.shape) * 1e-7
    lhs = (U.maxpool2_bwd(x, g) * dx).sum()
    rhs = (g * (U.maxpool2(x + dx) - U.maxpool2(x))).sum()
    assert abs(lhs - rhs) < 1e-12
    h = rng.normal(size=(2, 4, 3, 3))
    k = rng.normal(size=(2, 8, 6, 3))
    assert abs((U.up2(h) * k).sum() - (h * U.up2_bwd(k)).sum()) < 1e-12
