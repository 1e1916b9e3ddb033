"""FP64 CPU reference of the learned WFF2 surrogate (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's baseline leg may import
this module; the library never does.  The network has no counterpart in the
reference codebase: it is pinned here by central finite differences
(tests/test_unet_ref.py), and the GPU kernels are checked against it.

Architecture (NHWC tensors, batch n):
  level channels C_l = base * 2^l, l = 0..levels-1 (32, 64, 128, 256)
  encoder l:  e_l = relu(conv3(relu(conv3(in_l)))),  in_0 = x (2 ch),
              in_{l+1} = maxpool2(e_l); the last level is the bottleneck
  decoder l (from levels-2 down to 0):
              d_l = relu(conv3(relu(conv3(cat(e_l, up2(d_{l+1}))))))
              (d_{levels-1} = e_{levels-1}; up2 = nearest-neighbour x2)
  out = x + conv1(d_0)   (2 channels: filtered cos, sin; residual on the input)
conv3: 3x3, zero padding 1, weights [Cout][3][3][Cin] + bias [Cout].
Loss: L = mean over pixels of |out - u|^2, u = (cos t, sin t) the unit
vector of the WFF2-filtered phase.
"""
import numpy as np


def layer_specs(levels=4, base=32, cin=2):
    """[(name, kind, Cin, Cout)] in forward order; kind 'c3' or 'c1'."""
    ch = [base << l for l in range(levels)]
    specs = []
    prev = cin
    for l in range(levels):
        specs.append((f"e{l}a", "c3", prev, ch[l]))
        specs.append((f"e{l}b", "c3", ch[l], ch[l]))
        prev = ch[l]
    for l in range(levels - 2, -1, -1):
        specs.append((f"d{l}a", "c3", ch[l] + ch[l + 1], ch[l]))
        specs.append((f"d{l}b", "c3", ch[l], ch[l]))
    specs.append(("out", "c1", ch[0], cin))
    return specs


def init_params(levels=4, base=32, seed=0, cin=2):
    """He-normal weights, zero biases; the 1x1 output layer starts at zero so
    the untrained net is the identity (out = x)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, kind, ci, co in layer_specs(levels, base, cin):
        k = 3 if kind == "c3" else 1
        std = np.sqrt(2.0 / (k * k * ci))
        w = rng.normal(0.0, std, (co, k, k, ci)) if name != "out" else np.zeros((co, k, k, ci))
        p[name + ".w"] = w
        p[name + ".b"] = np.zeros(co)
    return p


def conv3(x, w, b):
    n, H, W, ci = x.shape
    xp = np.zeros((n, H + 2, W + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    out = np.broadcast_to(b, (n, H, W, w.shape[0])).copy()
    for dy in range(3):
        for dx in range(3):
            out += xp[:, dy:dy + H, dx:dx + W] @ w[:, dy, dx, :].T
    return out


def conv3_bwd(x, w, g):
    """Gradients of sum(g * conv3(x, w, b)) w.r.t. x, w, b."""
    n, H, W, ci = x.shape
    xp = np.zeros((n, H + 2, W + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    gx = np.zeros_like(xp)
    gw = np.zeros_like(w)
    for dy in range(3):
        for dx in range(3):
            gw[:, dy, dx, :] = np.einsum("nhwo,nhwi->oi", g, xp[:, dy:dy + H, dx:dx + W])
            gx[:, dy:dy + H, dx:dx + W] += g @ w[:, dy, dx, :]
    return gx[:, 1:-1, 1:-1], gw, g.sum(axis=(0, 1, 2))


def maxpool2(x):
    n, H, W, c = x.shape
    v = x.reshape(n, H // 2, 2, W // 2, 2, c)
    return v.max(axis=(2, 4))


def maxpool2_bwd(x, g):
    """Route g to the first maximum of each 2x2 window (row-major order)."""
    n, H, W, c = x.shape
    v = x.reshape(n, H // 2, 2, W // 2, 2, c).transpose(0, 1, 3, 5, 2, 4).reshape(n, H // 2, W // 2, c, 4)
    idx = v.argmax(axis=-1)
    gx = np.zeros((n, H // 2, W // 2, c, 4))
    np.put_along_axis(gx, idx[..., None], g[..., None], axis=-1)
    gx = gx.reshape(n, H // 2, W // 2, c, 2, 2).transpose(0, 1, 4, 2, 5, 3)
    return gx.reshape(n, H, W, c)


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up2_bwd(g):
    n, H, W, c = g.shape
    return g.reshape(n, H // 2, 2, W // 2, 2, c).sum(axis=(2, 4))


def forward(p, x, levels=4, keep=False):
    """x: (n, H, W, 2) -> out (n, H, W, 2); with keep=True also the cache."""
    cache = {}
    h = x
    enc = []
    for l in range(levels):
        a = np.maximum(conv3(h, p[f"e{l}a.w"][:, :, :, :], p[f"e{l}a.b"]), 0)
        e = np.maximum(conv3(a, p[f"e{l}b.w"], p[f"e{l}b.b"]), 0)
        cache[f"e{l}a_in"], cache[f"e{l}a"], cache[f"e{l}"] = h, a, e
        enc.append(e)
        if l < levels - 1:
            h = maxpool2(e)
    d = enc[-1]
    for l in range(levels - 2, -1, -1):
        c = np.concatenate([enc[l], up2(d)], axis=-1)
        a = np.maximum(conv3(c, p[f"d{l}a.w"], p[f"d{l}a.b"]), 0)
        d = np.maximum(conv3(a, p[f"d{l}b.w"], p[f"d{l}b.b"]), 0)
        cache[f"d{l}a_in"], cache[f"d{l}a"], cache[f"d{l}"] = c, a, d
    w1 = p["out.w"][:, 0, 0, :]
    out = x + d @ w1.T + p["out.b"]
    cache["d0_final"] = d
    return (out, cache) if keep else out


def loss_and_grads(p, x, u, levels=4):
    """L = mean_pixels |out - u|^2 and dL/dparams (dict like p)."""
    out, c = forward(p, x, levels, keep=True)
    n, H, W, _ = x.shape
    npx = n * H * W
    r = out - u
    L = float((r * r).sum() / npx)
    g_out = 2.0 * r / npx
    gp = {}
    d0 = c["d0_final"]
    w1 = p["out.w"][:, 0, 0, :]
    gp["out.w"] = np.einsum("nhwk,nhwc->kc", g_out, d0)[:, None, None, :]
    gp["out.b"] = g_out.sum(axis=(0, 1, 2))
    g = g_out @ w1                                   # d L / d d_0 (post-relu)
    g_enc = {l: None for l in range(levels)}
    ch = [p[f"e{l}b.b"].shape[0] for l in range(levels)]
    for l in range(0, levels - 1):
        # decoder level l: d_l = relu(conv(relu(conv(cat))))
        g = g * (c[f"d{l}"] > 0)
        ga, gp[f"d{l}b.w"], gp[f"d{l}b.b"] = conv3_bwd(c[f"d{l}a"], p[f"d{l}b.w"], g)
        ga = ga * (c[f"d{l}a"] > 0)
        gc, gp[f"d{l}a.w"], gp[f"d{l}a.b"] = conv3_bwd(c[f"d{l}a_in"], p[f"d{l}a.w"], ga)
        g_enc[l] = gc[..., :ch[l]]
        g = up2_bwd(gc[..., ch[l]:])                  # d L / d d_{l+1}
    # g is now d L / d e_{levels-1} (the bottleneck output) from the decoder
    for l in range(levels - 1, -1, -1):
        ge = g if l == levels - 1 else g_enc[l] + g
        ge = ge * (c[f"e{l}"] > 0)
        ga, gp[f"e{l}b.w"], gp[f"e{l}b.b"] = conv3_bwd(c[f"e{l}a"], p[f"e{l}b.w"], ge)
        ga = ga * (c[f"e{l}a"] > 0)
        gi, gp[f"e{l}a.w"], gp[f"e{l}a.b"] = conv3_bwd(c[f"e{l}a_in"], p[f"e{l}a.w"], ga)
        if l > 0:
            g = maxpool2_bwd(c[f"e{l - 1}"], gi)       # d L / d e_{l-1} via the pool
    return L, gp
