"""Learned WFF2 surrogate: a U-Net on BF16 MFMA kernels (libglsu), trained on
(raw DPRA field, WFF2-filtered field) pairs that the plain method generates.

The network maps z = exp(j Delta) (the raw DPRA phase difference, as
cos/sin channels) to the unit vector of the WFF2-filtered field
F = WFF2(z) (WFT/WFT2_CPU.cpp:190-390, called from DPRA/dpra_cpu.cpp:238).
Loss: mean |out - F/|F||^2.  At run time DPRA.frame_learned replaces the
WFF2 call by the network and keeps WFF2 as a per-frame spot check and
fallback.  Forward, backward and Adam all run in libglsu's HIP kernels; torch
only holds device memory.
"""
import ctypes
import os

import numpy as np
import torch

from ._lib import check, check_out, lib
from .params import wff_params
from . import synth


def layer_specs(levels=4, base=32, cin=2):
    """[(name, kind, Cin, Cout)] in parameter order (kind 'c3' or 'c1')."""
    ch = [base << l for l in range(levels)]
    specs, prev = [], cin
    for l in range(levels):
        specs.append((f"e{l}a", "c3", prev, ch[l]))
        specs.append((f"e{l}b", "c3", ch[l], ch[l]))
        prev = ch[l]
    for l in range(levels - 2, -1, -1):
        specs.append((f"d{l}a", "c3", ch[l] + ch[l + 1], ch[l]))
        specs.append((f"d{l}b", "c3", ch[l], ch[l]))
    specs.append(("out", "c1", ch[0], cin))
    return specs


class UNet:
    def __init__(self, levels=4, base=32, seed=0, stream=None):
        self.levels, self.base = int(levels), int(base)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        check(lib().glsu_unet_create(ctypes.byref(h), self.levels, self.base, int(seed),
                                     ctypes.c_void_p(self.stream.cuda_stream)))
        self._h = h
        c = ctypes.c_long()
        check(lib().glsu_unet_param_count(h, ctypes.byref(c)))
        self.count = c.value

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            try:
                lib().glsu_unet_destroy(h)
            except Exception:  # interpreter shutdown
                pass

    __del__ = close

    # ---- parameters -----------------------------------------------------
    def get_flat(self, grads=False):
        a = np.empty(self.count, np.float32)
        fn = lib().glsu_unet_get_grads if grads else lib().glsu_unet_get_params
        check(fn(self._h, a.ctypes.data))
        return a

    def set_flat(self, a):
        a = np.ascontiguousarray(a, np.float32)
        if a.size != self.count:
            raise ValueError(f"expected {self.count} parameters, got {a.size}")
        check(lib().glsu_unet_set_params(self._h, a.ctypes.data))

    def _to_dict(self, flat):
        out, i = {}, 0
        for name, kind, ci, co in layer_specs(self.levels, self.base):
            k = 3 if kind == "c3" else 1
            nw = co * k * k * ci
            out[name + ".w"] = flat[i:i + nw].reshape(co, k, k, ci)
            i += nw
            out[name + ".b"] = flat[i:i + co]
            i += co
        return out

    def params(self):
        return self._to_dict(self.get_flat())

    def grads(self):
        return self._to_dict(self.get_flat(grads=True))

    def set_params(self, d):
        flat = []
        for name, kind, ci, co in layer_specs(self.levels, self.base):
            flat.append(np.asarray(d[name + ".w"], np.float32).reshape(-1))
            flat.append(np.asarray(d[name + ".b"], np.float32).reshape(-1))
        self.set_flat(np.concatenate(flat))

    def save(self, path):
        np.savez(path, levels=self.levels, base=self.base, params=self.get_flat())

    @staticmethod
    def load(path, stream=None):
        z = np.load(path)
        net = UNet(int(z["levels"]), int(z["base"]), stream=stream)
        net.set_flat(z["params"])
        return net

    # ---- compute --------------------------------------------------------
    @staticmethod
    def _c64(t, name):
        if t.dtype != torch.complex64 or not t.is_cuda or not t.is_contiguous() or t.dim() != 3:
            raise ValueError(f"{name} must be a contiguous complex64 CUDA tensor (n, H, W)")
        return t

    def forward(self, x, out=None):
        self._c64(x, "x")
        if out is None:
            out = torch.empty_like(x)
        check_out(out, x.shape, torch.complex64)
        n, H, W = x.shape
        check(lib().glsu_unet_forward(self._h, ctypes.c_void_p(x.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), n, H, W))
        return out

    def loss_grad(self, x, u):
        self._c64(x, "x")
        self._c64(u, "u")
        if u.shape != x.shape:
            raise ValueError(f"u has shape {tuple(u.shape)}, expected x's {tuple(x.shape)}")
        n, H, W = x.shape
        loss = ctypes.c_double()
        check(lib().glsu_unet_loss_grad(self._h, ctypes.c_void_p(x.data_ptr()),
                                        ctypes.c_void_p(u.data_ptr()), n, H, W,
                                        ctypes.byref(loss)))
        return loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().glsu_unet_adam_step(self._h, float(lr), float(beta1), float(beta2),
                                        float(eps)))


# ------------------------------------------------------------------ data
def make_pairs(nseq=8, nframes=25, size=1024, seed0=900000, params=None, log=None, net=None,
               keep_every=1):
    """(x, u) pairs on synthetic sequences with held-out seeds: x = raw DPRA
    field z of a frame, u = unit vector of WFF2(z).  The chain that produces
    the reference phase is the plain method, or, with `net`, the learned one
    (so that the pairs cover the states the learned chain visits; the labels
    are WFF2 either way).  Every `keep_every`-th frame is kept.  Per
    sequence: peaks amplitude in U[2, 5] (+0.1 t), noise std in U[5, 25].
    Returns two complex64 CUDA tensors (frames, size, size)."""
    from .dpra import DPRA, gen_frames
    from .wft2 import WFT2
    params = params if params is not None else wff_params()
    rng = np.random.default_rng(seed0)
    xs, us = [], []
    w = WFT2(size, size, params=params)
    for s in range(nseq):
        amp0, noise = float(rng.uniform(2, 5)), float(rng.uniform(5, 25))
        fr = gen_frames(size, size, nframes, seed0 + s, amp0=amp0, noise=noise)
        d = DPRA(synth.phase(size, size, amp0, 0.05), params=params)
        for t in range(nframes):
            if net is None:
                d.dpra_per_frame(fr[t])
            else:
                d.frame_learned(fr[t], net, tol=-1.0)
            d.update_ref_phi()
            if t % keep_every:
                continue
            z = d.last_z()
            F = w(z)
            xs.append(z)
            us.append(F / F.abs().clamp_min(1e-30))
        if log:
            log(f"pairs: sequence {s + 1}/{nseq} (amp0 {amp0:.2f}, noise {noise:.1f}"
                f"{', learned chain' if net is not None else ''})")
    return torch.stack(xs), torch.stack(us)


def random_crops(x, u, batch, crop, gen):
    """A batch of random crop x crop windows from the pair tensors."""
    N, H, W = x.shape
    idx = torch.randint(0, N, (batch,), generator=gen)
    ys = torch.randint(0, H - crop + 1, (batch,), generator=gen)
    xs = torch.randint(0, W - crop + 1, (batch,), generator=gen)
    bx = torch.stack([x[i, a:a + crop, b:b + crop] for i, a, b in zip(idx.tolist(), ys.tolist(), xs.tolist())])
    bu = torch.stack([u[i, a:a + crop, b:b + crop] for i, a, b in zip(idx.tolist(), ys.tolist(), xs.tolist())])
    return bx.contiguous(), bu.contiguous()


def train(net, x, u, crops=20000, batch=16, crop=256, lr=1e-3, seed=0, log=None, log_every=100,
          warmup=50):
    """Adam on random crops of the pairs until `crops` crops have been seen.
    Returns the list of per-step losses."""
    gen = torch.Generator().manual_seed(seed)
    steps = (crops + batch - 1) // batch
    losses = []
    for k in range(steps):
        bx, bu = random_crops(x, u, batch, crop, gen)
        # linear warm-up, then cosine decay to 5 % of lr
        lr_k = lr * min(1.0, (k + 1) / max(warmup, 1)) * (
            0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * k / max(steps - 1, 1))))
        losses.append(net.loss_grad(bx, bu))
        net.adam_step(lr_k)
        if log and (k % log_every == 0 or k == steps - 1):
            log(f"step {k + 1}/{steps}: loss {losses[-1]:.5f} lr {lr_k:.2e}")
    return losses


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_l4_b32.npz")


def phase_rms(a, b):
    """RMS of the wrapped phase difference of two complex tensors (radians)."""
    d = torch.angle(a * b.conj())
    return float(torch.sqrt(torch.mean(d * d)))


def main(argv=None):
    """python -m glsu.learned --out PATH [--crops N ...]: generate training
    pairs with the plain method on held-out seeds and train the network."""
    import argparse
    import sys
    import time
    ap = argparse.ArgumentParser(prog="glsu.learned")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--crops", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--nseq", type=int, default=16)
    ap.add_argument("--nframes", type=int, default=25)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--learned-chain", type=int, default=0,
                    help="extra sequences whose reference phase follows the learned chain "
                         "of the --init network")
    ap.add_argument("--keep-every", type=int, default=1)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)
    t0 = time.perf_counter()
    net = UNet.load(a.init) if a.init else UNet(4, 32, seed=a.seed)
    x, u = make_pairs(a.nseq, a.nframes, a.size, log=log, keep_every=a.keep_every)
    if a.learned_chain:
        x2, u2 = make_pairs(a.learned_chain, a.nframes, a.size, seed0=950000, log=log, net=net,
                            keep_every=a.keep_every)
        x, u = torch.cat([x, x2]), torch.cat([u, u2])
        del x2, u2
    torch.cuda.synchronize()
    log(f"pairs: {tuple(x.shape)} in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    losses = train(net, x, u, a.crops, a.batch, a.crop, a.lr, a.seed, log=log)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    log(f"trained on {a.crops} crops of {a.crop}^2 in {dt:.1f} s "
        f"({a.crops * a.crop * a.crop / dt / 1e6:.1f} Mpx/s); final loss {np.mean(losses[-20:]):.5f}")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    np.save(os.path.splitext(a.out)[0] + "_losses.npy", np.asarray(losses, np.float32))
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
