"""The batched (64 uneven sequences) and learned (trained U-Net at 1024^2)
workloads at their configured sizes, and the WFF2 threshold's first-call
caching on the GPU.

* Batched: 64 sequences, sizes drawn from {256, 512, 768, 1024, 1536,
  2048}^2, 30 frames each, per-sequence seed, amplitude U[2, 5] + 0.1 t and
  noise U[5, 25]; the same draw as bench.py's batch64 workload.  One sequence
  of every size class must equal a single-sequence handle bit for bit, and
  the last frame of a 2048^2 sequence is checked against the FP64 oracle's
  one step on tile windows (locality: every output pixel of a step depends
  only on its input within +-(2s + 1)).
* Learned: the committed weights on the 1024^2 learned chain with the WFF2
  spot check and fallback, 120 frames; the network's phase against the FP64
  WFF2 of the same raw field, and the chain's own per-frame filtered phase.
* Threshold: WFT/WFT2_CPU.cpp:1014-1015 computes thr = 6 sqrt(mean|f|^2/3)
  on the first call when thr < 0 and keeps it for later calls.
"""
import os

import numpy as np
import pytest

from _util import say, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402

SIZES = [256, 512, 768, 1024, 1536, 2048]


def _wrap(a):
    return np.angle(np.exp(1j * a))


def _batch_workload(T):
    """bench.py's batch64 draw (seed 20240607)."""
    n = 64
    rng = np.random.default_rng(20240607)
    sizes = [int(v) for v in rng.choice(SIZES, n)]
    amp0 = rng.uniform(2.0, 5.0, n)
    noise = rng.uniform(5.0, 25.0, n)
    frames, phis = [], []
    for i, S in enumerate(sizes):
        frames.append(glsu.gen_frames(S, S, T, 1000 + i, amp0=float(amp0[i]), amp1=0.1,
                                      carrier=0.05, noise=float(noise[i])))
        phis.append(glsu.synth.phase(S, S, float(amp0[i]), 0.05))
    return sizes, amp0, noise, frames, phis


def test_c4_batched_64_uneven_sequences(oracle):
    T = 30
    sizes, amp0, noise, frames, phis = _batch_workload(T)
    say(f"[c4] 64 sequences, sizes {sizes}, {T} frames")
    b = glsu.DPRABatch(phis, 1)
    outs = b(frames, graph=True, timed=True)
    px = sum(S * S for S in sizes) * T
    say(f"[c4] batched run: {b.last_ms:.1f} ms for {64 * T} frames ({px / b.last_ms / 1e6:.2f} "
        f"Gpx/s), classes {b.info()['classes']}")
    # one sequence of every size class against a single-sequence handle
    for S in SIZES:
        if S not in sizes:
            continue
        i = sizes.index(S)
        d = glsu.DPRA(phis[i], 1, tuning=b.tuning(i))
        ref = d(frames[i])
        same = torch.equal(outs[i], ref)
        say(f"[c4] seq {i} ({S}^2): bitwise equal to a single handle = {same}")
        assert same, (i, (outs[i] - ref).abs().max().item())
        del d, ref
    # last frame of a 2048^2 sequence: one step of the oracle from the GPU's
    # reference phase, on tile windows
    assert 2048 in sizes
    i = sizes.index(2048)
    S, n, s = 2048, 64, 30
    m = 2 * s + 1
    d = glsu.DPRA(phis[i], 1, tuning=b.tuning(i))
    d(frames[i][:T - 1].contiguous())
    phr, _ = d.get_state(dphiref=False)
    bphr, _ = b.get_state(i)  # the batch has run all T frames: compare after one more step
    pc = torch.empty((S, S), dtype=torch.float32, device="cuda")
    d.dpra_per_frame(frames[i][T - 1], phicurr=pc)
    d.update_ref_phi()
    assert np.array_equal(d.get_state(dphiref=False)[0], bphr)
    inc = _wrap(pc.cpu().numpy().astype(np.float64) - phr.astype(np.float64))
    host = frames[i][T - 1].cpu().numpy().astype(np.float64)
    for (x0, y0) in [(0, 0), (S - n, S - n), (S // 2 - 37, S // 3 + 11), (S - n, 5)]:
        wx0, wy0 = max(x0 - m, 0), max(y0 - m, 0)
        wx1, wy1 = min(x0 + n + m, S), min(y0 + n + m, S)
        od = oracle.DPRA(np.ascontiguousarray(phr[wy0:wy1, wx0:wx1], np.float64))
        ref, _ = od.dpra_per_frame(np.ascontiguousarray(host[wy0:wy1, wx0:wx1]))
        r = ref[y0 - wy0:y0 - wy0 + n, x0 - wx0:x0 - wx0 + n]
        rms, mx = wrapped_err(inc[y0:y0 + n, x0:x0 + n], r)
        say(f"[c4] seq {i} (2048^2, amp0 {amp0[i]:.2f}, noise {noise[i]:.1f}) frame {T - 1} "
            f"tile ({x0},{y0}): rms={rms:.2e} max={mx:.2e}")
        assert rms < 1e-3


def test_c5_learned_chain_with_committed_weights(oracle):
    path = glsu.learned.WEIGHTS
    if not os.path.exists(path):
        pytest.skip("no committed weights")
    W = H = 1024
    T = 120
    tol = 0.03
    rng = np.random.default_rng(31337)  # bench.py's learned sequence
    amp0, noise = float(rng.uniform(2, 5)), float(rng.uniform(5, 25))
    params = glsu.wff_params(10.0, -0.2, 0.2, 0.1)
    frames = glsu.gen_frames(W, H, T, 424242, amp0=amp0, noise=noise)
    phi0 = glsu.synth.phase(W, H, amp0, 0.05)
    net = glsu.UNet.load(path)
    d = glsu.DPRA(phi0, params=params)
    out = torch.empty((T, H, W), dtype=torch.float32, device="cuda")
    sample = [0, 1, 2, 30, 60, 90, T - 1]
    rows = []
    for t in range(T):
        tile = ((t * 389) % (W - 64), (t * 613) % (H - 64))
        res, fb = d.frame_learned(frames[t], net, out[t], tol=tol, tile=tile)
        if t in sample:
            z = d.last_z().cpu().numpy().astype(np.complex128)
            y = d.last_net().cpu().numpy().astype(np.complex128)
            aF = np.angle(oracle.WFF2(W, H, 10.0, -0.2, 0.2, 0.1, pair_parallel=True)(z))
            o = out[t].cpu().numpy().astype(np.float64)
            inc = o if t == 0 else o - out[t - 1].cpu().numpy().astype(np.float64)
            net_rms = wrapped_err(np.angle(y), aF)[0]
            raw_rms = wrapped_err(np.angle(z), aF)[0]
            chain_rms = wrapped_err(inc, aF)[0]
            rows.append((t, fb, res, net_rms, raw_rms, chain_rms))
            say(f"[c5] frame {t}: fallback={fb} spot-check {res:.4f} | vs FP64 WFF2: net "
                f"{net_rms:.4f}, unfiltered {raw_rms:.4f}, chain {chain_rms:.4f} rad RMS")
        d.update_ref_phi()
    st = d.learned_stats()
    say(f"[c5] {T} frames, noise {noise:.2f}, amp0 {amp0:.2f}: {st['fallbacks']} fallbacks, "
        f"spot-check mean {st['mean_rms']:.4f} max {st['max_rms']:.4f}")
    t0 = rows[0]
    assert t0[3] <= 0.02, f"net vs FP64 WFF2 on frame 0: {t0[3]}"
    assert t0[4] / t0[3] > 5.0, "gain over the unfiltered phase"
    for t, fb, res, net_rms, raw_rms, chain_rms in rows:
        # the chain's filtered phase is the network's where the spot check
        # passed and WFF2's where the frame fell back
        assert chain_rms <= (1e-3 if fb else net_rms + 1e-3)
        assert raw_rms / chain_rms > 5.0


@pytest.mark.parametrize("W,H", [(512, 512), (640, 384)])
def test_spot_check_residual_equals_full_frame_wff2(W, H):
    """The fallback's residual is computed from WFF2 on the tile's window
    (tile +- 2s, clipped; its rows read in place from the W-wide field); it
    must equal the RMS of angle(net * conj(F)) on the tile with F the
    full-frame WFF2 of the same raw field, on square and non-square images."""
    path = glsu.learned.WEIGHTS
    net = glsu.UNet.load(path) if os.path.exists(path) else glsu.UNet(4, 32, seed=4)
    params = glsu.wff_params(10.0, -0.2, 0.2, 0.1)
    frames = glsu.gen_frames(W, H, 1, 4711, amp0=3.5, noise=15.0)
    phi0 = glsu.synth.phase(W, H, 3.5, 0.05)
    for tile in [(0, 0), (W - 64, H - 64), (200, 131), (17, H - 64), (W - 64, 300)]:
        d = glsu.DPRA(phi0, params=params)
        res, fb = d.frame_learned(frames[0], net, tol=1e9, tile=tile)
        z = d.last_z()
        y = d.last_net().cpu().numpy()
        F = glsu.WFT2(W, H, params=params)(z).cpu().numpy()
        x0, y0 = tile
        a = y[y0:y0 + 64, x0:x0 + 64].astype(np.complex128)
        b = F[y0:y0 + 64, x0:x0 + 64].astype(np.complex128)
        e = np.angle(a * np.conj(b))
        full = float(np.sqrt(np.mean(e * e)))
        say(f"[spot check] tile {tile}: kernel residual {res:.6f}, full-frame WFF2 {full:.6f}, "
            f"diff {abs(res - full):.2e}")
        assert not fb
        assert abs(res - full) <= 1e-5


def test_threshold_computed_on_first_call_and_kept(oracle):
    """One WFT2 handle with thr < 0 called on two fields of different modulus:
    the threshold is 6 sqrt(mean|f1|^2/3) from the first and is kept for the
    second (WFT/WFT2_CPU.cpp:1014-1015); both outputs match the oracle's
    two-call results."""
    W, H = 200, 160
    rng = np.random.default_rng(9)
    ph = 0.5 * np.arange(W)[None, :] + 0.2 * np.arange(H)[:, None] + rng.normal(0, 0.4, (H, W))
    f1 = (np.exp(1j * ph)).astype(np.complex64)
    f2 = (3.0 * np.exp(1j * (ph + 0.3)) * (1 + 0.2 * rng.normal(size=(H, W)))).astype(np.complex64)
    w = glsu.WFT2(W, H, params=glsu.wff_params(10.0, -0.2, 0.2, 0.1))
    o = oracle.WFF2(W, H, 10.0, -0.2, 0.2, 0.1)
    g1 = w(torch.from_numpy(f1).cuda()).cpu().numpy()
    thr1 = w.info()["thr"]
    r1 = o(f1.astype(np.complex128))
    g2 = w(torch.from_numpy(f2).cuda()).cpu().numpy()
    thr2 = w.info()["thr"]
    r2 = o(f2.astype(np.complex128))
    expect1 = 6 * np.sqrt(np.mean(np.abs(f1.astype(np.complex128)) ** 2) / 3)
    expect2 = 6 * np.sqrt(np.mean(np.abs(f2.astype(np.complex128)) ** 2) / 3)
    fresh = glsu.WFT2(W, H, params=glsu.wff_params(10.0, -0.2, 0.2, 0.1))
    g2_fresh = fresh(torch.from_numpy(f2).cuda()).cpu().numpy()
    say(f"[thr] first call {thr1:.6f} (expect {expect1:.6f}), second call {thr2:.6f} "
        f"(a fresh handle: {fresh.info()['thr']:.6f} = {expect2:.6f}); oracle thr {o.thr:.6f}")
    assert abs(thr1 - expect1) <= 1e-5 * expect1
    assert thr2 == thr1 and abs(o.thr - expect1) <= 1e-12 * expect1
    assert abs(fresh.info()["thr"] - expect2) <= 1e-5 * expect2
    for g, r in ((g1, r1), (g2, r2)):
        rel = np.abs(g - r).max() / np.abs(r).max()
        # phase where the filtered field is not vanishingly small (the phase of
        # a near-zero value carries no information)
        big = np.abs(r) > 1e-2 * np.abs(r).max()
        rms, _ = wrapped_err(np.angle(g[big]), np.angle(r[big]))
        say(f"   vs oracle: rel {rel:.2e}, phase rms {rms:.2e} on {big.mean():.3f} of the pixels")
        assert rel < 1e-4 and rms < 1e-4
    # the kept threshold matters: a fresh handle's second-field output differs
    assert np.abs(g2_fresh - g2).max() > 1e-3 * np.abs(g2).max()
