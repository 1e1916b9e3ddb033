#!/usr/bin/env python3
"""Benchmark: frames/s of the plain FP32 DPRA + WFF2 chain on one MI355X.

Workload (the configuration the headline metric is quoted on): synthetic
fringes at 1024x1024, 200 frames, 3x3 DPRA, WFF2 sigma = 10 with wx, wy in
[-0.2, 0.2] step 0.1 (25 pairs), threshold 6*sqrt(mean|z|^2/3), rr = 1.
One step = one pass of the chain over the 200-frame sequence (frames resident
in HBM), replayed from captured HIP graphs.  value = frames/s.

Also reported: `roofline` for the dominant kernel (the column WFF2 kernel),
measured with HIP events on the stream it runs on; `cpu_baseline` = the FP64
oracle (the reference CPU algorithm on this project's CPU FFT) on a bounded
frame sample on the host cores; `parity` = wrapped-phase error of the GPU
output against that FP64 sample, frame by frame.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "g-ls3u_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "frames/s DPRA+WFF2 @1024²; learned vs plain at stated rad RMS vs FP64"
PEAK_FP32_TFLOPS = 157.3
# the same ALUs without packed FP32 instructions (the kernels are compiled without them)
PEAK_FP32_UNPACKED_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


# name -> (size, frames, WFF2 (sigma, lo, hi, step), generator (amp0, amp1, carrier, noise))
WORKLOADS = {
    # the headline configuration: 1024^2, 200 frames, 25 pairs
    "plain1024": (1024, 200, (10.0, -0.2, 0.2, 0.1), (3.0, 0.1, 0.05, 10.0)),
    # heavy WFF2: 2048^2, 50 frames, 441 pairs, noise 20, carrier 0.08
    "heavy2048": (2048, 50, (12.0, -0.5, 0.5, 0.05), (3.0, 0.1, 0.08, 20.0)),
    # 64 sequences, sizes uniform over {256..2048}^2, 30 frames each, batched
    "batch64": (None, 30, (10.0, -0.2, 0.2, 0.1), None),
    # learned WFF2 surrogate: 1024^2, 500 frames, noise std drawn from U[5, 25]
    "learned1024": (1024, 500, (10.0, -0.2, 0.2, 0.1), None),
    # AIA initial phase: 1024^2, 20 randomly phase-shifted uint8 frames, noise 10
    "aia1024": (1024, 20, None, None),
    # WFR2 (ridge mode) with the reference's WFR defaults: sigma 10, [-2, 2] step
    # 0.025 on both axes (161 x 161 pairs), one 1024^2 field per step
    "wfr1024": (1024, 1, (10.0, -2.0, 2.0, 0.025), None),
}
BATCH_SIZES = [256, 512, 768, 1024, 1536, 2048]


def bench_learned(args, glsu, torch, np):
    """Learned filter vs plain WFF2 on one 1024^2 sequence (500 frames, noise
    std drawn from U[5, 25], seeds disjoint from the training pairs').  value =
    frames/s of the learned chain (U-Net in place of WFF2, with the 64x64 WFF2
    spot check and fallback every frame); also the plain chain's frames/s on
    the same frames, and on sampled frames the wrapped-phase RMS of the
    network's filtered phase, and of the unfiltered DPRA phase, against the
    FP64 WFF2 of the same raw field (the CPU oracle)."""
    size, T, wff, _ = WORKLOADS["learned1024"]
    W = H = args.size or size
    T = args.frames or T
    rng = np.random.default_rng(31337)
    amp0, noise = float(rng.uniform(2, 5)), float(rng.uniform(5, 25))
    params = glsu.wff_params(*wff)
    frames = glsu.gen_frames(W, H, T, 424242, amp0=amp0, noise=noise)
    phi0 = glsu.synth.phase(W, H, amp0, 0.05)
    path = glsu.learned.WEIGHTS
    trained_here = not os.path.exists(path)
    if trained_here:  # no committed weights: a short training run (untimed)
        x, u = glsu.learned.make_pairs(8, 25, 1024)
        net = glsu.UNet(4, 32, seed=0)
        glsu.learned.train(net, x, u, crops=8000)
        del x, u
    else:
        net = glsu.UNet.load(path)
    tol = 0.03
    # plain chain, graph mode
    d = glsu.DPRA(phi0, params=params)
    out = torch.empty((T, H, W), dtype=torch.float32, device="cuda")
    d(frames[:10].contiguous(), out=out[:10], graph=True)
    d = glsu.DPRA(phi0, params=params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d(frames, out=out, graph=True)
    torch.cuda.synchronize()
    plain_fps = T / (time.perf_counter() - t0)
    # learned chain: per frame LS -> U-Net -> spot check (tile moves every frame)
    def tile(t):
        return ((t * 389) % (W - 64), (t * 613) % (H - 64))
    dl = glsu.DPRA(phi0, params=params)
    for t in range(3):
        dl.frame_learned(frames[t], net, out[t], tol=tol, tile=tile(t))
        dl.update_ref_phi()
    dl = glsu.DPRA(phi0, params=params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(T):  # asynchronous: the fallback decision stays on the device
        dl.frame_learned(frames[t], net, out[t], tol=tol, tile=tile(t), sync=False)
        dl.update_ref_phi()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    learned_fps = T / dt
    st = dl.learned_stats()
    # accuracy on sampled frames against the FP64 WFF2 of the same raw field
    acc = {}
    if not args.no_cpu:
        from oracle import oracle as O
        wf = O.WFF2(W, H, *wff, nthreads=min(os.cpu_count() or 1, 16))
        dc = glsu.DPRA(phi0, params=params)
        errs_net, errs_raw = [], []
        sample = [0, T // 4, T // 2, T - 1]
        for t in range(T):
            dc.frame_learned(frames[t], net, tol=-1.0)
            dc.update_ref_phi()
            if t in sample:
                z = dc.last_z().cpu().numpy().astype(np.complex128)
                F = wf(z)
                y = dc.last_net().cpu().numpy().astype(np.complex128)
                en = np.angle(y * np.conj(F))
                er = np.angle(z * np.conj(F))
                errs_net.append(float(np.sqrt(np.mean(en * en))))
                errs_raw.append(float(np.sqrt(np.mean(er * er))))
        acc = {"frames": sample, "net_rms_rad": errs_net, "unfiltered_rms_rad": errs_raw,
               "gain_over_unfiltered": round(float(np.mean(errs_raw) / np.mean(errs_net)), 2)}
    line = {
        "metric": METRIC, "value": round(learned_fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (net, FP32 accumulate) + f32",
        "data": "synthetic",
        "config": {"workload": f"learned WFF2 surrogate, {W}x{H}, {T} frames: DPRA LS -> 4-level "
                               "U-Net (32 base ch, BF16 MFMA) -> phase update, WFF2 on a "
                               f"64x64 tile every frame, fallback above {tol} rad RMS; "
                               f"noise std {noise:.2f}, amp0 {amp0:.2f}",
                   "name": "learned1024", "weights": "trained in this run" if trained_here
                   else os.path.relpath(path, ROOT)},
        "plain_frames_per_s": round(plain_fps, 2),
        "learned_over_plain": round(learned_fps / plain_fps, 3),
        "fallback_frames": st["fallbacks"],
        "spot_check_rms_rad": {"mean": round(st["mean_rms"], 5), "max": round(st["max_rms"], 5)},
        "accuracy_vs_fp64_wff2": acc,
    }
    print(json.dumps(line), flush=True)


def aia_frames(np, M, H, W, seed, noise=10.0, amp=3.0, carrier=0.05):
    """M uint8 frames I = 128 + 100 cos(phi + delta_k) + noise, phi = amp *
    peaks + 2 pi carrier x, shifts uniform in (-pi, pi]."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(-3, 3, W), np.linspace(-3, 3, H))
    peaks = (3 * (1 - X) ** 2 * np.exp(-X ** 2 - (Y + 1) ** 2)
             - 10 * (X / 5 - X ** 3 - Y ** 5) * np.exp(-X ** 2 - Y ** 2)
             - np.exp(-(X + 1) ** 2 - Y ** 2) / 3)
    phi = amp * peaks + 2 * np.pi * carrier * np.arange(W)[None, :]
    delta = rng.uniform(-np.pi, np.pi, M)
    f = np.empty((M, H, W), np.uint8)
    for k in range(M):
        v = 128 + 100 * np.cos(phi + delta[k]) + rng.normal(0, noise, (H, W))
        f[k] = np.clip(np.rint(v), 0, 255)
    return f


def bench_aia(args, glsu, torch, np):
    """AIA (the initial-phase estimator before the DPRA chain): one step = one
    full solve with the reference's defaults (20 iterations at most, 1e-4,
    AIA_CUDAF's error metric) on M uint8 frames resident in HBM.  value =
    solves/s.  Roofline: the pass kernel (one launch per iteration reads every
    frame once), timed with HIP events on its stream."""
    size, M, _, _ = WORKLOADS["aia1024"]
    W = H = args.size or size
    M = args.frames or M
    f = aia_frames(np, M, H, W, 4242)
    frames = torch.from_numpy(f).cuda()
    start = np.random.default_rng(7).normal(0, 1, M).astype(np.float32)
    start[0] = 0
    a = glsu.AIA(M, W, H)
    phi = torch.empty((H, W), dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        a(frames, deltas=start, phi=phi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, d, ms, it, err = a(frames, deltas=start, phi=phi)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sps = args.steps / dt
    it_ms = a.time_iter(frames, start, reps=20, stage=0)
    pass_ms = a.time_iter(frames, start, reps=20, stage=1)
    info = a.info()
    nblk = info["blocks"] * info["chunks"]
    # algorithmic bytes of one iteration launch: every frame once, plus the
    # FP64 partials (5 + 2*KC per block) and the weights
    b_alg = M * W * H + 8 * (5 + 2 * info["frames_per_chunk"]) * nblk + 8 * M
    achieved = b_alg / (pass_ms * 1e-3) / 1e9
    line = {
        "metric": "AIA solves/s @1024^2 x 20 frames (uint8)", "value": round(sps, 2),
        "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / max(args.steps, 1), 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"AIA initial phase, {W}x{H}, {M} uint8 frames with random "
                               "shifts, noise 10; max_iter 20, max_err 1e-4, error metric "
                               "max|delta - delta_old| (AIA_CUDAF)",
                   "name": "aia1024", "iterations": it, "final_err": err,
                   "launch": info},
        "gpu_ms_per_solve": round(ms, 4),
        "iteration_ms": round(it_ms, 5),
        "roofline": {"bound": "hbm", "kernel": "aia_pass_kernel", "avg_ms": round(pass_ms, 5),
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "B_alg_per_launch": b_alg,
                     "note": "frames are re-read every iteration and fit the 256 MB "
                             "last-level cache at this size, so the achieved rate is an "
                             "effective one"},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        nthreads = min(os.cpu_count() or 1, 16)
        t1 = time.perf_counter()
        p_ref, d_ref, it_ref, _ = O.aia(f.astype(np.float64), start.astype(np.float64), 20,
                                        1e-4, 0, nthreads=nthreads)
        cdt = time.perf_counter() - t1
        e = np.angle(np.exp(1j * (phi.cpu().numpy().astype(np.float64) - p_ref)))
        line["cpu_baseline"] = {
            "value": round(1.0 / cdt, 4), "unit": "solves/s", "cores": nthreads, "kind": "port",
            "sample": f"one full solve of the same frames, FP64 oracle (reference algorithm, "
                      f"OpenMP over pixels), {it_ref} iterations, {cdt:.2f} s"}
        line["parity"] = {"iterations_gpu": it, "iterations_oracle": it_ref,
                          "phase_rms_rad": float(np.sqrt(np.mean(e * e))),
                          "phase_max_abs_rad": float(np.abs(e).max()),
                          "shift_max_abs_rad": float(np.abs(np.angle(np.exp(1j * (d - d_ref)))).max())}
    print(json.dumps(line), flush=True)


def wfr_field(np, W, H, seed, noise=0.3, amp=3.0, carrier=0.05):
    """A fringe field b exp(j phi) + complex noise, phi = amp peaks + 2 pi carrier x."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(-3, 3, W), np.linspace(-3, 3, H))
    peaks = (3 * (1 - X) ** 2 * np.exp(-X ** 2 - (Y + 1) ** 2)
             - 10 * (X / 5 - X ** 3 - Y ** 5) * np.exp(-X ** 2 - Y ** 2)
             - np.exp(-(X + 1) ** 2 - Y ** 2) / 3)
    ph = amp * peaks + 2 * np.pi * carrier * np.arange(W)[None, :]
    f = (1 + 0.3 * np.cos(X)) * np.exp(1j * ph)
    f = f + noise * (rng.normal(size=(H, W)) + 1j * rng.normal(size=(H, W)))
    return f.astype(np.complex64)


def bench_wfr(args, glsu, torch, np):
    """WFR2 (the ridge mode of WFT2) with the reference's WFR defaults.  One
    step = the eight ridge maps of one field.  value = fields/s.  Roofline:
    the column kernel (ridge over wy), timed with HIP events on its stream."""
    size, _, wfr, _ = WORKLOADS["wfr1024"]
    W = H = args.size or size
    f = torch.from_numpy(wfr_field(np, W, H, 77)).cuda()
    params = glsu.wff_params(*wfr)
    w = glsu.WFT2(W, H, type=glsu.WFR, params=params)
    out = torch.empty((8, H, W), dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        w.wfr(f, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w.wfr(f, out=out)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fps = args.steps / dt
    info = w.info()
    nwx, nwy, Ny = info["nwx"], info["nwy"], info["ny"]
    stage_ms = [w.time_kernel(f, k, reps=3) for k in (1, 2, 3)]
    col_ms = stage_ms[1]
    # per (wx, column): 1 + nwy FFTs of length Ny, and per pair the product
    # with the real window spectrum (2 flop/point) and the magnitude compare (4)
    flops = nwx * W * ((1 + nwy) * 5.0 * Ny * math.log2(Ny) + nwy * Ny * 6.0)
    achieved = flops / (col_ms * 1e-3) / 1e12
    line = {
        "metric": "WFR2 fields/s @1024^2 (161 x 161 pairs)", "value": round(fps, 3),
        "unit": "fields/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / max(args.steps, 1), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"WFR2 ridge maps of one {W}x{H} fringe field, sigma {wfr[0]:g}, "
                               f"wx, wy in [{wfr[1]:g}, {wfr[2]:g}] step {wfr[3]:g} "
                               f"({nwx} x {nwy} pairs; the reference's WFR defaults)",
                   "name": "wfr1024", "fft_len": [info["nx"], Ny]},
        "kernel_ms": {"rows_fwd": round(stage_ms[0], 4), "wfr_cols": round(stage_ms[1], 4),
                      "ridge_curvature": round(stage_ms[2], 4)},
        "roofline": {"bound": "mfma", "kernel": "wfr_cols_kernel", "avg_ms": round(col_ms, 4),
                     "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "frac_unpacked": round(achieved / PEAK_FP32_UNPACKED_TFLOPS, 4),
                     "note": "FFT butterflies on the FP32 vector ALU, priced against the FP32 peak"},
    }
    if not args.no_cpu:
        # bounded sample: the FP64 oracle over the first k wy rows of the grid
        # (k * nwx of the nwx * nwy pairs; the oracle runs one wy row per
        # thread, so k = threads), scaled to whole fields per second; parity on
        # the same sample against the GPU run with the same grid
        from oracle import oracle as O
        nthreads = min(os.cpu_count() or 1, 16)
        k = nthreads
        wyh = wfr[1] + (k - 1) * wfr[3]
        fh = f.cpu().numpy()
        o = O.WFR2(W, H, sigx=wfr[0], wxl=wfr[1], wxh=wfr[2], wxi=wfr[3],
                   wyl=wfr[1], wyh=wyh, wyi=wfr[3], nthreads=nthreads)
        t1 = time.perf_counter()
        ref = o(fh.astype(np.complex128))
        cdt = time.perf_counter() - t1
        ws = glsu.WFT2(W, H, type=glsu.WFR,
                       params=glsu.wff_params(wfr[0], wfr[1], wfr[2], wfr[3], wyl=wfr[1],
                                              wyh=wyh, wyi=wfr[3]))
        zs = {a: b.cpu().numpy().astype(np.float64) for a, b in ws.wfr(f).items()}
        same = (np.abs(zs["wx"] - ref["wx"]) < 1e-6) & (np.abs(zs["wy"] - ref["wy"]) < 1e-6)
        e = np.angle(np.exp(1j * (zs["phase"] - ref["phase"])))[same]
        line["cpu_baseline"] = {
            "value": round(k / nwy / cdt, 6), "unit": "fields/s", "cores": nthreads,
            "kind": "port",
            "sample": f"FP64 oracle on the same field over {k} of the {nwy} wy rows "
                      f"({k * nwx} of {nwx * nwy} pairs, {cdt:.1f} s), scaled by {nwy}/{k}"}
        line["parity"] = {"sample_pairs": k * nwx, "pair_mismatch_frac": float(1 - same.mean()),
                          "r_rel_max": float(np.abs(zs["r"] - ref["r"]).max() / ref["r"].max()),
                          "phase_rms_rad": float(np.sqrt(np.mean(e * e)))}
    print(json.dumps(line), flush=True)


def bench_batch(args, glsu, torch, np):
    """64 independent sequences of uneven sizes in one batched run per step
    (persistent launches, one graph replay per frame).  value = frames/s over
    all sequences; also reported: the same sequences run one after another
    through single-sequence handles."""
    _, T, wff, _ = WORKLOADS["batch64"]
    T = args.frames or T
    n = 64
    rng = np.random.default_rng(20240607)
    sizes = [int(v) for v in rng.choice(BATCH_SIZES, n)]
    amp0 = rng.uniform(2.0, 5.0, n)
    noise = rng.uniform(5.0, 25.0, n)
    params = glsu.wff_params(*wff)
    frames, phis = [], []
    for i, S in enumerate(sizes):
        frames.append(glsu.gen_frames(S, S, T, 1000 + i, amp0=float(amp0[i]), amp1=0.1,
                                      carrier=0.05, noise=float(noise[i])))
        phis.append(glsu.synth.phase(S, S, float(amp0[i]), 0.05))
    tuned = {}
    if not args.no_tune:
        for S in sorted(set(sizes)):
            tuned[S] = glsu.autotune.tuning_only(glsu.autotune.tune(
                S, S, params, force=True, log=lambda m: print(m, file=sys.stderr, flush=True)))
    tunings = [tuned.get(S) for S in sizes] if tuned else None
    b = glsu.DPRABatch(phis, 1, params=params, tunings=tunings)
    outs = [torch.empty((T, S, S), dtype=torch.float32, device="cuda") for S in sizes]
    graph = args.mode == "graph"
    for _ in range(args.warmup):
        b(frames, outs=outs, graph=graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b(frames, outs=outs, graph=graph)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fps = args.steps * n * T / dt
    # the same work through one single-sequence handle per sequence, in turn
    singles = [glsu.DPRA(phis[i], 1, params=params, tuning=tunings[i] if tunings else None)
               for i in range(n)]
    for i in range(n):
        singles[i](frames[i], out=outs[i], graph=graph)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(n):
        singles[i](frames[i], out=outs[i], graph=graph)
    torch.cuda.synchronize()
    seq_fps = n * T / (time.perf_counter() - t1)
    px = sum(S * S for S in sizes)
    np_ref = sum(_opt(S + 60) ** 2 for S in sizes)
    b_alg = 28 * px + 8 * np_ref * 26
    line = {
        "metric": METRIC, "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / max(args.steps, 1), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched FP32 DPRA+WFF2: {n} sequences, sizes uniform over "
                               f"{{256..2048}}^2, {T} frames each, one batched run per step, "
                               "WFF2 sigma=10, [-0.2,0.2] step 0.1 (25 pairs), per-sequence "
                               "amp0 in U[2,5] (+0.1 t) and noise in U[5,25]",
                   "name": "batch64", "sizes": sizes, "plan_classes": b.info()["classes"],
                   "mode": args.mode},
        "sequential_single_handles_frames_per_s": round(seq_fps, 2),
        "batched_over_sequential": round(fps / seq_fps, 3),
        "megapixels_per_s": round(args.steps * T * px / dt / 1e6, 1),
        "roofline": {"bound": "hbm", "chain_B_alg_per_step_frame": b_alg,
                     "achieved": round(b_alg * (fps / n) / 1e9, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(b_alg * (fps / n) / 1e9 / PEAK_HBM_GBS, 4),
                     "traffic": None,
                     "note": "whole-chain algorithmic bytes (28 HW + 8 Np (1+P) per frame, "
                             "reference padding) over all sequences per step; per-kernel "
                             "figures are those of the plain1024 workload"},
    }
    print(json.dumps(line), flush=True)


def col_kernel_flops(W, H, nwx, nwy, Ny):
    """Algorithmic FLOP of one launch of the column kernel: per (wx, column)
    item, 2 + 2*nwy complex FFTs of length Ny (5 N log2 N each) plus per pair
    two products with the real window spectrum (2 flop each), the threshold
    test (3) and the accumulate (2) per point (9 flop/point/pair in total)."""
    per_item = (2 + 2 * nwy) * 5.0 * Ny * math.log2(Ny) + nwy * Ny * 9.0
    return nwx * W * per_item


def col_kernel_bytes(W, H, nwx):
    """Algorithmic HBM bytes of one launch: read the nwx row-filtered planes
    and write the nwx column-filtered planes (complex64, H x W each)."""
    return 2 * nwx * W * H * 8


def _opt(n):
    """Smallest even 2^a 3^b 5^c 7^d >= n (the reference's padded FFT size)."""
    m = max(n, 2)
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1 and m % 2 == 0:
            return m
        m += 1


def _traffic(W, H, T, tuning):
    """HBM bytes per launch of the column kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), when it holds an entry for this size and tuning."""
    path = os.path.join(ROOT, "profiles", f"wff_cols_{W}x{H}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    for e in t.get("entries", []):
        if e.get("tuning") == tuning and e.get("size") == [H, W]:
            return {"bytes": int(e["hbm_bytes_per_launch"]), "source": os.path.relpath(path, ROOT)}
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="plain1024")
    ap.add_argument("--size", type=int, default=None, help="override the workload's size")
    ap.add_argument("--frames", type=int, default=None, help="override the workload's frames")
    ap.add_argument("--mode", choices=["graph", "eager"], default="graph")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget for the FP64 CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="library default plans instead of the autotuned ones")
    ap.add_argument("--tuning", default=None,
                    help="fixed tuning as planx,plany,cols_waves,rows_waves (skips the tuner)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run one plain copy kernel (torch.neg) over this many MiB of float32 "
                         "before the timed region, as the HBM counter calibration for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: runs on exactly one GPU (the per-frame chain is sequential; "
              "sharding across GPUs is not built).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import glsu

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "batch64":
        return bench_batch(args, glsu, torch, np)
    if args.workload == "learned1024":
        return bench_learned(args, glsu, torch, np)
    if args.workload == "aia1024":
        return bench_aia(args, glsu, torch, np)
    if args.workload == "wfr1024":
        return bench_wfr(args, glsu, torch, np)
    size, T, wff, gen = WORKLOADS[args.workload]
    W = H = args.size or size
    T = args.frames or T
    seed = 20240607
    amp0, amp1, carrier, noise = gen
    params = glsu.wff_params(*wff)

    frames = glsu.gen_frames(W, H, T, seed, amp0, amp1, carrier, noise)
    phi0 = glsu.synth.phase(W, H, amp0, carrier)
    out = torch.empty((T, H, W), dtype=torch.float32, device="cuda")
    tuning = None
    if args.tuning:
        px, py, cw, rw = (int(v) for v in args.tuning.split(","))
        tuning = dict(planx=px, plany=py, cols_waves=cw, rows_waves=rw)
    elif not args.no_tune:  # untimed: picks FFT plans and launch shapes for this size
        tuning = glsu.autotune.tuning_only(glsu.autotune.tune(
            W, H, params, force=True, log=lambda m: print(m, file=sys.stderr, flush=True)))
    d = glsu.DPRA(phi0, params=params, tuning=tuning)
    graph = args.mode == "graph"
    if args.calib_copy_mb:
        src = torch.ones(args.calib_copy_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        d(frames, out=out, graph=graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d(frames, out=out, graph=graph)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / max(args.steps, 1)
    fps = args.steps * T / dt

    # --- roofline of the dominant kernel (the column WFF2 kernel) ----------------
    # HIP events on the handle's stream around back-to-back launches of that
    # kernel on the benchmark's own buffers (after the timed region).
    kms = d.profile(frames[:8])
    col_ms = d.time_kernel(2, reps=20)
    info = glsu.WFT2(W, H, params=params, tuning=tuning).info()
    nwx, nwy, Ny = info["nwx"], info["nwy"], info["ny"]
    flops = col_kernel_flops(W, H, nwx, nwy, Ny)
    achieved = flops / (col_ms * 1e-3) / 1e12
    hbm = col_kernel_bytes(W, H, nwx) / (col_ms * 1e-3) / 1e9
    # whole-chain HBM figure: B_alg = 28 HW + 8 Np (1 + P) bytes per frame, with
    # Np the padded size of the reference's rule (even 2^a 3^b 5^c 7^d >= size + 2s)
    np_ref = _opt(H + 2 * info["sy"]) * _opt(W + 2 * info["sx"])
    b_alg = 28 * W * H + 8 * np_ref * (1 + nwx * nwy)
    roofline = {
        "bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
        "kernel": "wff_cols_kernel", "avg_ms": round(col_ms, 5),
        "note": "FFT butterflies on the FP32 vector ALU (no dense contraction, so no MFMA); "
                "priced against the FP32 peak, which MI355X's matrix and vector units share; "
                "that peak assumes packed FP32, which the kernels are compiled without (frac_unpacked "
                "is against the unpacked vector rate)",
        "frac_unpacked": round(achieved / PEAK_FP32_UNPACKED_TFLOPS, 4),
        "hbm_GBps": round(hbm, 1), "hbm_frac": round(hbm / PEAK_HBM_GBS, 4),
        "chain_hbm": {"B_alg_per_frame": b_alg, "GBps": round(b_alg * fps / 1e9, 1),
                      "frac": round(b_alg * fps / 1e9 / PEAK_HBM_GBS, 4)},
        "kernel_ms_per_frame": {k: round(v / 8, 5) for k, v in
                                zip(["dpra_ls", "rows_fwd", "cols", "rows_inv_update"], kms)},
    }

    traffic = _traffic(W, H, T, d.tuning())
    if traffic is not None:
        roofline["traffic"] = traffic["bytes"]
        roofline["traffic_over_algorithmic"] = round(traffic["bytes"] / col_kernel_bytes(W, H, nwx), 3)
        roofline["traffic_source"] = traffic["source"]

    line = {
        "metric": METRIC, "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain FP32 DPRA+WFF2 chain, {W}x{H}, {T} frames per step, "
                               f"3x3 DPRA, WFF2 sigma={wff[0]:g}, wx,wy in [{wff[1]:g},{wff[2]:g}] "
                               f"step {wff[3]:g} ({nwx * nwy} pairs), thr=6*sqrt(mean|z|^2/3), "
                               f"rr=1, noise {noise:g}, carrier {carrier:g}",
                   "name": args.workload,
                   "height": H, "width": W, "frames_per_step": T, "pairs": nwx * nwy,
                   "fft_len": [info["nx"], info["ny"]], "mode": args.mode,
                   "tuning": d.tuning()},
        "roofline": roofline,
    }

    # --- CPU baseline (FP64 oracle) and parity on the same frames ---------------
    if not args.no_cpu:
        from oracle import oracle as O
        nthreads = min(os.cpu_count() or 1, 16)
        host = frames.cpu().numpy().astype(np.float64)
        od = O.DPRA(phi0, sigx=wff[0], wxl=wff[1], wxh=wff[2], wxi=wff[3], nthreads=nthreads)
        cpu_out = []
        t0 = time.perf_counter()
        n = 0
        while n < T:
            dphi, _ = od.dpra_per_frame(host[n])
            od.update_ref_phi()
            cpu_out.append(dphi)
            n += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        cdt = time.perf_counter() - t0
        gpu = glsu.DPRA(phi0, params=params, tuning=tuning)(frames[:n].contiguous()).cpu().numpy()
        errs = []
        for k in range(n):
            e = np.angle(np.exp(1j * (gpu[k].astype(np.float64) - cpu_out[k])))
            errs.append((float(np.sqrt(np.mean(e * e))), float(np.abs(e).max())))
        line["cpu_baseline"] = {
            "value": round(n / cdt, 4), "unit": "frames/s", "cores": nthreads, "kind": "port",
            "sample": f"first {n} frames of the same {W}x{H} sequence, FP64 oracle "
                      f"(reference algorithm, OpenMP over wy: effective parallelism "
                      f"min({nthreads}, {nwy}) in WFF2), {cdt:.1f} s",
        }
        line["parity"] = {
            "frames": n, "rms_rad_max_over_frames": max(e[0] for e in errs),
            "max_abs_rad": max(e[1] for e in errs),
            "per_frame_rms": [float("%.3e" % e[0]) for e in errs],
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
