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


def host_info():
    """Host cores usable by this process, the thread count the CPU legs use,
    and the CPU model string."""
    try:
        nproc = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        nproc = os.cpu_count() or 1
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "16"))
    except ValueError:
        share = 16
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "os_cpu_count": os.cpu_count(), "threads": max(1, min(nproc, share)),
            "cpu_model": model}


LEARNED_TOL = 0.03


def _learned_tile(W, H, t):
    """Spot-check tile origin (x, y) of frame t: moves every frame."""
    return ((t * 389) % (W - 64), (t * 613) % (H - 64))


def learned_leg(args, glsu, torch, np, W=1024, T=500, with_accuracy=True, nthreads=1):
    """The learned WFF2 surrogate on one 1024^2 sequence (500 frames, noise
    std drawn from U[5, 25], seeds disjoint from the training pairs') with the
    committed weights: frames/s of the learned chain (DPRA LS -> U-Net ->
    phase update, with the 64x64 WFF2 spot check and fallback every frame) and
    of the plain chain on the same frames.  Accuracy of the timed chain: the
    chain is re-run (bitwise equal to the timed run: checked), and on sampled
    frames its per-frame filtered phase -- the increment of its cumulative
    output -- is compared with the FP64 WFF2 of the same raw field (the CPU
    oracle), as is the unfiltered raw phase.  Returns a dict, or None when
    the weights are missing."""
    _, _, wff, _ = WORKLOADS["learned1024"]
    H = W
    path = glsu.learned.WEIGHTS
    if not os.path.exists(path):
        return None
    rng = np.random.default_rng(31337)
    amp0, noise = float(rng.uniform(2, 5)), float(rng.uniform(5, 25))
    params = glsu.wff_params(*wff)
    frames = glsu.gen_frames(W, H, T, 424242, amp0=amp0, noise=noise)
    phi0 = glsu.synth.phase(W, H, amp0, 0.05)
    net = glsu.UNet.load(path)
    tol = LEARNED_TOL
    out = torch.empty((T, H, W), dtype=torch.float32, device="cuda")
    # plain chain on the same frames, graph mode (the first call captures)
    d = glsu.DPRA(phi0, params=params)
    d(frames[:10].contiguous(), out=out[:10], graph=True)
    d = glsu.DPRA(phi0, params=params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d(frames, out=out, graph=True)
    torch.cuda.synchronize()
    plain_fps = T / (time.perf_counter() - t0)
    del d
    # learned chain (warm-up on a separate handle, then the timed one)
    dl = glsu.DPRA(phi0, params=params)
    for t in range(3):
        dl.frame_learned(frames[t], net, out[t], tol=tol, tile=_learned_tile(W, H, t))
        dl.update_ref_phi()
    dl = glsu.DPRA(phi0, params=params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(T):  # asynchronous: the fallback decision stays on the device
        dl.frame_learned(frames[t], net, out[t], tol=tol, tile=_learned_tile(W, H, t), sync=False)
        dl.update_ref_phi()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = dl.learned_stats()
    del dl
    res = {
        "workload": f"learned WFF2 surrogate, {W}x{H}, {T} frames: DPRA LS -> 4-level U-Net "
                    "(32 base ch, BF16 MFMA, FP32 accumulate) -> phase update, WFF2 on a "
                    f"64x64 tile every frame, fallback above {tol} rad RMS; noise std "
                    f"{noise:.2f}, amp0 {amp0:.2f}",
        "weights": os.path.relpath(path, ROOT),
        "frames_per_s": round(T / dt, 2), "ms_per_frame": round(dt * 1e3 / T, 4),
        "plain_frames_per_s": round(plain_fps, 2),
        "learned_over_plain": round(plain_fps and (T / dt) / plain_fps, 3),
        "fallback_frames": st["fallbacks"], "frames": T,
        "spot_check_rms_rad": {"mean": round(st["mean_rms"], 5), "max": round(st["max_rms"], 5)},
    }
    if not with_accuracy:
        return res
    from oracle import oracle as O
    out_t = out.cpu().numpy()
    sample = sorted(set([0, 1, 2] + list(range(25, T, 50)) + [T - 1]))
    dc = glsu.DPRA(phi0, params=params)
    out2 = torch.empty_like(out)
    rows = []
    for t in range(T):
        r = dc.frame_learned(frames[t], net, out2[t], tol=tol, tile=_learned_tile(W, H, t),
                             sync=t in sample)
        if t in sample:
            z = dc.last_z().cpu().numpy().astype(np.complex128)
            y = dc.last_net().cpu().numpy().astype(np.complex128)
            F = O.WFF2(W, H, *wff, nthreads=nthreads, pair_parallel=True)(z)
            aF = np.angle(F)
            o2 = out2[t].cpu().numpy().astype(np.float64)
            inc = o2 if t == 0 else o2 - out2[t - 1].cpu().numpy().astype(np.float64)
            rows.append({"frame": t, "fallback": bool(r[1]),
                         "chain_rms_rad": round(O.wrapped_err(inc, aF)[0], 5),
                         "net_rms_rad": round(O.wrapped_err(np.angle(y), aF)[0], 5),
                         "unfiltered_rms_rad": round(O.wrapped_err(np.angle(z), aF)[0], 5)})
        dc.update_ref_phi()
    same = bool(np.array_equal(out2.cpu().numpy(), out_t))
    ch = [r["chain_rms_rad"] for r in rows]
    raw = [r["unfiltered_rms_rad"] for r in rows]
    res["accuracy_vs_fp64_wff2"] = {
        "note": "per sampled frame: chain = the timed chain's own filtered phase (increment of "
                "its output; WFF2's where the frame fell back), net = the network's output on "
                "the same raw field, unfiltered = the raw DPRA phase; all against the FP64 "
                "WFF2 of that raw field",
        "rerun_bitwise_equal_to_timed": same,
        "chain_rms_rad_mean": round(float(np.mean(ch)), 5),
        "chain_rms_rad_max": round(float(np.max(ch)), 5),
        "gain_over_unfiltered": round(float(np.mean(raw) / max(np.mean(ch), 1e-30)), 2),
        "frames": rows,
    }
    return res


def bench_learned(args, glsu, torch, np):
    """Learned filter vs plain WFF2 on one 1024^2 sequence: its own JSON line
    (value = frames/s of the learned chain)."""
    size, T, _, _ = WORKLOADS["learned1024"]
    W = args.size or size
    T = args.frames or T
    hi = host_info()
    res = learned_leg(args, glsu, torch, np, W=W, T=T, with_accuracy=not args.no_cpu,
                      nthreads=hi["threads"])
    if res is None:
        print("bench.py: no committed U-Net weights (train them with python -m glsu.learned)",
              file=sys.stderr)
        sys.exit(1)
    line = {
        "metric": METRIC, "value": res["frames_per_s"], "unit": "frames/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(res["ms_per_frame"] * T, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (net, FP32 accumulate) + f32", "data": "synthetic",
        "config": {"workload": res.pop("workload"), "name": "learned1024",
                   "weights": res.pop("weights")},
    }
    line.update(res)
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
                               "amp0 in U[2,5] (+0.1 t) and noise in U[5,25]; each step "
                               "continues the sequences from the state the previous step left "
                               "(repeated operator() calls)",
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
    if not args.no_cpu:
        # a documented per-sequence subset: the first sequence of each size up to
        # 1024^2, first 2 frames, FP64 oracle with the reference's loop (CPU
        # baseline) against the batched outputs of the same frames (parity)
        from oracle import oracle as O
        hi = host_info()
        sub = [sizes.index(S) for S in BATCH_SIZES if S <= 1024 and S in sizes]
        k = min(2, T)
        # a new batch of those sequences from their phi0 (the timed batch's state
        # has moved on, and its outputs were overwritten by the single-handle runs)
        bsub = glsu.DPRABatch([phis[i] for i in sub], 1, params=params,
                              tunings=[tunings[i] for i in sub] if tunings else None)
        osub = bsub([frames[i][:k].contiguous() for i in sub], graph=graph)
        t0 = time.perf_counter()
        px_done = 0
        rows = []
        for j, i in enumerate(sub):
            S = sizes[i]
            host = frames[i][:k].cpu().numpy().astype(np.float64)
            od = O.DPRA(phis[i], sigx=wff[0], wxl=wff[1], wxh=wff[2], wxi=wff[3],
                        nthreads=hi["threads"])
            ref = od.run(host)
            px_done += k * S * S
            got = osub[j].cpu().numpy()
            for t in range(k):
                r, m = O.wrapped_err(got[t], ref[t])
                rows.append({"seq": i, "size": S, "frame": t, "rms_rad": r, "max_rad": m})
        cdt = time.perf_counter() - t0
        px_all = sum(S * S for S in sizes) * T
        line["cpu_baseline"] = {
            "value": round(px_done / cdt / px_all * n * T, 4), "unit": "frames/s",
            "cores": hi["threads"], "kind": "port",
            "sample": f"first {k} frames of sequences {sub} (sizes {[sizes[i] for i in sub]}), "
                      f"FP64 oracle with the reference's loop (OpenMP over wy), {cdt:.1f} s; "
                      "scaled to the whole workload by pixel count",
            "host": hi}
        line["parity"] = {"frames": rows, "rms_rad_max": max(r["rms_rad"] for r in rows),
                          "note": "batched output of the first frames of a sequence of each "
                                  "size up to 1024^2 against the FP64 oracle chain"}
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


def _col_entry(W, H, tuning):
    """The committed rocprofv3 summary entry for the column kernel at this size
    and with this kernel's own launch choices (plany, cols_waves; the other
    tuning fields pick other kernels), or None."""
    path = os.path.join(ROOT, "profiles", f"wff_cols_{W}x{H}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, path
    for e in t.get("entries", []):
        et = e.get("tuning", {})
        if e.get("size") == [H, W] and all(et.get(k) == tuning.get(k)
                                           for k in ("plany", "cols_waves")):
            return e, path
    return None, path


def _traffic(W, H, T, tuning):
    """HBM bytes per launch of the column kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), when it holds an entry for this size and tuning."""
    e, path = _col_entry(W, H, tuning)
    if e is None:
        return None
    return {"bytes": int(e["hbm_bytes_per_launch"]), "source": os.path.relpath(path, ROOT)}


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
    ap.add_argument("--parity-frames", type=int, default=None,
                    help="frames of the sequence checked against the FP64 oracle (default: all)")
    ap.add_argument("--parity-floor", action="store_true",
                    help="also run the oracle chain from phi0 rounded to float32")
    ap.add_argument("--no-learned", action="store_true",
                    help="leave out the learned-surrogate object of the default line")
    ap.add_argument("--no-tune", action="store_true",
                    help="library default plans instead of the autotuned ones")
    ap.add_argument("--tuning", default=None,
                    help="fixed tuning as planx,plany,cols_waves,rows_waves[,ls_tile] (skips the tuner)")
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
        vals = [int(v) for v in args.tuning.split(",")]
        tuning = dict(zip(("planx", "plany", "cols_waves", "rows_waves", "ls_tile"), vals))
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

    # Every step is the whole sequence from its start: before each step (outside
    # the timed interval) the handle is reset to phi_ref = phi0 and a zero
    # cumulative phase, as a freshly constructed object holds them
    # (DPRA/dpra_cpu.cpp:19, :39-43); the WFF threshold stays cached from the
    # first frame, as the reference caches it (WFT/WFT2_CPU.cpp:1014-1015).
    phr0, dpr0 = d.get_state()

    def reset():
        d.set_state(phr0, dpr0)
        torch.cuda.synchronize()

    # set-up, untimed: capture the chain's graphs (the first call also computes
    # the threshold, so the second call captures the steady-state graph)
    for _ in range(2):
        reset()
        d(frames, out=out, graph=graph)
    for _ in range(args.warmup):
        reset()
        d(frames, out=out, graph=graph)
    step_s, step_ms = [], []
    for _ in range(args.steps):
        reset()
        t0 = time.perf_counter()
        d(frames, out=out, graph=graph, timed=True)
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - t0)
        step_ms.append(d.last_ms)
    dt = sum(step_s)
    ms_per_step = dt * 1e3 / max(args.steps, 1)
    fps = args.steps * T / dt
    gpu_out = out.cpu().numpy() if not args.no_cpu else None  # the last timed step's output

    # --- roofline of the dominant kernel (the column WFF2 kernel) ----------------
    # One more pass of the same sequence from phi0 (after the timed region) with
    # HIP events on the handle's stream around each kernel of every frame (these
    # intervals also hold the dispatch gap before each kernel) and the count of
    # skipped column transforms; then 20 back-to-back launches of the column
    # kernel on the state the pass ends in, between two events: its average
    # launch duration.
    reset()
    kms, skipped, transforms = d.profile_skips(frames)
    col_chain_ms = kms[2] / T
    col_ms = d.time_kernel(2, reps=20)
    info = glsu.WFT2(W, H, params=params, tuning=tuning).info()
    nwx, nwy, Ny = info["nwx"], info["nwy"], info["ny"]
    flops = col_kernel_flops(W, H, nwx, nwy, Ny)
    achieved = flops / (col_ms * 1e-3) / 1e12
    # FLOP the launches executed: a skipped transform removes one forward FFT
    # and the accumulate (5 N log2 N + 2 N per skipped (wx, column, wy))
    skip_flops = skipped / T * (5.0 * Ny * math.log2(Ny) + 2.0 * Ny)
    hbm = col_kernel_bytes(W, H, nwx) / (col_ms * 1e-3) / 1e9
    # whole-chain HBM figure: B_alg = 28 HW + 8 Np (1 + P) bytes per frame, with
    # Np the padded size of the reference's rule (even 2^a 3^b 5^c 7^d >= size + 2s)
    np_ref = _opt(H + 2 * info["sy"]) * _opt(W + 2 * info["sx"])
    b_alg = 28 * W * H + 8 * np_ref * (1 + nwx * nwy)
    roofline = {
        "bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
        "kernel": "wff_cols_kernel", "avg_ms": round(col_ms, 5),
        "avg_ms_source": "HIP events on the handle's stream around 20 back-to-back launches "
                         f"on the state after one pass of the {T}-frame sequence",
        "per_frame_in_chain_ms": round(col_chain_ms, 5),
        "per_frame_in_chain_note": "events around the column kernel of every frame of that "
                                   "pass, each interval including the dispatch gap before it",
        "skipped_transforms": {"skipped": skipped, "of": transforms,
                               "frac": round(skipped / max(transforms, 1), 4),
                               "executed_TFLOPs_per_frame_in_chain": round((flops - skip_flops) / (col_chain_ms * 1e-3)
                                                         / 1e12, 2),
                               "note": "forward transforms of a (wx, column, wy) whose thresholded "
                                       "column is zero on every image row add exactly nothing and "
                                       "are skipped; achieved counts them as done"},
        "note": "FFT butterflies on the FP32 vector ALU (no dense contraction, so no MFMA); "
                "priced against the FP32 peak, which assumes packed FP32, which the kernels "
                "are compiled without (frac_unpacked is against the unpacked vector rate)",
        "frac_unpacked": round(achieved / PEAK_FP32_UNPACKED_TFLOPS, 4),
        "hbm_GBps": round(hbm, 1), "hbm_frac": round(hbm / PEAK_HBM_GBS, 4),
        "chain_hbm": {"B_alg_per_frame": b_alg, "GBps": round(b_alg * fps / 1e9, 1),
                      "frac": round(b_alg * fps / 1e9 / PEAK_HBM_GBS, 4)},
        "kernel_ms_per_frame": {k: round(v / T, 5) for k, v in
                                zip(["dpra_ls", "rows_fwd", "cols", "rows_inv_update"], kms)},
    }
    prof = _profiled(W, H, d.tuning())
    if prof is not None:
        a_p = flops / (prof["avg_ms"] * 1e-3) / 1e12
        roofline["rocprof"] = {"avg_ms": prof["avg_ms"], "achieved": round(a_p, 2),
                               "frac": round(a_p / PEAK_FP32_TFLOPS, 4), "source": prof["source"]}
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
                               f"rr=1, noise {noise:g}, carrier {carrier:g}; every step from "
                               "phi0 and a zero cumulative phase",
                   "name": args.workload,
                   "height": H, "width": W, "frames_per_step": T, "pairs": nwx * nwy,
                   "fft_len": [info["nx"], info["ny"]], "mode": args.mode,
                   "tuning": d.tuning()},
        "step_ms": {"wall": [round(v * 1e3, 3) for v in step_s],
                    "hip_events": [round(v, 3) for v in step_ms]},
        "roofline": roofline,
    }
    if args.workload == "heavy2048":
        line["config"]["precision"] = (
            "FP32 bulk arithmetic; the three sign-changing normal-equation sums accumulated "
            "in FP64 (exact for the nine FP32 terms, in place of a compensated FP32 two-sum); "
            "FP64 3x3 solve where |det A| < 1e-6 * A00*A11*A22")
    del d

    # --- CPU baseline (FP64 oracle, the reference's loop structure) -------------
    if not args.no_cpu:
        from oracle import oracle as O
        hi = host_info()
        nthreads = hi["threads"]
        host = frames.cpu().numpy()
        od = O.DPRA(phi0, sigx=wff[0], wxl=wff[1], wxh=wff[2], wxi=wff[3], nthreads=nthreads)
        t0 = time.perf_counter()
        n = 0
        while n < T:
            od.dpra_per_frame(host[n].astype(np.float64))
            od.update_ref_phi()
            n += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        cdt = time.perf_counter() - t0
        del od
        line["cpu_baseline"] = {
            "value": round(n / cdt, 4), "unit": "frames/s", "cores": nthreads, "kind": "port",
            "sample": f"first {n} frames of the same {W}x{H} sequence, FP64 oracle (the "
                      f"reference algorithm on this project's CPU FFT; OpenMP over wy as "
                      f"WFT/WFT2_CPU.cpp:227-237, effective WFF2 parallelism "
                      f"min({nthreads}, {nwy}) = {min(nthreads, nwy)}), {cdt:.1f} s",
            "host": hi,
        }
        # --- parity: the whole sequence against the FP64 oracle, frame by frame --
        # the FP64 oracle needs ~0.5 s per 1024^2 frame at 25 pairs on 16 threads, and
        # about a minute per 2048^2 frame at 441 pairs: the heavy workload checks a
        # leading prefix (valid: the chain is sequential)
        default_nf = 2 if args.workload == "heavy2048" else T
        nf = min(default_nf if args.parity_frames is None else args.parity_frames, T)
        if nf > 0:
            line["parity"] = chain_parity(np, O, phi0, host, gpu_out, nf, wff, nthreads,
                                          floor=args.parity_floor)
            line["parity"]["one_step"] = one_step_parity(
                np, O, glsu, torch, phi0, frames, host, nf, params, tuning, wff, nthreads)
    if args.workload == "plain1024" and not args.no_learned:
        line["learned"] = learned_leg(args, glsu, torch, np, with_accuracy=not args.no_cpu,
                                      nthreads=host_info()["threads"])
    print(json.dumps(line), flush=True)


def chain_parity(np, O, phi0, host, gpu_out, nf, wff, nthreads, floor=False):
    """Frames 0..nf-1 of the GPU chain against the FP64 oracle chain from the
    same phi0 (DPRA_CPU's loop, DPRA/dpra_cpu.cpp:58-89, rr = 1), frame by
    frame: wrapped-phase RMS and max |error| in radians.  The oracle's WFF2
    threads take (wy, wx) pairs so the 200-frame chain fits the run; with
    floor=True the oracle chain is also run from phi0 rounded to float32 (the
    GPU's storage precision) and its own distance is reported per frame."""
    t0 = time.perf_counter()
    kw = dict(sigx=wff[0], wxl=wff[1], wxh=wff[2], wxi=wff[3], nthreads=nthreads,
              pair_parallel=True)
    od = O.DPRA(phi0, **kw)
    o32 = O.DPRA(phi0.astype(np.float32).astype(np.float64), **kw) if floor else None
    rms, mx, fl = [], [], []
    for t in range(nf):
        f = host[t].astype(np.float64)
        ref, _ = od.dpra_per_frame(f)
        od.update_ref_phi()
        r, m = O.wrapped_err(gpu_out[t], ref)
        rms.append(r)
        mx.append(m)
        if o32 is not None:
            r32, _ = o32.dpra_per_frame(f)
            o32.update_ref_phi()
            fl.append(O.wrapped_err(r32, ref)[0])
        if t % 20 == 0:
            print(f"parity frame {t}: rms {r:.3e} max {m:.3e}", file=sys.stderr, flush=True)
    res = {
        "frames": nf, "rms_rad_max_over_frames": max(rms), "max_abs_rad": max(mx),
        "frames_over_1e-3_rms": int(sum(r > 1e-3 for r in rms)),
        "per_frame_rms": [float("%.3e" % v) for v in rms],
        "per_frame_max": [float("%.3e" % v) for v in mx],
        "oracle": "FP64 chain from the exact phi0, WFF2 threads over (wy, wx) pairs",
        "note": "chained: frame t of the GPU chain against frame t of the FP64 chain; the "
                "recursion carries every earlier frame's difference forward and amplifies "
                "it where the 3x3 systems are near-singular, so this grows with t for any "
                "two implementations that are not bitwise equal; one_step holds each "
                "frame's own arithmetic against FP64",
        "cpu_s": round(time.perf_counter() - t0, 1),
    }
    if fl:
        res["oracle_floor_rms"] = [float("%.3e" % v) for v in fl]
    return res


def one_step_parity(np, O, glsu, torch, phi0, frames, host, nf, params, tuning, wff, nthreads,
                    n=64):
    """Frames 0..nf-1 one step at a time: the GPU chain is run frame by frame
    (same kernels and tuning as the timed run), its reference phase phi_r(t)
    is read back before frame t, and the FP64 oracle computes frame t's phase
    increment from that same phi_r(t) on the windows of two 64x64 tiles (one
    corner, rotating, and one drawn per frame).  Every output pixel of a step
    depends only on its input within +-(2s + 1), so the window reproduces the
    full-image step on the tile.  This is each frame's own arithmetic against
    FP64, without the recursion carrying earlier differences forward."""
    t0 = time.perf_counter()
    H, W = host.shape[1:]
    sx = int(round(3 * wff[0]))
    m = 2 * sx + 1
    d = glsu.DPRA(phi0, params=params, tuning=tuning)
    pc = torch.empty((H, W), dtype=torch.float32, device="cuda")
    kw = dict(sigx=wff[0], wxl=wff[1], wxh=wff[2], wxi=wff[3], nthreads=nthreads,
              pair_parallel=True)
    orcs = {}
    corners = [(0, 0), (W - n, 0), (0, H - n), (W - n, H - n)]
    rms, mx = [], []
    for t in range(nf):
        phr, _ = d.get_state(dphiref=False)
        d.dpra_per_frame(frames[t], phicurr=pc)
        d.update_ref_phi()
        inc = pc.cpu().numpy().astype(np.float64) - phr.astype(np.float64)
        rng = np.random.default_rng(7000 + t)
        tiles = [corners[t % 4], (int(rng.integers(0, W - n + 1)), int(rng.integers(0, H - n + 1)))]
        errs = []
        for (x0, y0) in tiles:
            wx0, wy0 = max(x0 - m, 0), max(y0 - m, 0)
            wx1, wy1 = min(x0 + n + m, W), min(y0 + n + m, H)
            key = (wy1 - wy0, wx1 - wx0)
            if key not in orcs:
                orcs[key] = O.DPRA(np.zeros(key), **kw)
            od = orcs[key]
            od.set_state(phr[wy0:wy1, wx0:wx1], np.zeros(key))
            ref, _ = od.dpra_per_frame(host[t, wy0:wy1, wx0:wx1].astype(np.float64))
            e = inc[y0:y0 + n, x0:x0 + n] - ref[y0 - wy0:y0 - wy0 + n, x0 - wx0:x0 - wx0 + n]
            errs.append(np.angle(np.exp(1j * e)).ravel())
        e = np.concatenate(errs)
        rms.append(float(np.sqrt(np.mean(e * e))))
        mx.append(float(np.abs(e).max()))
        if t % 40 == 0:
            print(f"one-step parity frame {t}: rms {rms[-1]:.3e} max {mx[-1]:.3e}",
                  file=sys.stderr, flush=True)
    return {"frames": nf, "tiles_per_frame": 2, "tile": n,
            "rms_rad_max_over_frames": max(rms), "rms_rad_mean": float(np.mean(rms)),
            "max_abs_rad": max(mx), "frames_over_1e-3_rms": int(sum(r > 1e-3 for r in rms)),
            "per_frame_rms": [float("%.3e" % v) for v in rms],
            "cpu_s": round(time.perf_counter() - t0, 1)}


def _profiled(W, H, tuning):
    """Average column-kernel duration from the committed rocprofv3 kernel
    trace summary, when it holds an entry for this size and tuning."""
    e, path = _col_entry(W, H, tuning)
    if e is None or "trace_avg_ms" not in e:
        return None
    return {"avg_ms": float(e["trace_avg_ms"]), "source": os.path.relpath(path, ROOT)}


if __name__ == "__main__":
    main()
