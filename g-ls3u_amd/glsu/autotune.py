"""Autotuner for the WFF2 plan and launch shapes.

The reference picks its cuFFT sizes with the OPT table and has no launch
tuning (WFT/WFT2_CUDAf.cu:61-120).  Here several FFT plans cover a padded
length (any N >= size + half window is exact) and the row / column kernels
can run with different wavefront counts, so the fastest combination is picked
by timing on the device.  Every candidate computes the same result up to FP32
rounding; the choice only affects speed.

The search is staged because the kernels are independent:
  1. columns: (plany, cols_waves), timed on the column kernel alone;
  2. rows:    (planx, rows_waves), timed on the forward + inverse row kernels;
  3. DPRA least squares: the pixel tile shape (ls_tile), timed on that kernel
     (the reference fixes its DPRA blocks, DPRA/dpra_cudaf.cu:121-122; the
     shape does not change the result).
Timings come from glsu_dpra_profile (HIP events between kernels) over a few
synthetic frames of the requested size, taken at three points of a synthetic
sequence with the chain's own state there (the times are summed; "ms" in the
result is that sum).  Results are cached per device, size
and WFF parameters in a JSON file.
"""
import json
import os

import numpy as np
import torch

from ._lib import LS_TILES, TUNING_FIELDS, fft_plans
from .dpra import DPRA, gen_frames
from .params import wff_params
from . import synth

CACHE_ENV = "GLSU_TUNE_CACHE"
_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "glsu", "autotune.json")


def _c_round(v):
    return int(np.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def candidates(need, t64_only, limit=3):
    """FFT plan ids with N >= need, shortest first (at most `limit` distinct N)."""
    c = sorted((p["N"], p["id"]) for p in fft_plans()
               if p["N"] >= need and (p["T"] == 64 or not t64_only))
    out, lengths = [], []
    for n, i in c:
        if n not in lengths:
            if len(lengths) == limit:
                break
            lengths.append(n)
        out.append(i)
    return out


def key(width, height, params, device=None):
    dev = torch.cuda.get_device_name(device) if torch.cuda.is_available() else "none"
    p = params
    return (f"{dev}|{width}x{height}|{p.sigma_x:.6g},{p.wxl:.9g},{p.wxh:.9g},{p.wxi:.9g}|"
            f"{p.sigma_y:.6g},{p.wyl:.9g},{p.wyh:.9g},{p.wyi:.9g}")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, table):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time_one(phi0, params, frames, tuning, reps):
    d = DPRA(phi0, 1, params=params, tuning=tuning)
    d.profile(frames[:1])  # warm-up: threshold, first launches
    best = None
    for _ in range(reps):
        ms = d.profile(frames)
        best = ms if best is None else [min(a, b) for a, b in zip(best, ms)]
    used = d.tuning()
    d.close()
    return used, [m / frames.shape[0] for m in best]


def _time(samples, params, tuning, reps):
    """Per-frame kernel times of `tuning`, summed over the sample sequences
    [(reference phase, frames), ...]."""
    total, used = None, None
    for phi0, frames in samples:
        used, ms = _time_one(phi0, params, frames, tuning, reps)
        total = ms if total is None else [a + b for a, b in zip(total, ms)]
    return used, total


# Timing samples: frames at three points of a synthetic sequence (its start,
# frame 100 and frame 180), each with the reference phase the chain itself
# reached there.  The column kernel skips transforms of columns whose
# thresholded signal is zero; how many there are, and how they cluster,
# changes as the chain's reference phase evolves, and the launch shape that
# wins on the first frames can lose on the later ones (measured at 1024^2: 2
# vs 4 column wavefronts swap order after about 70 frames), so the samples
# span a 200-frame sequence the way its average does.
_SAMPLE_T0 = (0, 100, 180)


def _samples(width, height, frames, params):
    phi0 = synth.phase(width, height, 3.0, 0.05)
    last = _SAMPLE_T0[-1]
    seq = gen_frames(width, height, last + frames, seed=1234)
    out, t, d = [], 0, DPRA(phi0, 1, params=params)  # default launch shapes
    for t0 in _SAMPLE_T0:
        if t0 > t:
            d(seq[t:t0])  # advance the chain to frame t0
            t = t0
        phr = phi0 if t0 == 0 else d.get_state()[0].astype(np.float64)
        out.append((phr, seq[t0:t0 + frames]))
    d.close()
    return out


def tune(width, height, params=None, frames=4, reps=3, cache=None, force=False, log=None):
    """Pick (planx, plany, cols_waves, rows_waves, ls_tile) for a width x height WFF2 /
    DPRA with `params`.  Returns a dict usable as `tuning=` for glsu.DPRA and
    glsu.WFT2, with the per-frame kernel times it measured (key "ms")."""
    params = params if params is not None else wff_params()
    path = cache or os.environ.get(CACHE_ENV, _DEFAULT_CACHE)
    k = key(width, height, params)
    table = _load(path) if path else {}
    if not force and k in table:
        return dict(table[k])
    sx, sy = _c_round(3 * params.sigma_x), _c_round(3 * params.sigma_y)
    samples = _samples(width, height, frames, params)
    say = log or (lambda *a: None)

    # 1. columns (row settings at their defaults)
    best_c = None
    for py in candidates(height + sy, False):
        for cw in (1, 2, 4):
            used, ms = _time(samples, params, dict(plany=py, cols_waves=cw), reps)
            if used["cols_waves"] != cw:  # LDS limit lowered it: same as a smaller one
                continue
            say(f"cols plany={py} waves={cw}: {ms[2] * 1e3:.1f} us")
            if best_c is None or ms[2] < best_c[0]:
                best_c = (ms[2], py, cw)
    # 2. rows
    best_r = None
    for px in candidates(width + sx, True):
        for rw in (1, 2, 4):
            used, ms = _time(samples, params, dict(planx=px, plany=best_c[1],
                                                   cols_waves=best_c[2], rows_waves=rw), reps)
            if used["rows_waves"] != rw:
                continue
            t = ms[1] + ms[3]
            say(f"rows planx={px} waves={rw}: {t * 1e3:.1f} us")
            if best_r is None or t < best_r[0]:
                best_r = (t, px, rw)
    tuning = dict(planx=best_r[1], plany=best_c[1], cols_waves=best_c[2], rows_waves=best_r[2])
    # 3. least-squares tile shape
    best_l = None
    for lt, shape in LS_TILES.items():
        _, ms = _time(samples, params, dict(tuning, ls_tile=lt), reps)
        say(f"least squares tile {shape[0]}x{shape[1]}: {ms[0] * 1e3:.1f} us")
        if best_l is None or ms[0] < best_l[0]:
            best_l = (ms[0], lt)
    tuning["ls_tile"] = best_l[1]
    _, ms = _time(samples, params, tuning, reps)
    res = dict(tuning, ms=ms)
    if path:
        table = _load(path)
        table[k] = res
        _save(path, table)
    return dict(res)


def tuning_only(res):
    """Strip measurement fields from a tune() result (fields a cached result
    predates keep the library default)."""
    return {k: res.get(k, -1) for k in TUNING_FIELDS}
