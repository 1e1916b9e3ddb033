"""Autotuner: the tuned choice is valid, cached, and results-invariant.

Every (plan, launch shape) the tuner may pick must give the same WFF2 / DPRA
result, within FP32 rounding, as the default choice and the FP64 oracle.
"""
import numpy as np
import pytest

from _util import say, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402
from glsu import autotune  # noqa: E402


def _all_tunings(W, H, params):
    sx = sy = int(np.floor(3 * params.sigma_x + 0.5))
    out = []
    for py in autotune.candidates(H + sy, False):
        for cw in (1, 2, 4):
            out.append(dict(plany=py, cols_waves=cw))
    for px in autotune.candidates(W + sx, True):
        for rw in (1, 2, 4):
            out.append(dict(planx=px, rows_waves=rw))
    return out


def test_every_tuning_gives_the_same_wff2(oracle):
    W, H = 300, 212
    params = glsu.wff_params(10.0, -0.2, 0.2, 0.1)
    rng = np.random.default_rng(3)
    ph = 0.3 * np.arange(W)[None, :] + 0.1 * np.arange(H)[:, None] + rng.normal(0, 0.4, (H, W))
    f = np.exp(1j * ph).astype(np.complex64)
    ref = oracle.WFF2(W, H, 10.0, -0.2, 0.2, 0.1)(f.astype(np.complex128))
    fd = torch.from_numpy(f).cuda()
    base = glsu.WFT2(W, H, params=params)(fd).cpu().numpy()
    for t in _all_tunings(W, H, params):
        w = glsu.WFT2(W, H, params=params, tuning=t)
        out = w(fd).cpu().numpy()
        used = w.tuning()
        rel = np.abs(out - ref).max() / np.abs(ref).max()
        relb = np.abs(out - base).max() / np.abs(base).max()
        say(f"[tuning {t}] used={used} rel(oracle)={rel:.2e} rel(default)={relb:.2e}")
        for k, v in t.items():
            assert used[k] == v or k.endswith("waves")  # waves may be capped by LDS
        assert rel < 1e-4 and relb < 1e-4


def test_tune_picks_valid_cached_choice(oracle, tmp_path):
    W, H, T = 384, 256, 4
    cache = str(tmp_path / "tune.json")
    res = autotune.tune(W, H, frames=3, reps=2, cache=cache, log=say)
    say(f"[tune] {res}")
    plans = {p["id"]: p for p in glsu.fft_plans()}
    assert plans[res["planx"]]["T"] == 64 and plans[res["planx"]]["N"] >= W + 30
    assert plans[res["plany"]]["N"] >= H + 30
    assert res["cols_waves"] in (1, 2, 4) and 1 <= res["rows_waves"] <= 4
    assert res["ls_tile"] in glsu._lib.LS_TILES
    assert len(res["ms"]) == 4 and all(m > 0 for m in res["ms"])
    # second call: served from the cache
    again = autotune.tune(W, H, cache=cache)
    assert autotune.tuning_only(again) == autotune.tuning_only(res)
    # the tuned DPRA follows the oracle like the default one does
    frames, phi0 = oracle.gen_sequence(W, H, T, 21, dtype=np.float32)
    od = oracle.DPRA(phi0)
    ref = od.run(frames.astype(np.float64))
    fr = torch.from_numpy(frames).cuda()
    outs = []
    for t in (None, autotune.tuning_only(res)):
        d = glsu.DPRA(phi0, tuning=t)
        outs.append(d(fr).cpu().numpy())
    for i in range(T):
        rms, mx = wrapped_err(outs[1][i], ref[i])
        rms2, _ = wrapped_err(outs[1][i], outs[0][i])
        say(f"   frame {i}: tuned vs oracle rms={rms:.2e} max={mx:.2e}; vs default rms={rms2:.2e}")
        assert rms < 1e-3
        assert rms2 < 1e-4


def test_every_ls_tile_gives_bitwise_identical_chains(oracle):
    """The least-squares tile shape changes only how pixels are grouped into
    workgroups: every shape gives the default's chain output bit for bit
    (odd size, so partial tiles at both edges)."""
    W, H, T = 333, 211, 3
    frames, phi0 = oracle.gen_sequence(W, H, T, 23, dtype=np.float32)
    fr = torch.from_numpy(frames).cuda()
    base = glsu.DPRA(phi0)(fr).cpu().numpy()
    for tile, shape in glsu._lib.LS_TILES.items():
        d = glsu.DPRA(phi0, tuning=dict(ls_tile=tile))
        assert d.tuning()["ls_tile"] == tile
        out = d(fr).cpu().numpy()
        say(f"[ls tile {tile} = {shape}] bitwise equal to the default: {np.array_equal(out, base)}")
        assert np.array_equal(out, base), np.abs(out - base).max()
