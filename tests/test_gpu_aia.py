"""AIA on the GPU (libglsu, glsu.AIA) against the FP64 oracle and the
reference's golden vectors (Test/{1,2,3}.bmp with Test/Phi_CPU.csv, fixture
tests/golden/aia_bmp.npz).  Bars: the phase step and the final phase within
1e-4 rad RMS of the oracle, the shifts within 2e-5 rad, the same iteration
count; FP32 per-pixel arithmetic with FP64 reductions and solves."""
import os

import numpy as np
import pytest

from _util import aia_frames, say, wrap, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "aia_bmp.npz")


def test_golden_phase_step_on_reference_bitmaps():
    z = np.load(GOLD)
    fr = torch.from_numpy(z["frames"]).cuda()
    d0 = z["deltas"]
    x = z["x_cpu"].astype(np.float64)
    say("[aia golden] 3 x 256 x 256 uint8, max_iter 0")
    a = glsu.AIA(3, 256, 256)
    phi, d, ms, it, err = a(fr, deltas=d0, max_iter=0)
    assert it == 0 and err == pytest.approx(2e-4)
    # phi + delta_0 from the reference's per-pixel solutions
    ref = wrap(np.arctan2(-x[:, 2], x[:, 1]) + np.float64(d0[0])).reshape(256, 256)
    rms, mx = wrapped_err(phi.cpu().numpy(), ref)
    say(f"   phase vs Phi_CPU.csv: rms {rms:.2e} max {mx:.2e} ({ms:.3f} ms)")
    assert rms < 1e-5 and mx < 1e-4
    assert np.abs(d - wrap(d0.astype(np.float64) - d0[0])).max() < 1e-6


@pytest.mark.parametrize("err_mode", [0, 1])
def test_reference_bitmaps_full_run_matches_oracle(oracle, err_mode):
    z = np.load(GOLD)
    f = z["frames"]
    d0 = z["deltas"].astype(np.float64)
    # Test/AIA_CPU_Test.cpp:157-158: 55 iterations, 1e-4
    p_ref, d_ref, it_ref, err_ref = oracle.aia(f.astype(np.float64), d0, 55, 1e-4, err_mode)
    say(f"[aia bitmaps] err_mode {err_mode}: oracle {it_ref} iterations, err {err_ref:.3e}")
    a = glsu.AIA(3, 256, 256)
    phi, d, ms, it, err = a(torch.from_numpy(f).cuda(), deltas=z["deltas"], max_iter=55,
                            max_err=1e-4, err_mode=err_mode)
    rms, mx = wrapped_err(phi.cpu().numpy(), p_ref)
    say(f"   gpu {it} iterations, err {err:.3e}; shifts {np.abs(wrap(d - d_ref)).max():.2e}; "
        f"phase rms {rms:.2e} max {mx:.2e} ({ms:.3f} ms)")
    assert it == it_ref
    assert np.abs(wrap(d - d_ref)).max() < 2e-5
    assert rms < 1e-4


CASES = [  # (M, H, W, dtype, err_mode)
    (4, 256, 256, np.uint8, 0),
    (7, 333, 257, np.float32, 1),    # odd pixel count: scalar loads
    (20, 512, 512, np.uint8, 0),
    (9, 200, 96, np.float64, 1),
    (40, 128, 160, np.float32, 0),   # two frame chunks (32 + 8)
    (70, 96, 64, np.uint8, 1),       # three chunks
]


@pytest.mark.parametrize("M,H,W,dtype,err_mode", CASES)
def test_synthetic_matches_oracle(oracle, M, H, W, dtype, err_mode):
    f, phi_true, delta_true = aia_frames(M, H, W, 100 + M, dtype=dtype)
    start = np.random.default_rng(M).normal(0, 1, M)
    start[0] = 0
    p_ref, d_ref, it_ref, err_ref = oracle.aia(f.astype(np.float64), start, 20, 1e-4, err_mode)
    say(f"[aia synthetic] M={M} {H}x{W} {np.dtype(dtype).name} err_mode {err_mode}: "
        f"oracle {it_ref} it, err {err_ref:.2e}")
    a = glsu.AIA(M, W, H)
    say(f"   launch {a.info()}")
    phi, d, ms, it, err = a(torch.from_numpy(f).cuda(), deltas=start.astype(np.float32),
                            max_iter=20, max_err=1e-4, err_mode=err_mode)
    rms, mx = wrapped_err(phi.cpu().numpy(), p_ref)
    dd = np.abs(wrap(d - d_ref)).max()
    say(f"   gpu {it} it, err {err:.2e}; shifts {dd:.2e}; phase rms {rms:.2e} max {mx:.2e} "
        f"({ms:.3f} ms); vs truth {wrapped_err(p_ref, wrap(phi_true + delta_true[0]))[0]:.2e}")
    # the float32 start shifts are what the GPU starts from; the oracle's
    # start is the same float32 values
    assert it == it_ref
    assert dd < 2e-5
    assert rms < 1e-4 and mx < 2e-3


def test_host_entry_and_determinism(oracle):
    M, H, W = 6, 120, 100
    f, _, _ = aia_frames(M, H, W, 7, dtype=np.uint8)
    a = glsu.AIA(M, W, H)
    say("[aia host/device] 6 x 120 x 100 uint8, seeded start")
    p1, d1, _, it1, e1 = a(torch.from_numpy(f).cuda(), seed=11)
    p2, d2, _, it2, e2 = a(torch.from_numpy(f).cuda(), seed=11)
    p3, d3, _, it3, e3 = a(list(f), seed=11)
    assert torch.equal(p1, p2) and np.array_equal(d1, d2) and it1 == it2 and e1 == e2
    assert np.array_equal(p1.cpu().numpy(), p3) and np.array_equal(d1, d3) and it1 == it3
    # the seeded start: delta_0 = 0, N(0,1) draws -- same oracle answer from it
    _, dseed, _, _, _ = a(torch.from_numpy(f).cuda(), seed=11, max_iter=0)
    p_ref, d_ref, it_ref, _ = oracle.aia(f.astype(np.float64), dseed.astype(np.float64), 20, 1e-4)
    assert it1 == it_ref and np.abs(wrap(d1 - d_ref)).max() < 2e-5


def test_many_iterations_poll_and_errors(oracle):
    M, H, W = 5, 64, 96
    f, _, _ = aia_frames(M, H, W, 21, noise=8.0, dtype=np.float32)
    start = np.array([0, 1.0, -1.0, 2.0, -2.5], np.float32)
    a = glsu.AIA(M, W, H)
    say("[aia poll] 5 x 64 x 96 float32, max_iter 60, max_err 1e-9")
    phi, d, ms, it, err = a(torch.from_numpy(f).cuda(), deltas=start, max_iter=60, max_err=1e-9)
    p_ref, d_ref, it_ref, _ = oracle.aia(f.astype(np.float64), start.astype(np.float64), 60, 1e-9)
    say(f"   gpu {it} it (err {err:.2e}), oracle {it_ref} it")
    assert it <= 60
    assert np.abs(wrap(d - d_ref)).max() < 2e-5
    assert wrapped_err(phi.cpu().numpy(), p_ref)[0] < 1e-4
    # equal shifts: the phase step's normal matrix is singular
    with pytest.raises(glsu.NotPositiveDefiniteError):
        a(torch.from_numpy(f).cuda(), deltas=np.full(M, 0.4, np.float32))
    # the handle still works afterwards
    phi2, d2, _, it2, _ = a(torch.from_numpy(f).cuda(), deltas=start, max_iter=60, max_err=1e-9)
    assert torch.equal(phi, phi2) and it2 == it
    # iteration timing: all launches run, the handle stays usable
    assert a.time_iter(torch.from_numpy(f).cuda(), start, reps=5, stage=0) > 0
    assert a.time_iter(torch.from_numpy(f).cuda(), start, reps=5, stage=1) > 0
    phi3, _, _, it3, _ = a(torch.from_numpy(f).cuda(), deltas=start, max_iter=60, max_err=1e-9)
    assert torch.equal(phi, phi3) and it3 == it
    with pytest.raises(glsu.GlsuError):
        glsu.AIA(2, W, H)
    with pytest.raises(ValueError):
        a(torch.from_numpy(f[:4]).cuda())
