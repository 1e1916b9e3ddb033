"""Edge cases of the drop-in boundary against the FP64 oracle: the frame
dtypes the reference's interfaces take (uint8 cv::Mat data, DPRA/dpra_cpu.cpp:181;
float and double), degenerate image sizes, rr > 1 with repeated calls, the
host-pointer entry points, zero-frame runs, and the C++ wrapper (include/glsu.hpp)
driving the library end to end."""
import os
import subprocess

import numpy as np
import pytest

from _util import say, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dtype", [np.uint8, np.float32, np.float64])
def test_frame_dtypes_match_oracle(oracle, dtype):
    W, H, T = 200, 150, 3
    frames, phi0 = oracle.gen_sequence(W, H, T, 61, dtype=np.float64)
    if dtype == np.uint8:
        frames = np.clip(np.rint(frames), 0, 255)
    frames = frames.astype(dtype)
    ref = oracle.DPRA(phi0).run(frames.astype(np.float64))
    out = glsu.DPRA(phi0)(torch.from_numpy(frames).cuda()).cpu().numpy()
    for t in range(T):
        rms, mx = wrapped_err(out[t], ref[t])
        say(f"[dtype {np.dtype(dtype).name}] frame {t}: rms={rms:.2e} max={mx:.2e}")
        assert rms < 1e-3


@pytest.mark.parametrize("W,H", [(1, 1), (1, 40), (40, 1), (2, 2), (3, 5), (64, 64), (65, 63)])
def test_degenerate_sizes_match_oracle(oracle, W, H):
    sig, lo, hi, st = 2.0, -0.6, 0.6, 0.3
    rng = np.random.default_rng(W * 100 + H)
    ph = 0.5 * np.arange(W)[None, :] + 0.3 * np.arange(H)[:, None] + rng.normal(0, 0.3, (H, W))
    f = np.exp(1j * ph).astype(np.complex64)
    ref = oracle.WFF2(W, H, sig, lo, hi, st)(f.astype(np.complex128))
    out = glsu.WFT2(W, H, params=glsu.wff_params(sig, lo, hi, st))(torch.from_numpy(f).cuda())
    out = out.cpu().numpy()
    scale = max(np.abs(ref).max(), 1e-30)
    rel = np.abs(out - ref).max() / scale
    say(f"[size {W}x{H}] WFT2 rel={rel:.2e}")
    assert rel < 1e-4
    # the DPRA chain on the same size (3x3 neighbourhoods clipped everywhere);
    # the generator's peaks surface spans [-3, 3] whatever the size, so small
    # images get a low amplitude to stay sampled (an aliased phase makes 3x3
    # systems near-singular, where FP32 and FP64 rounding of phi0 alone
    # decide the answer)
    frames, phi0 = oracle.gen_sequence(W, H, 2, 62, amp0=0.3, amp1=0.03, carrier=0.08,
                                       dtype=np.float32)
    od = oracle.DPRA(phi0, sigx=sig, wxl=lo, wxh=hi, wxi=st)
    r = od.run(frames.astype(np.float64))
    g = glsu.DPRA(phi0, params=glsu.wff_params(sig, lo, hi, st))(torch.from_numpy(frames).cuda())
    for t in range(2):
        rms, mx = wrapped_err(g[t].cpu().numpy(), r[t])
        say(f"   DPRA frame {t}: rms={rms:.2e} max={mx:.2e}")
        assert rms < 1e-3


def test_rr3_repeated_calls_and_host_entry_points(oracle):
    W, H = 160, 120
    frames, phi0 = oracle.gen_sequence(W, H, 6, 63, dtype=np.float32)
    od = oracle.DPRA(phi0, rr=3)
    ref = np.concatenate([od.run(frames[:4].astype(np.float64)),
                          od.run(frames[4:].astype(np.float64))])
    d = glsu.DPRA(phi0, rr=3)
    fr = torch.from_numpy(frames).cuda()
    out = torch.cat([d(fr[:4].contiguous()), d(fr[4:].contiguous())]).cpu().numpy()
    for t in range(6):
        rms, _ = wrapped_err(out[t], ref[t])
        assert rms < 1e-3, t
    # zero frames: a no-op
    assert d(fr[:0].contiguous()).shape[0] == 0
    # host-pointer per-frame entry point (the reference's dpra_per_frame form)
    import ctypes
    from glsu._lib import lib, check
    h2 = glsu.DPRA(phi0)
    host = np.empty((H, W), np.float32)
    ms = ctypes.c_double()
    od1 = oracle.DPRA(phi0)
    for t in range(2):
        f = np.ascontiguousarray(frames[t])
        check(lib().glsu_dpra_frame_host(h2._h, f.ctypes.data, glsu.F32, host.ctypes.data,
                                         ctypes.byref(ms)))
        check(lib().glsu_dpra_update_ref(h2._h))
        r, _ = od1.dpra_per_frame(frames[t].astype(np.float64))
        od1.update_ref_phi()
        rms, _ = wrapped_err(host, r)
        say(f"[host entry] frame {t}: rms={rms:.2e} ({ms.value:.3f} ms)")
        assert rms < 1e-3


def test_cpp_wrapper_example_runs():
    exe = os.path.join(ROOT, "g-ls3u_amd", "lib", "example_dpra")
    if not os.path.exists(exe):
        pytest.skip("example_dpra not built")
    r = subprocess.run([exe, "256", "192", "6"], capture_output=True, text=True, timeout=120)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    # "centre dphi <v> (model <m>)": the recovered phase change follows the model
    toks = r.stdout.split()
    v = float(toks[toks.index("dphi") + 1])
    m = float(toks[toks.index("(model") + 1].rstrip(");"))
    assert abs(v - m) < 0.05
    # the same loop through DPRA_HYBRID_HIPF (the DPRA_HYBRIDF constructor:
    # the reference's WFF parameters, sigma 20, +-0.15 step 0.05, thr 15)
    r = subprocess.run([exe, "256", "192", "6", "hybrid"], capture_output=True, text=True,
                       timeout=120)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    toks = r.stdout.split()
    v = float(toks[toks.index("dphi") + 1])
    m = float(toks[toks.index("(model") + 1].rstrip(");"))
    assert "hybrid" in r.stdout and abs(v - m) < 0.05
    # the constructor's padding error is the reference's int -1 throw
    r = subprocess.run([exe, "4096", "16", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "threw -1" in r.stdout


def test_cpp_aia_and_wfr_classes_example_runs():
    # AIA::AIA_HIPF on six shifted 8-bit frames and WFT2_HIPF in WFR mode on a
    # plane wave, from C++: the program checks the recovered shifts, phase,
    # ridge frequency and zero curvature itself and exits 0 when they hold
    exe = os.path.join(ROOT, "g-ls3u_amd", "lib", "example_classes")
    if not os.path.exists(exe):
        pytest.skip("example_classes not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert "example_classes: ok" in r.stdout


def test_profile_counts_skipped_column_transforms(oracle):
    """glsu_dpra_profile_skips: with thr = 0 every thresholded sample survives
    (no transform is skipped); with a threshold above every |sample| every
    column is zero after thresholding (all are skipped, and the output is
    the zero field's phase).  The default threshold lies in between."""
    W, H, T = 160, 96, 2
    frames, phi0 = oracle.gen_sequence(W, H, T, 5, dtype=np.float32)
    fr = torch.from_numpy(frames).cuda()
    total = T * 5 * W * 5
    for thr, expect in ((0.0, 0), (1e30, total), (-1.0, None)):
        d = glsu.DPRA(phi0, params=glsu.wff_params(10.0, -0.2, 0.2, 0.1, thr=thr))
        ms, skipped, transforms = d.profile_skips(fr)
        say(f"[skips] thr={thr}: {skipped} of {transforms} column transforms skipped")
        assert transforms == total
        if expect is not None:
            assert skipped == expect
        else:
            assert 0 <= skipped < total


def test_caller_tensors_checked_before_launch():
    """Tensors whose raw addresses reach the library (AIA's phase output, the
    U-Net loss target, the 8-bit output stage, the frame generator's output)
    are refused in Python when their device, shape or dtype is wrong."""
    say("[tensor checks] AIA 4x64x48, UNet 2x32, phase_to_u8, gen_frames")
    M, W, H = 4, 64, 48
    a = glsu.AIA(M, W, H)
    frames = torch.zeros((M, H, W), dtype=torch.float32, device="cuda")
    for bad in (torch.empty((H, W), dtype=torch.float32),
                torch.empty((H, W + 1), dtype=torch.float32, device="cuda"),
                torch.empty((H, W), dtype=torch.float64, device="cuda")):
        with pytest.raises(ValueError):
            a(frames, phi=bad)
    with pytest.raises(ValueError):
        a.time_iter(frames[:, :, :W // 2].contiguous(), None)
    net = glsu.UNet(2, 32)
    x = torch.zeros((1, 32, 32), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        net.loss_grad(x, torch.zeros((1, 32, 16), dtype=torch.complex64, device="cuda"))
    phi = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        glsu.phase_to_u8(phi, out=torch.empty((H, W), dtype=torch.uint8))
    with pytest.raises(ValueError):
        glsu.gen_frames(W, H, 2, 1, out=torch.empty((1, H, W), dtype=torch.float32, device="cuda"))
    assert glsu.phase_to_u8(phi).shape == (H, W)
    assert glsu.gen_frames(W, H, 2, 1).shape == (2, H, W)
