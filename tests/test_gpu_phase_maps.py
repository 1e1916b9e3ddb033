"""The drop-in on the reference's own measured phase maps (Experiments/imgs/
1/1_phi.bin 256x256 and 2/2_phi.bin 366x371, fixture tests/golden/phi_maps.npz):
WFF2 of exp(j phi) and a DPRA chain seeded with the map, against the FP64 oracle.
The frames are synthetic (the reference's frame sequences are not in its
tree): I_t = 128 + 100 cos(phi + t d) + N(0, 5^2), d a smooth phase change."""
import os

import numpy as np
import pytest

from _util import say, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "phi_maps.npz")


@pytest.mark.parametrize("key", ["phi1", "phi2"])
def test_wff2_of_measured_map(oracle, key):
    phi = np.load(GOLDEN)[key].astype(np.float64)
    H, W = phi.shape
    f = np.exp(1j * phi).astype(np.complex64)
    ref = oracle.WFF2(W, H)(f.astype(np.complex128))
    out = glsu.WFT2(W, H)(torch.from_numpy(f).cuda()).cpu().numpy()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    # where the map's local fringe frequency lies outside the (wx, wy) grid every
    # pair is thresholded away and F ~ 0: its angle is rounding noise on both
    # sides, so the phase is compared where |F| is at least 1 % of its maximum
    live = np.abs(ref) > 1e-2 * np.abs(ref).max()
    rms, mx = wrapped_err(np.angle(out[live]), np.angle(ref[live]))
    say(f"[{key} {W}x{H}] WFT2 rel={rel:.2e}; phase on {live.mean():.0%} of pixels: "
        f"rms={rms:.2e} max={mx:.2e}")
    assert rel < 1e-4 and rms < 1e-4


@pytest.mark.parametrize("key", ["phi1", "phi2"])
def test_dpra_chain_on_measured_map(oracle, key):
    phi = np.load(GOLDEN)[key].astype(np.float64)
    H, W = phi.shape
    T = 4
    y, x = np.mgrid[0:H, 0:W]
    d = 0.12 * np.cos(2 * np.pi * x / W) * np.sin(np.pi * y / H)
    rng = np.random.default_rng(7)
    frames = np.stack([128 + 100 * np.cos(phi + t * d) + rng.normal(0, 5, (H, W))
                       for t in range(T)]).astype(np.float32)
    ref = oracle.DPRA(phi).run(frames.astype(np.float64))
    out = glsu.DPRA(phi)(torch.from_numpy(frames).cuda()).cpu().numpy()
    for t in range(T):
        rms, mx = wrapped_err(out[t], ref[t])
        say(f"[{key}] frame {t}: rms={rms:.2e} max={mx:.2e}")
        assert rms < 1e-3
