"""Pins the FP64 oracle's AIA (oracle/glsu_oracle.c, orc_aia*) -- CPU only.

  golden    the reference's own vectors: Test/Phi_CPU.csv holds the per-pixel
            solutions x = A1^-1 b of the phase step on Test/{1,2,3}.bmp for the
            initial shifts of Test/AIA_CPU_Test.cpp:150-153 (fixture
            tests/golden/aia_bmp.npz, made by make_aia_fixtures.py);
  closed    noise-free frames I = a + b cos(phi + delta): the phase step with the
  forms     true shifts returns the true phase, the shift step with the true
            phase returns the true shifts, and the full loop from a perturbed
            start returns phi + delta_0 and delta - delta_0;
  identities  M = 3 fits every pixel exactly; a common offset c on the shifts
            moves the phase by -c; the loop's iteration count, error modes and
            final wrap against half-steps driven from Python.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "aia_bmp.npz")
REF = "/root/reference"


def wrap(x):
    return np.angle(np.exp(1j * np.asarray(x)))


def synth(M, H, W, seed, noise=0.0, vary=None):
    """Frames M x H x W of I = a + b cos(phi + delta) with random shifts.
    vary=None: a = 120, b = 80 everywhere -- the model of both steps, so the
    true (phi, delta) is a fixed point of the loop; vary="pixel": a, b vary
    over the image (the phase step's model, I_k(j) = a_j + b_j cos(..)); vary=
    "frame": a, b vary from frame to frame (the shift step's model)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:H, 0:W]
    a, b = np.full((1, H, W), 120.0), np.full((1, H, W), 80.0)
    if vary == "pixel":
        a = (120 + 20 * np.sin(x / 17.0) * np.cos(y / 23.0))[None]
        b = (80 + 15 * np.cos(x / 11.0 + y / 29.0))[None]
    elif vary == "frame":
        a = rng.uniform(100, 140, (M, 1, 1))
        b = rng.uniform(60, 100, (M, 1, 1))
    phi = wrap(0.002 * ((x - W / 2) ** 2 + (y - H / 3) ** 2) + 0.31 * x - 0.17 * y)
    delta = rng.uniform(-np.pi, np.pi, M)
    f = a + b * np.cos(phi[None] + delta[:, None, None])
    if noise:
        f = f + rng.normal(0, noise, f.shape)
    return f, phi, delta


def test_golden_phase_step_matches_reference_csv(oracle):
    z = np.load(GOLD)
    f = z["frames"].astype(np.float64)
    d = z["deltas"].astype(np.float64)
    x = z["x_cpu"].astype(np.float64)
    phi, X = oracle.aia_phi(f, d, return_x=True)
    # the CSV has 6 significant digits (values up to ~400) from an FP32 solve
    assert np.abs(X - x).max() < 3e-3
    assert np.linalg.norm(X - x) / np.linalg.norm(x) < 1e-5
    e = wrap(phi.ravel() - np.arctan2(-x[:, 2], x[:, 1]))
    assert np.sqrt(np.mean(e ** 2)) < 1e-5 and np.abs(e).max() < 1e-4


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "Test")), reason="reference tree absent")
def test_golden_fixture_matches_reference_files():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mk", os.path.join(os.path.dirname(GOLD), "make_aia_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    z = np.load(GOLD)
    for k in (1, 2, 3):
        assert np.array_equal(mk.read_bmp8(os.path.join(REF, "Test", f"{k}.bmp")), z["frames"][k - 1])
    v = np.loadtxt(os.path.join(REF, "Test", "Phi_CPU.csv")).reshape(-1, 3).astype(np.float32)
    assert np.array_equal(v, z["x_cpu"])


def test_half_steps_closed_form(oracle):
    f, phi, delta = synth(7, 40, 56, 1, vary="pixel")
    p = oracle.aia_phi(f, delta)
    assert np.abs(wrap(p - phi)).max() < 1e-10
    # a sign or index slip (e.g. atan2(x2, x1), cos/sin swapped) fails this:
    assert np.abs(wrap(-p - phi)).max() > 0.1
    f, phi, delta = synth(7, 40, 56, 2, vary="frame")
    d = oracle.aia_delta(f, phi)
    assert np.abs(wrap(d - delta)).max() < 1e-10
    # per-pixel background breaks the shift step's model: it must then differ
    f, phi, delta = synth(7, 40, 56, 2, vary="pixel")
    assert np.abs(wrap(oracle.aia_delta(f, phi) - delta)).max() > 1e-4


@pytest.mark.parametrize("M,err_mode", [(4, 0), (5, 1), (12, 0)])
def test_full_loop_known_answer(oracle, M, err_mode):
    # M >= 4: with three frames every shift triple fits each pixel exactly, so
    # the loop stops at a start-dependent fixed point (not a known answer)
    f, phi, delta = synth(M, 48, 40, 10 + M)
    start = delta + np.random.default_rng(M).normal(0, 0.3, M)
    p, d, it, err = oracle.aia(f, start, max_iter=200, max_err=1e-12, err_mode=err_mode)
    assert it < 200 and err <= 1e-12
    assert np.abs(wrap(d - (delta - delta[0]))).max() < 1e-8
    assert np.abs(wrap(p - (phi + delta[0]))).max() < 1e-8
    assert d[0] == 0.0


def test_three_frames_fit_exactly(oracle):
    rng = np.random.default_rng(3)
    f = rng.uniform(0, 255, (3, 9, 11))
    d = np.array([0.3, 1.9, -2.2])
    _, X = oracle.aia_phi(f, d, return_x=True)
    c, s = np.cos(d), np.sin(d)
    rec = X[:, 0][None] + X[:, 1][None] * c[:, None] + X[:, 2][None] * s[:, None]
    assert np.abs(rec - f.reshape(3, -1)).max() < 1e-9


def test_shift_offset_moves_phase(oracle):
    f, _, delta = synth(6, 20, 24, 5, noise=3.0)
    p0 = oracle.aia_phi(f, delta)
    p1 = oracle.aia_phi(f, delta + 0.7)
    assert np.abs(wrap(p1 - (p0 - 0.7))).max() < 1e-10


@pytest.mark.parametrize("err_mode", [0, 1])
def test_loop_semantics_against_half_steps(oracle, err_mode):
    f, _, _ = synth(5, 30, 26, 7, noise=4.0)
    start = np.array([0.0, 0.9, -1.2, 2.4, 0.2])
    max_iter, max_err = 50, 1e-6
    d, err, it = start.copy(), 2 * max_err, 0
    while err > max_err and it < max_iter:
        old = d.copy()
        d = oracle.aia_delta(f, oracle.aia_phi(f, d))
        it += 1
        err = (np.abs((d - d[0]) - (old - old[0])).max() if err_mode
               else np.abs(d - old).max())
    phi = wrap(oracle.aia_phi(f, d) + d[0])
    dn = wrap(d - d[0])
    p2, d2, it2, err2 = oracle.aia(f, start, max_iter, max_err, err_mode)
    assert it2 == it and err2 == pytest.approx(err, rel=1e-9, abs=1e-15)
    assert np.abs(wrap(p2 - phi)).max() < 1e-12 and np.abs(d2 - dn).max() < 1e-12


def test_no_iterations_and_not_positive_definite(oracle):
    f, _, _ = synth(4, 10, 12, 9)
    start = np.array([0.5, 1.0, -1.0, 2.0])
    p, d, it, err = oracle.aia(f, start, max_iter=0, max_err=1e-4)
    assert it == 0 and err == 2e-4
    assert np.abs(wrap(p - (oracle.aia_phi(f, start) + 0.5))).max() < 1e-12
    assert np.abs(d - wrap(start - 0.5)).max() < 1e-15
    # max_err <= 0: the loop never starts (err = 2*max_err is not > max_err)
    assert oracle.aia(f, start, max_iter=10, max_err=0.0)[2] == 0
    with pytest.raises(ArithmeticError):
        oracle.aia_phi(f, np.full(4, 0.3))
