"""Pins of the oracle's DPRA (DPRA/dpra_cpu.cpp:9-258): the 3x3 least squares
against numpy.linalg.solve (including border pixels and the A00 = 9 quirk),
closed forms on noise-free fringes, invariances, the singular case, the phase
update and the i % rr reference-update cadence."""
import numpy as np
import pytest

import _bruteforce as bf


def test_ls_matches_numpy_solve(oracle):
    rng = np.random.default_rng(1)
    H, W = 9, 11
    phr = rng.uniform(-np.pi, np.pi, (H, W))
    I = rng.uniform(0, 255, (H, W))
    z = oracle.DPRA(phr, sigx=1.0, wxl=-0.5, wxh=0.5, wxi=0.5, nthreads=2).ls(I)
    ref = bf.dpra_ls(I, phr)
    assert np.abs(z - ref).max() < 1e-10


def test_solve3_matches_numpy(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        M = rng.standard_normal((3, 3))
        A = M @ M.T + 0.1 * np.eye(3)
        b = rng.standard_normal(3)
        x, sing = oracle.solve3_sym(A, b)
        assert not sing
        assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-10
    # only the upper triangle is read (column-major 'U', as LAPACKE_dsysv)
    A = np.array([[4.0, 1.0, 0.5], [99.0, 3.0, 0.2], [99.0, 99.0, 2.0]])
    x, _ = oracle.solve3_sym(A, [1.0, 2.0, 3.0])
    As = np.triu(A) + np.triu(A, 1).T
    assert np.abs(x - np.linalg.solve(As, [1.0, 2.0, 3.0])).max() < 1e-12


def _carrier_phase(H, W):
    y, x = np.mgrid[0:H, 0:W]
    return 0.45 * x + 0.3 * y + 0.002 * (x - W / 2) ** 2


def test_ls_closed_form_noise_free(oracle):
    # I = a + b cos(phr + Delta) at interior pixels -> the solve returns Delta exactly
    H, W = 20, 24
    phr = _carrier_phase(H, W)
    delta = 0.37
    for a, b in [(128.0, 100.0), (10.0, 3.0)]:
        I = a + b * np.cos(phr + delta)
        z = oracle.DPRA(phr, sigx=1.0, wxl=-0.5, wxh=0.5, wxi=0.5).ls(I)
        e = np.angle(z[1:-1, 1:-1] * np.exp(-1j * delta))
        assert np.abs(e).max() < 1e-9


def test_ls_gain_offset_invariance_and_border_quirk(oracle):
    H, W = 16, 18
    phr = _carrier_phase(H, W)
    rng = np.random.default_rng(3)
    I = 128 + 100 * np.cos(phr + 0.2 + rng.normal(0, 0.3, (H, W)))
    d = oracle.DPRA(phr, sigx=1.0, wxl=-0.5, wxh=0.5, wxi=0.5)
    z0 = d.ls(I)
    z1 = d.ls(2.5 * I + 40.0)
    inner = (slice(1, -1), slice(1, -1))
    assert np.abs(z1[inner] - z0[inner]).max() < 1e-10          # interior: invariant
    z2 = d.ls(2.5 * I)                                           # gain only: invariant everywhere
    assert np.abs(z2 - z0).max() < 1e-10
    # border pixels use A00 = 9 with 4 or 6 neighbours: an offset changes them
    assert np.abs(z1[0, :] - z0[0, :]).max() > 1e-6


def test_ls_singular_defined_behaviour(oracle):
    # constant reference phase: A has rank 1 -> Delta := 0 (z = 1)
    H, W = 6, 7
    phr = np.full((H, W), 0.8)
    I = np.full((H, W), 100.0)
    z = oracle.DPRA(phr, sigx=1.0, wxl=-0.5, wxh=0.5, wxi=0.5).ls(I)
    assert np.all(z == 1.0)


@pytest.mark.parametrize("rr", [1, 2, 3])
def test_chain_closed_form_cumulative(oracle, rr):
    # noise-free frames with a uniform phase step delta per frame: WFF2 of the
    # constant field exp(j delta) keeps its phase, so the accumulated phase
    # change is (t + 1) delta for every update cadence rr.  Border pixels are
    # biased (A00 = 9 with 4 or 6 neighbours); through the reference phase that
    # bias moves 2s pixels inward per frame, so frame t is exact beyond
    # (t + 1) * 2s + 1 pixels from the border.
    H, W, T, delta, sig = 84, 88, 5, 0.21, 1.0
    s2 = 2 * 3
    phr = _carrier_phase(H, W)
    frames = np.stack([128 + 100 * np.cos(phr + (t + 1) * delta) for t in range(T)])
    d = oracle.DPRA(phr, rr=rr, sigx=sig, wxl=-0.3, wxh=0.3, wxi=0.3, float_cast=False)
    out = d.run(frames)
    for t in range(T):
        m = (t + 1) * s2 + 1
        e = out[t, m:-m, m:-m] - (t + 1) * delta
        assert np.abs(e).max() < 1e-9, (rr, t)


def test_float_cast_of_t(oracle):
    # dpra_cpu.cpp:246 stores t = atan2(F) in a float; the flag turns it off
    H, W, T = 40, 44, 3
    phr = _carrier_phase(H, W)
    rng = np.random.default_rng(4)
    frames = np.stack([128 + 100 * np.cos(phr + 0.1 * (t + 1) + rng.normal(0, 0.05, (H, W)))
                       for t in range(T)])
    a = oracle.DPRA(phr, sigx=2.0, wxl=-0.3, wxh=0.3, wxi=0.3).run(frames)
    b = oracle.DPRA(phr, sigx=2.0, wxl=-0.3, wxh=0.3, wxi=0.3, float_cast=False).run(frames)
    diff = np.abs(a - b).max()
    assert 0 < diff < 2.5e-7 * T * 4
    # with the cast, every per-frame increment is a float32 value
    inc = a[0] - 0.0
    assert np.all(inc.astype(np.float32).astype(np.float64) == inc)


def test_update_ref_semantics(oracle):
    # phicurr = wrap(phiref + t), dphi = dphiref + t; update copies both
    H, W = 60, 64
    phr = _carrier_phase(H, W)
    f1 = 128 + 100 * np.cos(phr + 0.3)
    d = oracle.DPRA(phr, sigx=1.0, wxl=-0.3, wxh=0.3, wxi=0.3, float_cast=False)
    dphi1, phic1 = d.dpra_per_frame(f1)
    assert np.abs(np.angle(np.exp(1j * (phic1 - phr - dphi1)))).max() < 1e-12
    assert np.all(np.abs(phic1) <= np.pi)
    # without update_ref_phi the same frame gives the same answer
    dphi1b, _ = d.dpra_per_frame(f1)
    assert np.array_equal(dphi1, dphi1b)
    d.update_ref_phi()
    dphi2, _ = d.dpra_per_frame(f1)          # same frame again: increment ~ 0
    m = 2 * 6 + 1                            # two frames of border-bias propagation (2s = 6)
    assert np.abs(np.angle(np.exp(1j * (dphi2 - dphi1))))[m:-m, m:-m].max() < 1e-9


def test_chain_amplifies_rounding_level_differences(oracle):
    """The DPRA recursion on the 256^2, 20-frame sequence (amplitude 3 + 0.1 t,
    carrier 0.05, noise 10) carries a difference of one part in 1e15 in phi0
    to > 1e-3 rad RMS by its last frame: where the peaks gradient cancels the
    carrier the 3x3 systems are near-singular and each frame amplifies what it
    is given.  A chained comparison of two implementations therefore cannot
    stay within 1e-3 rad to the end unless they agree bit for bit; per-frame
    (one-step) comparisons are what pin the arithmetic.  The per-frame phase
    increment is rounded to float (DPRA/dpra_cpu.cpp:246), so a different
    order of the FP64 pair sums leaves the chain unchanged."""
    W = H = 256
    T = 20
    frames, phi0 = oracle.gen_sequence(W, H, T, 7, dtype=np.float32)
    f64 = frames.astype(np.float64)
    a = oracle.DPRA(phi0, nthreads=4).run(f64)
    b = oracle.DPRA(phi0 * (1 + 1e-15), nthreads=4).run(f64)
    c = oracle.DPRA(phi0, nthreads=3, pair_parallel=True).run(f64)
    first, _ = oracle.wrapped_err(b[0], a[0])
    last, _ = oracle.wrapped_err(b[-1], a[-1])
    assert first < 1e-10 and last > 1e-3
    assert np.array_equal(a, c)
