"""Pins of the OPT padding table (WFT/WFT.h:18-32, WFT/WFT.cpp:10-23,
MatLab/FFT_OptimalSize.m:1-23), the (wx, wy) grid size rule
(WFT/WFT2_CPU.cpp:222-223) and the synthetic fringe generator."""
import numpy as np
import pytest


def test_opt_table_regenerated_from_matlab_rule(oracle):
    ref = sorted({2 ** i * 3 ** j * 5 ** k * 7 ** l
                  for i in range(13) for j in range(13) for k in range(13) for l in range(13)
                  if 2 ** i * 3 ** j * 5 ** k * 7 ** l <= 4096
                  and (2 ** i * 3 ** j * 5 ** k * 7 ** l) % 2 == 0})
    tab = oracle.opt_table()
    assert tab == ref
    assert tab[:8] == [2, 4, 6, 8, 10, 12, 14, 16] and tab[-1] == 4096 and len(tab) == 190


@pytest.mark.parametrize("x,expect", [(316, 320), (1084, 1120), (2108, 2160), (2120, 2160),
                                      (572, 576), (828, 840), (1596, 1600), (2, 2), (1, 2),
                                      (4096, 4096), (4097, -1), (1120, 1120), (1121, 1134)])
def test_opt_size_known_answers(oracle, x, expect):
    assert oracle.opt_size(x) == expect


def test_pair_count_known_answers(oracle):
    assert oracle.pair_count(-0.2, 0.2, 0.1) == 5
    assert oracle.pair_count(-0.5, 0.5, 0.05) == 21
    f = np.float32
    assert oracle.pair_count(float(f(-0.15)), float(f(0.15)), float(f(0.05))) == 7
    # WFF defaults of WFT2_cpu(W, H, WFF, z): [-2.3, 2.3] step 0.1
    assert oracle.pair_count(-2 - 3.0 / 10, 2 + 3.0 / 10, 1.0 / 10) == 47


def test_peaks_known_values(oracle):
    assert abs(oracle.lib().orc_peaks(0.0, 0.0) - 8 / (3 * np.e)) < 1e-15
    # X = 1, Y = -1: first term 0, X/5 - X^3 - Y^5 = 0.2
    assert abs(oracle.lib().orc_peaks(1.0, -1.0) - (-10 * 0.2 * np.exp(-2) - np.exp(-5) / 3)) < 1e-14


def test_normal_generator_statistics(oracle):
    v = np.array([oracle.normal(7, 3, p) for p in range(20000)])
    assert abs(v.mean()) < 0.03 and abs(v.std() - 1) < 0.03
    assert oracle.normal(7, 3, 11) == oracle.normal(7, 3, 11)
    assert oracle.normal(7, 3, 11) != oracle.normal(7, 4, 11)


def test_generator_frame_definition(oracle):
    W, H = 40, 30
    I, phi = oracle.gen_frame(W, H, 2, 5, amp0=3, amp1=0.1, carrier=0.05, noise=0.0)
    X = np.linspace(-3, 3, W)[None, :]
    Y = np.linspace(-3, 3, H)[:, None]
    pk = (3 * (1 - X) ** 2 * np.exp(-X ** 2 - (Y + 1) ** 2)
          - 10 * (X / 5 - X ** 3 - Y ** 5) * np.exp(-X ** 2 - Y ** 2)
          - np.exp(-(X + 1) ** 2 - Y ** 2) / 3)
    ref = 3.2 * pk + 2 * np.pi * 0.05 * np.arange(W)[None, :]
    assert np.abs(phi - ref).max() < 1e-12
    assert np.abs(I - (128 + 100 * np.cos(ref))).max() < 1e-10


def test_host_phase_matches_oracle(oracle):
    import glsu.synth as synth
    phi = oracle.gen_phase(64, 48, 3.0, 0.05)
    assert np.abs(synth.phase(64, 48, 3.0, 0.05) - phi).max() < 1e-12


def test_wrapped_error_metric(oracle):
    a = np.array([0.0, np.pi - 0.01, 1.0])
    b = np.array([2 * np.pi, -np.pi + 0.01, 1.0])
    rms, mx = oracle.wrapped_err(a, b)
    assert abs(mx - 0.02) < 1e-12
