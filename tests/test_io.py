"""The reference's matrix file formats through libglsu (no GPU needed):
binary matrices (Utils/matrixIO.h:18-35, Utils/matrixIO.inl:33-104) and
complex text (Utils/Utils.cpp:80-193), pinned against byte layouts built
independently with numpy and against the reference's own phase maps
(fixture tests/golden/phi_maps.npz, made by make_phi_fixtures.py)."""
import os
import struct

import numpy as np
import pytest

import glsu
from glsu import io as gio

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "phi_maps.npz")
REF = "/root/reference/Experiments/imgs"


def _numpy_bin(path, a):
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", *a.shape))
        f.write(np.ascontiguousarray(a).tobytes())


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8, np.int32])
def test_bin_round_trip_and_layout(tmp_path, dtype):
    rng = np.random.default_rng(1)
    a = (rng.normal(size=(7, 5)) * 50).astype(dtype)
    p1, p2 = str(tmp_path / "a.bin"), str(tmp_path / "b.bin")
    gio.write_bin(p1, a)
    _numpy_bin(p2, a)
    assert open(p1, "rb").read() == open(p2, "rb").read()     # same bytes as the format
    b = gio.read_bin(p2, dtype)
    assert b.dtype == np.dtype(dtype) and np.array_equal(a, b)


def test_bin_errors(tmp_path):
    with pytest.raises(glsu.GlsuError):
        gio.read_bin(str(tmp_path / "missing.bin"))
    p = str(tmp_path / "short.bin")
    with open(p, "wb") as f:
        f.write(struct.pack("<ii", 4, 4))
        f.write(np.zeros(10, np.float32).tobytes())            # 10 of 16 elements
    with pytest.raises(glsu.GlsuError):
        gio.read_bin(p)


def test_reference_phase_maps_read_exactly(tmp_path):
    g = np.load(GOLDEN)
    for k, shape in (("phi1", (256, 256)), ("phi2", (366, 371))):
        a = g[k]
        assert a.shape == shape and a.dtype == np.float32
        assert -np.pi <= a.min() and a.max() <= np.pi            # wrapped phase maps
        p = str(tmp_path / f"{k}.bin")
        _numpy_bin(p, a)
        assert np.array_equal(gio.read_bin(p), a)
        n = k[-1]
        ref = os.path.join(REF, n, f"{n}_phi.bin")
        if os.path.exists(ref):                                   # the original file itself
            assert np.array_equal(gio.read_bin(ref), a)


def test_complex_text_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(2)
    a = rng.normal(size=(3, 4)) + 1j * rng.normal(size=(3, 4))
    p = str(tmp_path / "a.fp")
    gio.write_fp(p, a, precision=17)
    lines = open(p).read().splitlines()
    assert lines[0] == "3,4" and len(lines) == 13
    re, im = (float(v) for v in lines[1].split(","))
    assert re == a[0, 0].real and im == a[0, 0].imag            # row-major, re,im per line
    assert np.array_equal(gio.read_fp(p), a)
    # the reference's stream default: 6 significant digits
    gio.write_fp(p, a)
    b = gio.read_fp(p)
    assert np.abs(b - a).max() <= 5e-6 * np.abs(a).max() and not np.array_equal(a, b)
    with open(p, "w") as f:
        f.write("2,2\n1,2\n3,4\n")                               # too few elements
    with pytest.raises(glsu.GlsuError):
        gio.read_fp(p)
