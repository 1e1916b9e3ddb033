"""C-ABI checks that need no GPU: libglsu.so loads, exports every entry point
include/glsu.h declares, and reports argument / padding errors the way the
reference constructors do (status codes instead of `throw -1`)."""
import ctypes
import os
import re
import subprocess

import pytest

import glsu
from glsu import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "glsu.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(glsu_\w+)\s*\(", txt)))


def test_header_declarations_match_binding():
    assert declared() == sorted(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (glsu_\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    L = glsu.lib()
    for s in declared():
        assert hasattr(L, s)


def test_version_and_plans():
    assert b"gfx950" in glsu.lib().glsu_version()
    plans = glsu.fft_plans()
    assert all(p["N"] == p["T"] * p["E"] for p in plans)
    assert any(p["N"] >= 1024 + 30 for p in plans)
    assert any(p["N"] >= 2048 + 36 for p in plans)
    # the first cross-lane stage goes through permlane swaps exactly for the
    # 64-lane plans whose first radix is 4 (the radix-8 and narrower plans use LDS)
    by_id = {p["id"]: p for p in plans}
    assert by_id[4]["exchange"] == 1 and by_id[4]["N"] == 1280
    assert by_id[1]["exchange"] == 0  # 64 x 8, radices 8 x 8
    assert all(p["exchange"] == 0 for p in plans if p["T"] != 64)
    assert glsu.lib().glsu_fft_plan_exchange(999) == glsu._lib.ERR_ARG


def test_param_helpers():
    p = glsu.wft2_default_params()
    assert (p.sigma_x, p.wxl, p.wxh, p.wxi, p.thr) == (10.0, -2.3, 2.3, 0.1, -1.0)
    q = glsu.dpra_reference_params()
    assert q.sigma_x == 20 and q.thr == 15 and q.wxl == float(ctypes.c_float(-0.15).value)


def _create_wft2(W, H, typ=0, params=None):
    h = ctypes.c_void_p()
    p = params or glsu.wff_params()
    return glsu.lib().glsu_wft2_create(ctypes.byref(h), W, H, typ, ctypes.byref(p), None), h


def test_padding_error_code():
    # W + 2s > 4096: the reference throws int -1 (WFT/WFT2_CPU.cpp:69-73)
    rc, h = _create_wft2(4096 - 60 + 1, 32)
    assert rc == _lib.ERR_PADDING
    assert b"4096" in glsu.lib().glsu_last_error()


def test_argument_errors():
    assert _create_wft2(0, 10)[0] == _lib.ERR_ARG
    # WFR mode: the curvature padding rule (W + 3 sx <= 4096, WFT/WFT2_CPU.cpp:848-858)
    # fails where WFF mode would pass, and a window beyond the curvature tiles
    assert _create_wft2(4020, 16, typ=1)[0] == _lib.ERR_PADDING
    assert _create_wft2(10, 10, typ=1, params=glsu.wff_params(sigma_x=40.0))[0] == _lib.ERR_UNSUPPORTED
    assert _create_wft2(10, 10, typ=7)[0] == _lib.ERR_ARG
    assert _create_wft2(10, 10, params=glsu.wff_params(wxi=0.0))[0] == _lib.ERR_ARG
    h = ctypes.c_void_p()
    phi = (ctypes.c_double * 4)()
    rc = glsu.lib().glsu_dpra_create(ctypes.byref(h), phi, 2, 2, 0, None, None)
    assert rc == _lib.ERR_ARG                                            # rr must be >= 1
    assert glsu.lib().glsu_dpra_frame(None, None, 1, None, None, None) == _lib.ERR_ARG


def test_tuning_errors():
    # tuning requests are checked on the host, before any device work
    p = glsu.wff_params()
    h = ctypes.c_void_p()

    def create(**kw):
        t = _lib.tuning_arg(kw)
        return glsu.lib().glsu_wft2_create_tuned(ctypes.byref(h), 200, 150, 0, ctypes.byref(p),
                                                 None, ctypes.byref(t))
    assert create(cols_waves=3) == _lib.ERR_ARG
    assert create(rows_waves=5) == _lib.ERR_ARG
    assert create(rows_waves=0) == _lib.ERR_ARG
    assert create(ls_tile=7) == _lib.ERR_ARG            # unknown least-squares tile shape
    plans = {q["id"]: q for q in glsu.fft_plans()}
    narrow = [i for i, q in plans.items() if q["T"] != 64 and q["N"] >= 200 + 30]
    short = [i for i, q in plans.items() if q["N"] < 150]
    assert create(planx=narrow[0]) == _lib.ERR_ARG      # x axis needs a T == 64 plan
    assert create(plany=short[0]) == _lib.ERR_ARG       # too short for H + half window
    assert create(planx=999) == _lib.ERR_ARG


def test_autotune_candidates():
    from glsu import autotune
    plans = {q["id"]: q for q in glsu.fft_plans()}
    c = autotune.candidates(1054, True)
    assert c and all(plans[i]["T"] == 64 and plans[i]["N"] >= 1054 for i in c)
    assert len({plans[i]["N"] for i in c}) <= 3
    assert plans[c[0]]["N"] == min(q["N"] for q in plans.values()
                                   if q["T"] == 64 and q["N"] >= 1054)
    c2 = autotune.candidates(300, False)
    assert any(plans[i]["T"] != 64 for i in c2)


def test_no_device_is_reported_not_faked():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    n = ctypes.c_int(-1)
    glsu.lib().glsu_device_count(ctypes.byref(n))
    assert n.value == 0
    rc, _ = _create_wft2(64, 64)
    assert rc == _lib.ERR_DEVICE


def test_cpp_wrapper_compiles():
    # include/glsu.hpp (reference-style C++ classes) compiles against the ABI
    for name in ("example_dpra.cpp", "example_classes.cpp"):
        src = os.path.join(ROOT, "g-ls3u_amd", "csrc", name)
        r = subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-fsyntax-only", src],
                           capture_output=True, text=True)
        assert r.returncode == 0, name + ": " + r.stderr[-2000:]


def test_output_tensors_are_validated_before_any_launch():
    """Caller-supplied outputs go to the C ABI as raw pointers; a tensor of the
    wrong shape, dtype, device or layout must be refused in Python first."""
    torch = pytest.importorskip("torch")
    chk = _lib.check_out
    with pytest.raises(ValueError, match="CUDA"):
        chk(torch.empty((3, 4, 5), dtype=torch.float32), (3, 4, 5), torch.float32)
    with pytest.raises(ValueError, match="CUDA"):
        chk(None, (3, 4, 5), torch.float32)

    class CudaLike:
        """Metadata of a CUDA tensor (the checks read nothing else)."""
        is_cuda = True

        def __init__(self, t):
            self.t, self.dtype, self.shape = t, t.dtype, t.shape

        def is_contiguous(self):
            return self.t.is_contiguous()

    good = CudaLike(torch.empty((3, 4, 5), dtype=torch.float32))
    assert chk(good, (3, 4, 5), torch.float32) is good
    with pytest.raises(ValueError, match="shape"):
        chk(good, (3, 4, 6), torch.float32)
    with pytest.raises(ValueError, match="float64"):
        chk(CudaLike(torch.empty((3, 4, 5), dtype=torch.float64)), (3, 4, 5), torch.float32)
    with pytest.raises(ValueError, match="contiguous"):
        chk(CudaLike(torch.empty((5, 4, 3)).transpose(0, 2)), (3, 4, 5), torch.float32)
