"""The device FFT's in-register DFTs (g-ls3u_amd/csrc/fft_device.h, Dft<N>)
compiled for the host and checked against a direct FP64 DFT (CPU only).
The cross-lane stages need a wavefront and are covered on the GPU by
test_every_fft_plan_matches_oracle."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "g-ls3u_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_in_register_dfts_match_direct_dft(tmp_path):
    exe = tmp_path / "fft_dft_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", CSRC, "-o", str(exe),
                    os.path.join(HERE, "fft_dft_check.cpp")], check=True,
                   capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
