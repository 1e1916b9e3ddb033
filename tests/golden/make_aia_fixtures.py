"""Writes tests/golden/aia_bmp.npz from data files of the reference's tests:
  frames   Test/1.bmp, Test/2.bmp, Test/3.bmp -- 256 x 256 8-bit palette
           bitmaps with a grey palette, read the way the reference's AIA tests
           do (cv::imread + cvtColor(BGR2GRAY), Test/AIA_CPU_Test.cpp:127-145,
           which for a grey palette is the palette value of each index);
  x_cpu    Test/Phi_CPU.csv -- 196608 values = 256*256 pixels x 3: the per-pixel
           solutions x = A1^-1 b of AIA's phase step, as AIA/aia_cpuf.cpp:169-178
           (commented) writes them, for the initial shifts of
           Test/AIA_CPU_Test.cpp:150-153.  Test/Phi_GPU.csv, written by
           AIA/aia_cudaf_v2.cu:578, holds the same solutions from the GPU class;
           the script checks that the two agree (to 1e-5 relative) and keeps
           the CPU one;
  deltas   those shifts {0.8622, 0.3188, -1.3077} (float literals).
Both files are parsed with plain numpy.  Run in a checkout with the reference
tree next to it:  python tests/golden/make_aia_fixtures.py"""
import os
import struct
import sys

import numpy as np

REF = os.environ.get("GLSU_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "aia_bmp.npz")


def read_bmp8(path):
    """Uncompressed 8-bit palette BMP (BITMAPINFOHEADER) -> grey uint8 image."""
    b = open(path, "rb").read()
    assert b[:2] == b"BM", path
    off = struct.unpack_from("<I", b, 10)[0]
    hsize, w, h = struct.unpack_from("<Iii", b, 14)
    bpp, comp = struct.unpack_from("<HI", b, 28)
    assert bpp == 8 and comp == 0, (path, bpp, comp)
    ncol = struct.unpack_from("<I", b, 46)[0] or 256
    pal = np.frombuffer(b, np.uint8, 4 * ncol, 14 + hsize).reshape(-1, 4)[:, :3]  # B, G, R
    assert np.all(pal[:, 0] == pal[:, 1]) and np.all(pal[:, 1] == pal[:, 2]), "not grey"
    stride = (w + 3) // 4 * 4
    idx = np.frombuffer(b, np.uint8, stride * abs(h), off).reshape(abs(h), stride)[:, :w]
    if h > 0:  # bottom-up rows
        idx = idx[::-1]
    return pal[idx, 0].copy()


def main():
    frames = np.stack([read_bmp8(os.path.join(REF, "Test", f"{k}.bmp")) for k in (1, 2, 3)])
    xs = {}
    for name in ("CPU", "GPU"):
        v = np.loadtxt(os.path.join(REF, "Test", f"Phi_{name}.csv"), dtype=np.float64)
        assert v.size == frames[0].size * 3, v.size
        xs[name] = v.reshape(-1, 3)
    rel = np.linalg.norm(xs["CPU"] - xs["GPU"]) / np.linalg.norm(xs["CPU"])
    assert rel < 1e-5, rel
    deltas = np.array([0.8622, 0.3188, -1.3077], np.float32)
    np.savez_compressed(OUT, frames=frames, deltas=deltas, x_cpu=xs["CPU"].astype(np.float32))
    print("wrote", OUT, frames.shape, f"(CPU vs GPU solutions: {rel:.1e} relative)")


if __name__ == "__main__":
    sys.exit(main())
