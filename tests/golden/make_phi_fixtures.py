"""Writes tests/golden/phi_maps.npz: the reference's two measured phase maps
(Experiments/imgs/1/1_phi.bin, 256 x 256, and Experiments/imgs/2/2_phi.bin,
366 rows x 371 cols; binary matrix format of Utils/matrixIO.h:18-35 /
Utils/matrixIO.inl:33-62: int32 rows, int32 cols, float32 row-major).  The
files are parsed here with plain numpy, independently of the library's
reader, so the fixture can pin that reader.  Run in a checkout that has the
reference tree next to it:  python tests/golden/make_phi_fixtures.py"""
import os
import sys

import numpy as np

REF = os.environ.get("GLSU_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "phi_maps.npz")


def parse(path):
    raw = np.fromfile(path, dtype=np.uint8)
    rows, cols = np.frombuffer(raw[:8].tobytes(), dtype="<i4")
    data = np.frombuffer(raw[8:].tobytes(), dtype="<f4")
    assert data.size == rows * cols, (path, rows, cols, data.size)
    return data.reshape(rows, cols)


def main():
    maps = {}
    for k in (1, 2):
        p = os.path.join(REF, "Experiments", "imgs", str(k), f"{k}_phi.bin")
        maps[f"phi{k}"] = parse(p)
        print(p, maps[f"phi{k}"].shape, float(maps[f"phi{k}"].min()), float(maps[f"phi{k}"].max()))
    np.savez_compressed(OUT, **maps)
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
