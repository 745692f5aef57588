"""Write excerpts of ODTbrain's FDTD example sinograms as test fixtures.

The reference's example pipelines (examples/backprop_from_fdtd_3d.py:35-46,
examples/backprop_from_fdtd_3d_tilted.py:62-71) run on the archives
examples/data/fdtd_3d_sino_A180_R6.500.tar.lzma and
fdtd_3d_sino_A220_R6.500_tiltyz0.2.tar.lzma (complex fields, 188^2
detector).  These fixtures keep every 6th (resp. 10th) projection and the
central 96 x 96 detector pixels as complex64, with the archives' angles and
parameters, so that the GPU pipeline can be compared with the oracle's on
realistic data at test time.  Inputs only: the expected values come from
the oracle when the test runs.

    python tests/golden/make_example_excerpts.py [REFERENCE_ROOT]
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from odtbrain_amd import example_data  # noqa: E402

CROP = 96
CASES = [("fdtd_3d_sino_A180_R6.500.tar.lzma", 6, "example_fdtd_3d_excerpt.npz"),
         ("fdtd_3d_sino_A220_R6.500_tiltyz0.2.tar.lzma", 10,
          "example_fdtd_3d_tilted_excerpt.npz")]


def excerpt(path, step):
    sino, angles, _phantom, cfg = example_data.load_data(str(path))
    n = sino.shape[-1]
    c0 = (n - CROP) // 2
    part = sino[::step, c0:c0 + CROP, c0:c0 + CROP].astype(np.complex64)
    out = {"sino": np.ascontiguousarray(part),
           "angles": np.asarray(angles[::step], dtype=np.float64)}
    for key in ("res", "nm", "lD", "tilt_yz"):
        if key in cfg:
            out[key] = np.float64(cfg[key])
    return out


def main():
    root = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    for name, step, out in CASES:
        arrays = excerpt(root / "examples" / "data" / name, step)
        np.savez_compressed(HERE / out, **arrays)
        print(out, arrays["sino"].shape, {k: float(v) for k, v in arrays.items()
                                          if v.ndim == 0})


if __name__ == "__main__":
    main()
