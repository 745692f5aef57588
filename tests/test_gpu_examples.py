"""The reference's 3D example pipelines on excerpts of its FDTD example data
(tests/golden/example_fdtd_3d*_excerpt.npz, written by
tests/golden/make_example_excerpts.py): sinogram_as_rytov ->
backpropagate_3d (examples/backprop_from_fdtd_3d.py:35-46) or
backpropagate_3d_tilted with the axis tilted by the archive's tilt_yz
(examples/backprop_from_fdtd_3d_tilted.py:62-71) -> odt_to_ri, on the GPU in
FP64 and FP32, against the same chain of the oracle in complex128.

Tolerances: object function FP64 relative L2 <= 1e-9, FP32 <= 1e-5; the
refractive index (its deviation from the medium, n - nm) likewise."""
import os

import numpy as np
import pytest

import odtbrain_amd as G
from common import GOLDEN, rel_l2
from oracle import odt_oracle as O
from oracle import prepare as P
from oracle import translate as T

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

WORKERS = max(1, min(16, os.cpu_count() or 1))


def _load(name):
    with np.load(GOLDEN / name, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name,tilted", [
    ("example_fdtd_3d_excerpt.npz", False),
    ("example_fdtd_3d_tilted_excerpt.npz", True)])
def test_example_pipeline_vs_oracle(name, tilted):
    d = _load(name)
    sino, angles = d["sino"], d["angles"]
    res, nm, lD = float(d["res"]), float(d["nm"]), float(d["lD"])
    kw = dict(lD=lD)
    if tilted:
        t = float(d["tilt_yz"])
        kw["tilted_axis"] = [0, np.cos(t), np.sin(t)]
    g_fn = G.backpropagate_3d_tilted if tilted else G.backpropagate_3d
    o_fn = O.backpropagate_3d_tilted if tilted else O.backpropagate_3d

    u = sino.astype(np.complex128)
    uR_ref = P.sinogram_as_rytov(u)
    f_ref = O.angle_parallel(o_fn, uR_ref, angles, res, nm, workers=WORKERS,
                             **kw)
    n_ref = T.odt_to_ri(f_ref, res, nm)

    for dt, ctype, tol in ((np.float64, torch.complex128, 1e-9),
                           (np.float32, torch.complex64, 1e-5)):
        d_u = torch.from_numpy(u).cuda().to(ctype)
        uR = G.sinogram_as_rytov(d_u)
        f = g_fn(uR, angles, res, nm, dtype=dt, **kw)
        n = G.odt_to_ri(f, res, nm)
        f, n = f.cpu().numpy(), n.cpu().numpy()
        ef = rel_l2(f, f_ref)
        en = rel_l2(n - nm, n_ref - nm)
        print(f"{name} {dt.__name__}: f rel_l2 {ef:.3e}, n - nm rel_l2 "
              f"{en:.3e}", flush=True)
        assert ef < tol
        assert en < tol
