"""odt_to_ri / opt_to_ri on the GPU against the reference's fixtures
(tests/test_processing.py:52-76) and the oracle."""
import numpy as np
import pytest

import odtbrain_amd as G
from oracle import translate as T

from common import load_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["odt_to_ri", "opt_to_ri"])
def test_translate_fixture(name):
    f, res, nm = T.processing_test_data()
    ri = getattr(G, name)(f=f, res=res, nm=nm)
    assert ri.dtype == np.float64 and ri.shape == f.shape
    assert np.allclose(np.array(ri).flatten().view(float),
                       load_golden("processing__test_" + name))
    # 2D input gives the same slice
    assert np.allclose(getattr(G, name)(f=f[0], res=res, nm=nm), ri[0])


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-15),
                                       (np.complex64, 2e-7)])
def test_odt_to_ri_complex_vs_oracle(dtype, tol):
    rng = np.random.default_rng(0)
    res, nm = 6.0, 1.333
    km = 2 * np.pi * nm / res
    shape = (33, 17, 9)
    f = (km ** 2 * rng.uniform(-2.0, 1.0, shape) +
         1j * km ** 2 * rng.uniform(-0.5, 0.5, shape)).astype(dtype)
    ri = G.odt_to_ri(f, res, nm)
    ref = T.odt_to_ri(f.astype(np.complex128), res, nm)
    assert ri.dtype == dtype
    assert np.all(ri.real >= 0)
    err = np.abs(ri - ref) / np.abs(ref)
    assert err.max() < 5 * tol, err.max()


def test_odt_to_ri_branch_cut_and_special_values():
    """Negative real radicands (f < -km^2) with +0, -0 and tiny imaginary
    parts take numpy's principal root (f / km^2 + 1 turns a -0 imaginary
    part into +0 first; a tiny negative one keeps its sign); zero radicand
    gives 0; real negative radicands give NaN."""
    res, nm = 6.0, 1.333
    km2 = (2 * np.pi * nm / res) ** 2
    vals = np.array([-3 * km2 + 0j, complex(-3 * km2, -0.0),
                     complex(-3 * km2, 1e-300), complex(-3 * km2, -1e-300),
                     -km2 + 0j, 0j, complex(5 * km2, -0.0)],
                    dtype=np.complex128)
    ri = G.odt_to_ri(vals, res, nm)
    with np.errstate(invalid="ignore"):
        ref = T.odt_to_ri(vals.copy(), res, nm)
    assert np.allclose(ri, ref, rtol=1e-15, atol=0)
    assert np.array_equal(np.signbit(ri.imag), np.signbit(ref.imag))
    real = np.array([-3 * km2, -km2, 0.0, 2.0 * km2])
    with np.errstate(invalid="ignore"):
        ref_r = T.odt_to_ri(real.copy(), res, nm)
    ri_r = G.odt_to_ri(real, res, nm)
    assert np.isnan(ri_r[0]) and np.isnan(ref_r[0])
    assert np.allclose(ri_r[1:], ref_r[1:], rtol=1e-15)


def test_translate_device_tensors_and_float32():
    import torch
    rng = np.random.default_rng(1)
    f = rng.uniform(0, 5, (4, 5, 6)).astype(np.float32)
    d = torch.from_numpy(f).cuda()
    for name in ("odt_to_ri", "opt_to_ri"):
        out = getattr(G, name)(d, 6.0, 1.333)
        assert isinstance(out, torch.Tensor) and out.is_cuda
        assert out.dtype == torch.float32
        ref = getattr(T, name)(f.astype(np.float64), 6.0, 1.333)
        assert np.allclose(out.cpu().numpy(), ref, rtol=3e-7, atol=0)
    c = (f + 1j * f[::-1]).astype(np.complex64)
    out = G.opt_to_ri(torch.from_numpy(c).cuda(), 6.0, 1.333).cpu().numpy()
    assert np.allclose(out, T.opt_to_ri(c.astype(np.complex128), 6.0, 1.333),
                       rtol=3e-7)
