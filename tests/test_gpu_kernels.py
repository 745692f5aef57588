"""GPU parity of the standalone kernels, called through the C ABI:
transforms vs numpy.fft, the IIR prefilter vs scipy.ndimage.spline_filter1d,
rotate-accumulate vs scipy.ndimage.rotate and the reference's mprotate
fixture (alg3d_bpp__test_3d_mprotate)."""
import ctypes

import numpy as np
import pytest
import scipy.ndimage as ndi

from common import load_golden
from odtbrain_amd import _core
from odtbrain_amd import _native as N

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("prec", [N.ODT_FP32, N.ODT_FP64])
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024,
                               2048, 4096,
                               # Bluestein (not powers of two, <= 2048):
                               # small primes, composites, the largest
                               3, 5, 7, 10, 12, 22, 97, 100, 250, 251, 1000,
                               1023, 2047,
                               # direct DFT beyond
                               2049])
def test_fft_c2c_matches_numpy(prec, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((7, n)) + 1j * rng.standard_normal((7, n))
    cdt = torch.complex64 if prec == N.ODT_FP32 else torch.complex128
    tol = 2e-6 * max(1, np.log2(n)) if prec == N.ODT_FP32 else 1e-13
    d_in = torch.from_numpy(x).to(cdt).cuda()
    for direction in (-1, 1):
        d_out = torch.empty_like(d_in)
        N.check(N.lib().odt_fft_c2c(prec, 7, n, direction, _ptr(d_in),
                                    _ptr(d_out), _stream()), "fft")
        torch.cuda.synchronize()
        ref = np.fft.fft(x, axis=-1) if direction < 0 else \
            np.fft.ifft(x, axis=-1, norm="forward")
        got = d_out.cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < tol, (n, direction, err)


@pytest.mark.parametrize("prec", [N.ODT_FP32, N.ODT_FP64])
@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_spline_prefilter_matches_scipy(prec, order):
    rng = np.random.default_rng(order)
    x = rng.standard_normal((5, 33))
    tdt = torch.float32 if prec == N.ODT_FP32 else torch.float64
    tol = 1e-5 if prec == N.ODT_FP32 else 1e-12
    for axis in (0, 1):
        d = torch.from_numpy(x).to(tdt).cuda()
        if axis == 1:   # lines are rows
            nl, n, ls, es = 5, 33, 33, 1
        else:           # lines are columns
            nl, n, ls, es = 33, 5, 1, 33
        N.check(N.lib().odt_spline_prefilter(prec, order, nl, n, ls, es,
                                             _ptr(d), _stream()), "prefilter")
        torch.cuda.synchronize()
        ref = ndi.spline_filter1d(x, order, axis=axis, mode="mirror")
        got = d.cpu().numpy()
        assert np.abs(got - ref).max() < tol * np.abs(ref).max()


def _rotate_gpu(coef, angles_deg, order, prec, accumulate=None):
    """coef (nang, ny, nz, nx) prefiltered -> sum of rotated planes."""
    nang, ny, nz, nx = coef.shape
    tdt = torch.float32 if prec == N.ODT_FP32 else torch.float64
    d_c = torch.from_numpy(np.ascontiguousarray(coef)).to(tdt).cuda()
    if accumulate is None:
        d_o = torch.empty((nz, ny, nx), dtype=tdt, device="cuda")
        acc = 0
    else:
        d_o = torch.from_numpy(accumulate).to(tdt).cuda()
        acc = 1
    params = np.ascontiguousarray(_core.planar_params(angles_deg, nz, nx))
    N.check(N.lib().odt_rotate_accumulate(prec, order, nang, ny, nz, nx,
                                          _ptr(d_c), N.dptr(params), _ptr(d_o),
                                          acc, _stream()), "rotate")
    torch.cuda.synchronize()
    return d_o.cpu().numpy()


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
def test_rotate_accumulate_matches_scipy_fp64(order):
    """Exact out-of-bounds set (including 90-degree multiples, where input
    coordinates land exactly on the border) and values to 1e-10."""
    rng = np.random.default_rng(20 + order)
    nz, ny, nx = 17, 3, 17
    angles = [-2.0, 90.0, -270.0, 37.5, 180.0, -123.25]
    planes = rng.standard_normal((len(angles), ny, nz, nx))
    coef = np.empty_like(planes)
    ref = np.zeros((nz, ny, nx))
    for a, ang in enumerate(angles):
        for y in range(ny):
            p = planes[a, y]
            coef[a, y] = ndi.spline_filter(p, order, mode="mirror") \
                if order > 1 else p
            ref[:, y, :] += ndi.rotate(p, ang, reshape=False, order=order,
                                       mode="constant", cval=0)
    got = _rotate_gpu(coef, angles, order, N.ODT_FP64)
    assert np.allclose(got, ref, rtol=0, atol=1e-10)
    # single angles: the exact zero set of scipy's constant mode
    for a, ang in enumerate(angles[:3]):
        one = _rotate_gpu(coef[a:a + 1], [ang], order, N.ODT_FP64)
        r1 = np.stack([ndi.rotate(planes[a, y], ang, reshape=False,
                                  order=order, mode="constant", cval=0)
                       for y in range(ny)], axis=1)
        assert np.array_equal(one == 0, r1 == 0)


def test_rotate_accumulate_mprotate_fixture():
    """Reference fixture: arange(20^3) rotated by 2 degrees (angle -2) at
    order 2 for y in [0, 10) (tests/test_alg3d_bpp.py:116-134); the
    prefilter runs on the GPU too."""
    ln2 = 20
    arr = np.arange(ln2 ** 3, dtype=float).reshape(ln2, ln2, ln2)
    planes = np.ascontiguousarray(arr[:, :10, :].transpose(1, 0, 2))[None]
    d = torch.from_numpy(planes).cuda()
    # prefilter along x (rows of each plane) then z (columns)
    N.check(N.lib().odt_spline_prefilter(N.ODT_FP64, 2, 10 * ln2, ln2, ln2, 1,
                                         _ptr(d), _stream()), "pf x")
    for y in range(10):
        sub = d[0, y]
        N.check(N.lib().odt_spline_prefilter(N.ODT_FP64, 2, ln2, ln2, 1, ln2,
                                             _ptr(sub), _stream()), "pf z")
    torch.cuda.synchronize()
    got = _rotate_gpu(d.cpu().numpy(), [-2.0], 2, N.ODT_FP64)
    arr[:, :10, :] = got
    assert np.allclose(arr.flatten(), load_golden("alg3d_bpp__test_3d_mprotate"))


def test_rotate_accumulate_fp32_and_accumulate_mode():
    rng = np.random.default_rng(7)
    coef = rng.standard_normal((4, 2, 32, 32))
    angles = [10.0, -50.0, 200.0, 333.3]
    ref = _rotate_gpu(coef, angles, 2, N.ODT_FP64)
    got = _rotate_gpu(coef, angles, 2, N.ODT_FP32)
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()
    base = rng.standard_normal((32, 2, 32))
    acc = _rotate_gpu(coef, angles, 2, N.ODT_FP64, accumulate=base)
    assert np.allclose(acc, base + ref, rtol=0, atol=1e-12)
