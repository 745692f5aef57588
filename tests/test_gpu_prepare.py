"""sinogram_as_rytov / sinogram_as_radon on the GPU: the reference's
fixtures and consistency checks (tests/test_processing.py:79-137), and the
oracle's restatement on noisy inputs (phase residues included)."""
import numpy as np
import pytest

import odtbrain_amd as G
from oracle import prepare as P

from common import load_golden

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi


def _mod_rest(a, b):
    q = np.array(a / b, dtype=int)
    r = a - b * q
    wrong = np.where(np.abs(r) > b / 2)
    r[wrong] -= b * np.sign(r[wrong])
    return r


def test_sino_radon_fixture_and_2d_consistency():
    sino = P.processing_test_sino()
    rad = G.sinogram_as_radon(sino)
    assert rad.dtype == np.float64 and rad.shape == sino.shape
    if rad.min() > TWOPI:
        rad -= TWOPI
    assert np.allclose(np.array(rad).flatten().view(float),
                       load_golden("processing__test_sino_radon"))
    for sl in ((slice(None), slice(None), 0), (slice(None), 0, slice(None))):
        r2 = G.sinogram_as_radon(sino[sl])
        assert np.allclose(0, _mod_rest(r2 - rad[sl], TWOPI), atol=1e-6)


def test_sino_rytov_fixture_and_2d_consistency():
    sino = P.processing_test_sino(rytov=True)
    ryt = G.sinogram_as_rytov(sino)
    assert ryt.dtype == np.complex128
    if ryt.imag.min() > TWOPI:
        ryt.imag -= TWOPI
    assert np.allclose(np.array(ryt).flatten().view(float),
                       load_golden("processing__test_sino_rytov"))
    for sl in ((slice(None), slice(None), 0), (slice(None), 0, slice(None))):
        r2 = G.sinogram_as_rytov(sino[sl])
        d = r2 - ryt[sl]
        assert np.allclose(0, (d.real + 1j * _mod_rest(d.imag, TWOPI))
                           .view(float), atol=1e-6)


def _noisy_field(shape, seed, turns=4.0, noise=0.6):
    """Complex field whose phase spans several turns, with phase noise
    large enough to create residues."""
    rng = np.random.default_rng(seed)
    A = shape[0]
    grids = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape[1:]],
                        indexing="ij")
    r2 = sum(g ** 2 for g in grids)
    phase = np.empty(shape)
    for a in range(A):
        phase[a] = TWOPI * turns * np.exp(-r2 / 0.4) * (1 + 0.1 * a) + \
            rng.normal(0, noise, shape[1:])
    amp = rng.uniform(0.5, 1.5, shape)
    return amp * np.exp(1j * phase)


def _residues(phase):
    """Number of 2x2 loops of a wrapped image whose wrapped differences do
    not sum to zero."""
    w = lambda d: (d + np.pi) % TWOPI - np.pi  # noqa: E731
    dx = w(np.diff(phase, axis=1))
    dy = w(np.diff(phase, axis=0))
    loop = dx[:-1, :] + dy[:, 1:] - dx[1:, :] - dy[:, :-1]
    return int(np.sum(np.abs(loop) > 1e-6))


@pytest.mark.parametrize("shape", [(3, 12, 14), (2, 25, 9), (4, 30)])
@pytest.mark.parametrize("align", [True, False])
def test_prepare_matches_oracle_with_residues(shape, align):
    u = _noisy_field(shape, seed=len(shape) + shape[1])
    if len(shape) == 3:
        assert sum(_residues(np.angle(im)) for im in u) > 0
    ref = P.sinogram_as_rytov(u, align=align)
    got = G.sinogram_as_rytov(u, align=align)
    assert np.allclose(got, ref, rtol=0, atol=1e-12)
    ref_r = P.sinogram_as_radon(u, align=align)
    got_r = G.sinogram_as_radon(u, align=align)
    assert np.allclose(got_r, ref_r, rtol=0, atol=1e-12)


@pytest.mark.parametrize("shape", [(4, 256, 256), (3, 97, 130), (1, 255, 257),
                                   (2, 2, 50), (2, 50, 3), (2, 2, 3)])
@pytest.mark.parametrize("align", [True, False])
def test_unwrap_lds_merge_equals_global_merge(shape, align, monkeypatch):
    """Images of at most 65536 pixels merge with their group sizes in LDS,
    64 root finds at a time; the global union-find (one thread per image)
    applies the same merges in the same order: the same output bits, on
    noisy fields full of residues."""
    u = _noisy_field(shape, seed=11, noise=1.2)
    if shape[1] * shape[2] >= 10000:
        assert _residues(np.angle(u[0])) > 500
    got = G.sinogram_as_rytov(u, align=align)
    monkeypatch.setenv("ODTBRAIN_AMD_UW_GLOBAL", "1")
    ref = G.sinogram_as_rytov(u, align=align)
    assert np.array_equal(got.view(float), ref.view(float))
    got_r = G.sinogram_as_radon(u[:2], align=align)
    monkeypatch.delenv("ODTBRAIN_AMD_UW_GLOBAL")
    assert np.array_equal(got_r, G.sinogram_as_radon(u[:2], align=align))


def test_prepare_u0_array_and_float32_and_tensors():
    import torch
    u = _noisy_field((3, 10, 12), seed=3, noise=0.2)
    rng = np.random.default_rng(4)
    u0 = (rng.uniform(0.8, 1.2, (10, 12)) *
          np.exp(1j * rng.uniform(-0.3, 0.3, (10, 12))))
    ref = P.sinogram_as_rytov(u, u0=u0)
    got = G.sinogram_as_rytov(u, u0=u0)
    assert np.allclose(got, ref, rtol=0, atol=1e-12)
    d = torch.from_numpy(u.astype(np.complex64)).cuda()
    g32 = G.sinogram_as_rytov(d)
    assert isinstance(g32, torch.Tensor) and g32.dtype == torch.complex64
    ref32 = P.sinogram_as_rytov(u.astype(np.complex64).astype(np.complex128))
    assert np.allclose(g32.cpu().numpy(), ref32, rtol=0, atol=2e-5)
