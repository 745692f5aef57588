"""The offline example-data loaders (examples/example_helper.py:67-143) on
archives written here in the same formats: member order, parameter lines,
the angle grid, the return order of the zip loader, and the offline error."""
import io
import lzma
import tarfile
import zipfile

import numpy as np
import pytest

from odtbrain_amd import example_data as E


def _txt(a):
    b = io.StringIO()
    np.savetxt(b, a)
    return b.getvalue().encode()


def _add(t, name, data):
    info = tarfile.TarInfo(name)
    info.size = len(data)
    t.addfile(info, io.BytesIO(data))


def test_tar_lzma_archive(tmp_path):
    rng = np.random.default_rng(0)
    A, N = 5, 6
    re = rng.standard_normal((A, N, N))
    im = rng.standard_normal((A, N, N))
    ph = rng.standard_normal((3, N, N))
    raw = io.BytesIO()
    with tarfile.open(fileobj=raw, mode="w") as t:
        # written out of order: the loader sorts the members by name
        for a in reversed(range(A)):
            _add(t, f"field_{a:03d}_imag.txt", _txt(im[a]))
            _add(t, f"field_{a:03d}_real.txt", _txt(re[a]))
        for k in reversed(range(3)):
            _add(t, f"phantom_{k:03d}.txt", _txt(ph[k]))
        _add(t, "fdtd_info.txt",
             b"res = 6.5\nnm = 1.333\nlD = 0\ncomment line\na = b = c\n")
    p = tmp_path / "sino.tar.lzma"
    p.write_bytes(lzma.compress(raw.getvalue(), format=lzma.FORMAT_ALONE))
    sino, angles, phantom, parms = E.load_data(p)
    assert sino.shape == (A, N, N) and np.iscomplexobj(sino)
    assert np.allclose(sino, re + 1j * im, rtol=0, atol=1e-15)
    assert np.allclose(phantom, ph, rtol=0, atol=1e-15)
    assert parms == {"res": 6.5, "nm": 1.333, "lD": 0.0}
    assert np.array_equal(angles, np.linspace(0, 2 * np.pi, A, endpoint=False))


def test_zip_archive_return_order(tmp_path):
    rng = np.random.default_rng(1)
    s = rng.standard_normal((2, 4, 7))
    ang = np.linspace(0, np.pi, 4)
    ph = rng.standard_normal((7, 7))
    p = tmp_path / "sino.zip"
    with zipfile.ZipFile(p, "w") as z:
        z.writestr("re.txt", _txt(s[0]))
        z.writestr("im.txt", _txt(s[1]))
        z.writestr("ang.txt", _txt(ang))
        z.writestr("ph.txt", _txt(ph))
        z.writestr("info.txt", b"res = 13\nnm = 1.333\nlD = 0.5\n")
    sino, angles, phantom, cfg = E.load_data(
        p, f_sino_real="re.txt", f_sino_imag="im.txt", f_angles="ang.txt",
        f_phantom="ph.txt", f_info="info.txt")
    assert np.allclose(sino, s[0] + 1j * s[1], rtol=0, atol=1e-15)
    assert np.allclose(angles, ang, rtol=0, atol=1e-15)
    assert np.allclose(phantom, ph, rtol=0, atol=1e-15)
    assert cfg == {"res": 13.0, "nm": 1.333, "lD": 0.5}
    only = E.load_data(p, f_sino_real="re.txt", f_sino_imag="im.txt")
    assert len(only) == 1 and only[0].shape == (4, 7)


def test_offline_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        E.load_data(tmp_path / "missing.tar.lzma")
    other = tmp_path / "data.npy"
    other.write_bytes(b"")
    assert E.load_data(other) is None
