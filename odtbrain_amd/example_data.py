"""Loaders of the reference's example data archives (examples/example_helper.py:
67-143): FDTD / QLSI sinograms as ``.tar.lzma`` archives of text arrays and
Mie / FDTD sinograms as ``.zip`` archives, so that those realistic inputs can
feed the reconstruction offline.  Local files only: where the reference
downloads a missing archive (example_helper.py:44-64) these raise
FileNotFoundError.

Host-side file parsing (numpy); the arrays it returns go to the GPU path
like any other input.
"""
import io
import lzma
import pathlib
import tarfile
import zipfile

import numpy as np


def _path(fname):
    p = pathlib.Path(fname)
    if not p.exists():
        raise FileNotFoundError(f"example data file {p} not found (offline "
                                "loader: the archive must be present locally)")
    return p


def _parse_info(lines):
    """``key = value`` lines with a float value (example_helper.py:96-100,
    136-141); other lines are ignored."""
    cfg = {}
    for ln in lines:
        ln = ln.decode() if isinstance(ln, bytes) else ln
        if ln.count("=") == 1:
            key, val = ln.split("=")
            cfg[key.strip()] = float(val.strip())
    return cfg


def load_tar_lzma_data(tlfile):
    """Sinogram, angles, phantom and parameters of a ``.tar.lzma`` archive
    (example_helper.py:76-116).

    Members are taken in sorted name order: ``fdtd_info*`` holds the
    parameters, ``phantom*`` the phantom slices, ``field*real.txt`` and
    ``field*imag.txt`` one projection each.  Returns ``(sino, angles,
    phantom, parms)`` with ``sino`` complex (A, ...) and ``angles`` A values
    uniform on [0, 2 pi) (``np.linspace(0, 2 pi, A, endpoint=False)``).
    """
    path = _path(tlfile)
    with open(path, "rb") as fh:
        data = lzma.decompress(fh.read())
    fields_real, fields_imag, phantom, parms = [], [], [], {}
    with tarfile.open(fileobj=io.BytesIO(data), mode="r:") as t:
        members = sorted(t.getmembers(), key=lambda m: m.name)
        for m in members:
            n = m.name
            f = t.extractfile(m)
            if f is None:
                continue
            if n.startswith("fdtd_info"):
                parms.update(_parse_info(f.readlines()))
            elif n.startswith("phantom"):
                phantom.append(np.loadtxt(f))
            elif n.startswith("field"):
                if n.endswith("imag.txt"):
                    fields_imag.append(np.loadtxt(f))
                elif n.endswith("real.txt"):
                    fields_real.append(np.loadtxt(f))
    phantom = np.array(phantom)
    sino = np.array(fields_real) + 1j * np.array(fields_imag)
    angles = np.linspace(0, 2 * np.pi, sino.shape[0], endpoint=False)
    return sino, angles, phantom, parms


def load_zip_data(zipname, f_sino_real, f_sino_imag, f_angles=None,
                  f_phantom=None, f_info=None):
    """Arrays of a ``.zip`` archive by member name (example_helper.py:
    119-143): ``[sino, angles?, phantom?, info?]`` in that order, each
    optional entry only when its member name is given."""
    ret = []
    with zipfile.ZipFile(str(_path(zipname))) as arc:
        sino = np.loadtxt(arc.open(f_sino_real)) + \
            1j * np.loadtxt(arc.open(f_sino_imag))
        ret.append(sino)
        if f_angles:
            ret.append(np.loadtxt(arc.open(f_angles)))
        if f_phantom:
            ret.append(np.loadtxt(arc.open(f_phantom)))
        if f_info:
            with arc.open(f_info) as info:
                ret.append(_parse_info(info.readlines()))
    return ret


def load_data(fname, **kwargs):
    """Dispatch on the suffix as example_helper.load_data (:67-73):
    ``.lzma`` -> load_tar_lzma_data, ``.zip`` -> load_zip_data(**kwargs);
    None for any other suffix, as there."""
    p = _path(fname)
    if p.suffix == ".lzma":
        return load_tar_lzma_data(p)
    if p.suffix == ".zip":
        return load_zip_data(p, **kwargs)
    return None
