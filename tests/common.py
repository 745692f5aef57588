"""Test inputs and fixture access.

The sinogram generators restate the reference test helpers
(tests/common_methods.py:11-118 of ODTbrain) so that the committed golden
fixtures (copied data files of the reference's tests/data) can be
reproduced from the same inputs.  ``cutout`` keeps the reference's 3D mask
exactly, including its reshape typo (common_methods.py:200), because the
fixtures were written through it.
"""
import pathlib
import zipfile

import numpy as np

GOLDEN = pathlib.Path(__file__).parent / "golden"


def normalize(av, vmin=0., vmax=1.):
    """common_methods.py:246-257."""
    if vmin == vmax:
        return np.ones_like(av) * vmin
    if vmax < vmin:
        vmin, vmax = vmax, vmin
    one = (av - np.min(av)) / (np.max(av) - np.min(av))
    return one * (vmax - vmin) + vmin


def create_test_sino_2d(A=9, N=22, max_phase=5.0, ampl_range=(1.0, 1.0)):
    """Off-centre Gaussian phase object, 2D (common_methods.py:11-60)."""
    out = np.zeros((A, N), dtype=np.complex128)
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    x = np.linspace(-N / 2, N / 2, N, endpoint=True)
    dev = np.sqrt(N / 2)
    off = N / 7
    for i in range(A):
        x0 = np.cos(angles[i]) * off
        phase = normalize(np.exp(-(x - x0) ** 2 / dev ** 2), vmax=max_phase)
        if ampl_range[0] == ampl_range[1]:
            ampl = ampl_range[0]
        else:
            adev = dev / 5
            aoff = off * .3
            ampl = np.exp(-(x - x0 - aoff) ** 2 / adev ** 2) + \
                np.exp(-(x - x0 + aoff) ** 2 / adev ** 2)
            ampl = normalize(ampl, vmin=ampl_range[0], vmax=ampl_range[1])
        out[i] = ampl * np.exp(1j * phase)
    return out, angles


def create_test_sino_3d(A=9, Nx=22, Ny=22, max_phase=5.0,
                        ampl_range=(1.0, 1.0)):
    """Off-centre Gaussian phase object, 3D (common_methods.py:63-118)."""
    out = np.zeros((A, Ny, Nx), dtype=np.complex128)
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    x = np.linspace(-Nx / 2, Nx / 2, Nx, endpoint=True).reshape(1, -1)
    y = np.linspace(-Ny / 2, Ny / 2, Ny, endpoint=True).reshape(-1, 1)
    dev = min(np.sqrt(Nx / 2), np.sqrt(Ny / 2))
    off = Nx / 7
    for i in range(A):
        x0 = np.cos(angles[i]) * off
        phase = np.exp(-(x - x0) ** 2 / dev ** 2) * np.exp(-y ** 2 / dev ** 2)
        phase = normalize(phase, vmax=max_phase)
        if ampl_range[0] == ampl_range[1]:
            ampl = ampl_range[0]
        else:
            adev = dev / 5
            aoff = off * .3
            ampl = np.exp(-(x - x0 - aoff) ** 2 / adev ** 2) + \
                np.exp(-(x - x0 + aoff) ** 2 / adev ** 2)
            ampl = normalize(ampl, vmin=ampl_range[0], vmax=ampl_range[1])
        out[i] = ampl * np.exp(1j * phase)
    return out, angles


def create_test_sino_3d_tilted(A=9, Nx=22, Ny=22, max_phase=5.0,
                               tilt_plane=0.0):
    """In-plane rotated variant (common_methods.py:121-184), constant amplitude."""
    from scipy.ndimage import rotate
    out = np.zeros((A, Ny, Nx), dtype=np.complex128)
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    x = np.linspace(-Nx / 2, Nx / 2, Nx, endpoint=True).reshape(1, -1)
    y = np.linspace(-Ny / 2, Ny / 2, Ny, endpoint=True).reshape(-1, 1)
    dev = min(np.sqrt(Nx / 2), np.sqrt(Ny / 2))
    off = Nx / 7
    for i in range(A):
        x0 = np.cos(angles[i]) * off
        phase = np.exp(-(x - x0) ** 2 / dev ** 2) * np.exp(-y ** 2 / dev ** 2)
        phase = normalize(phase, vmax=max_phase)
        ampl = np.ones((Nx, Ny))
        ampl = rotate(ampl, np.rad2deg(tilt_plane), reshape=False, cval=1)
        phase = rotate(phase, np.rad2deg(tilt_plane), reshape=False, cval=0)
        out[i] = ampl * np.exp(1j * phase)
    return out, angles


def cutout(a):
    """Circle/sphere mask (common_methods.py:187-205), 3D typo kept."""
    x = np.arange(a.shape[0])
    c = a.shape[0] / 2
    if len(a.shape) == 2:
        x = x.reshape(-1, 1)
        y = x.reshape(1, -1)
        zero = ((x - c) ** 2 + (y - c) ** 2) < c ** 2
    elif len(a.shape) == 3:
        x = x.reshape(-1, 1, 1)
        y = x.reshape(1, -1, 1)
        z = x.reshape(1, -1, 1)
        zero = ((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) < c ** 2
    else:
        raise ValueError("Cutout array must have dimension 2 or 3!")
    a *= zero
    return a


def get_test_parameter_set(set_number=1):
    """common_methods.py:231-243."""
    res, lD, nm = 2.1, 0, 1.333
    params = []
    for _ in range(set_number):
        params.append({"res": res, "lD": lD, "nm": nm})
        res += .1
        lD += np.pi
        nm *= 1.01
    return params


def load_golden(name):
    """Values of a golden fixture (zip holding data.txt, %.10f floats)."""
    with zipfile.ZipFile(GOLDEN / (name + ".zip")) as arc:
        with arc.open("data.txt") as fh:
            return np.loadtxt(fh)


def rel_l2(x, ref):
    x = np.asarray(x, dtype=np.complex128).ravel()
    ref = np.asarray(ref, dtype=np.complex128).ravel()
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


ROOT_DIR = pathlib.Path(__file__).resolve().parent.parent


def quarter_turn_case():
    """Sinogram without mirror symmetry in y (random phase on the reference's
    Gaussian object) and angles whose differences avoid exact quadrants: at
    multiples of 90 degrees the rotated coordinates of boundary voxels sit
    exactly on the plane edge, where rounding decides scipy's out-of-bounds
    test."""
    sino, angles = create_test_sino_3d(A=12, Nx=18, Ny=18)
    rng = np.random.default_rng(0)
    sino = sino * np.exp(1j * 0.3 * rng.standard_normal(sino.shape))
    # the tilted geometry uses angles relative to the first one: jitter
    # every angle so that no difference is a whole quadrant
    return sino, angles + rng.uniform(0.05, 0.2, angles.shape)
