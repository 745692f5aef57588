"""2D backpropagation on the GPU (drop-in for odtbrain.backpropagate_2d,
odtbrain/_alg2d_bpp.py:8-326)."""
import numpy as np

from . import _core
from . import _native as N
from . import util


def backpropagate_2d(uSin, angles, res, nm, lD=0, coords=None,
                     weight_angles=True,
                     onlyreal=False, padding=True, padval=0,
                     count=None, max_count=None, verbose=0, *, dtype=None,
                     prefilter="fused"):
    r"""2D backpropagation with the Fourier diffraction theorem (GPU).

    Same parameters, defaults, validation, errors and return value as
    odtbrain.backpropagate_2d (odtbrain/_alg2d_bpp.py:8-131): ``uSin``
    (A, N), ``angles`` (A,) radians; padding to the power of two >= 2.1 N
    (at least 64), ``padval=None`` pads with edge values, a number pads with
    a linear ramp to it; rotation is the cubic B-spline of
    scipy.ndimage.rotate.  Returns (N, N) float64, or complex128 unless
    ``onlyreal``.

    Extensions (keyword only): ``dtype`` -- None/float64 (reference
    precision) or float32 (complex64 arithmetic, complex64/float32 result);
    ``prefilter`` as in backpropagate_3d.
    """
    A = angles.shape[0]
    if max_count is not None:
        max_count.value += A + 2
    assert len(uSin.shape) == 2, "Input data `uB` must have shape (A,N)!"
    assert len(uSin) == A, "`len(angles)` must be  equal to `len(uSin)`!"
    if coords is not None:
        # the reference's message expression (`+ + "..."`) raises TypeError
        raise TypeError("bad operand type for unary +: 'str'")
    if dtype is None:
        dtype = np.float64
    dtype = np.dtype(dtype)
    if dtype.name not in ["float32", "float64"]:
        raise ValueError("dtype must be float32 or float64!")
    precision = N.ODT_FP32 if dtype == np.float32 else N.ODT_FP64
    cdt = np.complex64 if precision == N.ODT_FP32 else np.complex128
    km = (2 * np.pi * nm) / res
    weights = util.compute_angle_weights_1d(angles) if weight_angles \
        else np.ones(A)
    ln = uSin.shape[1]
    order = max(64., 2 ** np.ceil(np.log(ln * 2.1) / np.log(2)))
    pad = order - ln if padding else 0
    padl = int(np.ceil(pad / 2))
    padr = int(pad - padl)
    lN = ln + padl + padr
    if padl + padr == 0:
        pad_mode, pv = N.ODT_PAD_NONE, 0j
    elif padval is None:
        pad_mode, pv = N.ODT_PAD_EDGE, 0j
        if verbose > 0:
            print("......Padding with edge values.")
    else:
        pad_mode, pv = N.ODT_PAD_RAMP, complex(padval)
        if verbose > 0:
            print("......Verifying padding value: {}".format(padval))
    # scipy.ndimage.rotate(sino, -angles[i]*180/np.pi, reshape=False),
    # default axes (1, 0) -> plane (rows, cols) (_alg2d_bpp.py:311-321)
    ang = np.asarray(angles, dtype=float).reshape(-1)
    params = _core.planar_params([-a * 180 / np.pi for a in ang], ln, ln)
    if count is not None:
        count.value += 2
    d_usin, was_tensor = _core.device_input(uSin, cdt)
    d_usin = d_usin.reshape(A, 1, ln)
    fields = dict(
        kind=N.ODT_KIND_2D, precision=precision, A=A, Ny=1, Nx=ln,
        lNy=1, lNx=lN, padyl=0, padxl=padl, pad_mode=pad_mode,
        intp_order=3, onlyreal=int(bool(onlyreal)),
        flags=N.ODT_FLAG_IIR_PREFILTER if prefilter == "iir" else 0,
        angle_param_stride=6, padval_re=pv.real, padval_im=pv.imag,
        km=float(km), lD=float(lD), z0=-ln / 2.0 + .5, ramp_u=0.0,
        ramp_v=1.0)
    esz = (4 if precision == N.ODT_FP32 else 8) * (1 if onlyreal else 2)
    fields["slab_budget_bytes"] = _core.slab_budget(ln * ln * esz, A,
                                                    d_usin.device)
    out = _core.run(fields, weights, params, d_usin, (ln, 1, ln),
                    bool(onlyreal), precision, was_tensor)
    if count is not None:
        count.value += A
    return out.reshape(ln, ln)
