"""3D backpropagation about the y axis on the GPU (drop-in for
odtbrain.backpropagate_3d, odtbrain/_alg3d_bpp.py:84-555)."""
import multiprocessing as mp

import numpy as np

from . import _core
from . import _native as N
from . import util

ncores = mp.cpu_count()


def _precision(dtype):
    """dtype rule of _alg3d_bpp.py:254-260 -> (real dtype, complex dtype)."""
    if dtype is None:
        dtype = np.float64
    dtype = np.dtype(dtype)
    if dtype.name not in ["float32", "float64"]:
        raise ValueError("dtype must be float32 or float64!")
    cdt = np.dtype(np.complex64 if dtype == np.float32 else np.complex128)
    return dtype, cdt


def _pads(lny, lnx, padding, padfac):
    """Pad widths of _alg3d_bpp.py:285-304."""
    padx = _core.padded_order(lnx, padfac) - lnx if padding[0] else 0
    pady = _core.padded_order(lny, padfac) - lny if padding[1] else 0
    padyl = int(np.ceil(pady / 2))
    padxl = int(np.ceil(padx / 2))
    return padyl, padxl, lny + pady, lnx + padx


def _pad_mode(padding, padval, lny, lnx, lNy, lNx):
    """Padding mode of the projection load (_alg3d_bpp.py:493-506)."""
    if not (padding[0] and padding[1]):
        if (lNy, lNx) != (lny, lnx):
            # the reference assigns the unpadded projection into the padded
            # buffer, which numpy refuses
            raise ValueError(
                "could not broadcast input array from shape ({},{}) into "
                "shape ({},{})".format(lny, lnx, lNy, lNx))
        return N.ODT_PAD_NONE, 0.0, 0.0
    if isinstance(padval, str) and padval == "edge":
        return N.ODT_PAD_EDGE, 0.0, 0.0
    pv = complex(padval)
    return N.ODT_PAD_RAMP, pv.real, pv.imag


def _validate_common(uSin, A, padding, coords, num_cores, limit):
    if len(uSin.shape) != 3:
        raise ValueError("Input data `uSin` must have shape (A,Ny,Nx).")
    if len(uSin) != A:
        raise ValueError("`len(angles)` must be  equal to `len(uSin)`.")
    if len(list(padding)) != 2:
        raise ValueError("`padding` must be boolean tuple of length 2!")
    if np.array(padding).dtype is not np.dtype(bool):
        raise ValueError("Parameter `padding` must be boolean tuple.")
    if coords is not None:
        raise NotImplementedError("Setting coordinates is not yet supported.")
    if num_cores > limit:
        raise ValueError("`num_cores` must not exceed number "
                         + "of physical cores: {}".format(limit))


def _as_input(uSin, copy):
    """numpy copy semantics of ``np.array(uSin, copy=copy)`` (:272) for
    host arrays; device tensors pass through untouched."""
    try:
        import torch
        if isinstance(uSin, torch.Tensor):
            return uSin
    except ImportError:  # pragma: no cover - torch is a hard requirement
        pass
    return np.array(uSin, copy=copy)


def _plan_fields(angles, res, nm, lD, lny, lnx, padyl, padxl, lNy, lNx,
                 pad_mode, pv_re, pv_im, intp_order, onlyreal, prefilter,
                 precision, device):
    """Plan description and rotation parameters (host float64)."""
    A = angles.size
    ln = lnx
    km = (2 * np.pi * nm) / res
    # scipy.ndimage.rotate(arr, angle=-phi0, axes=(0, 2)) per y-plane
    # (_alg3d_bpp.py:73-81, :534)
    phi0 = np.rad2deg(np.asarray(angles, dtype=float).reshape(-1))
    params = _core.planar_params([-p for p in phi0], ln, lnx)
    fields = dict(
        kind=N.ODT_KIND_3D, precision=precision, A=A, Ny=lny, Nx=lnx,
        lNy=lNy, lNx=lNx, padyl=padyl, padxl=padxl, pad_mode=pad_mode,
        intp_order=int(intp_order), onlyreal=int(bool(onlyreal)),
        flags=N.ODT_FLAG_IIR_PREFILTER if prefilter == "iir" else 0,
        angle_param_stride=6, padval_re=pv_re, padval_im=pv_im,
        km=float(km), lD=float(lD), z0=-ln / 2.0, ramp_u=0.0, ramp_v=1.0)
    esz = (4 if precision == N.ODT_FP32 else 8) * (1 if onlyreal else 2)
    fields["slab_budget_bytes"] = _core.slab_budget(
        ln * lny * lnx * esz, A, device)
    return fields, params


def backpropagate_3d(uSin, angles, res, nm, lD=0, coords=None,
                     weight_angles=True, onlyreal=False,
                     padding=(True, True), padfac=1.75, padval="edge",
                     intp_order=2, dtype=None,
                     num_cores=ncores,
                     save_memory=False,
                     copy=True,
                     count=None, max_count=None,
                     verbose=0, *, prefilter="fused"):
    r"""3D backpropagation (GPU).

    Same parameters, defaults, validation, errors and return value as
    odtbrain.backpropagate_3d (odtbrain/_alg3d_bpp.py:84-236): ``uSin``
    (A, Ny, Nx) sinogram of u_B/u_0, ``angles`` (A,) in radians, ``res``
    wavelength in pixels, ``nm`` medium index, ``lD`` detector distance.
    Returns the object function f of shape (Nx, Ny, Nx): complex64/128 by
    ``dtype`` (float32/float64, default float64), real if ``onlyreal``.

    ``num_cores`` and ``save_memory`` are validated but have no effect: the
    work runs on the GPU and the propagator is never tabulated.  If ``uSin``
    is a torch tensor on the GPU the result is returned as a GPU tensor.

    Extension (keyword only): ``prefilter`` -- "fused" (default) folds the
    B-spline prefilter into the inverse transforms, "iir" runs the separable
    recursive prefilter kernel on the slabs (identical results up to
    rounding).
    """
    A = angles.size
    _validate_common(uSin, A, padding, coords, num_cores, ncores)
    dtype, cdt = _precision(dtype)
    precision = N.ODT_FP32 if dtype == np.float32 else N.ODT_FP64
    if max_count is not None:
        max_count.value += A + 2
    uSin = _as_input(uSin, copy)
    lny, lnx = uSin.shape[1], uSin.shape[2]
    ln = lnx
    padyl, padxl, lNy, lNx = _pads(lny, lnx, padding, padfac)
    if verbose > 0:
        print("......Image size (x,y): {}x{}, padded: {}x{}".format(
            lnx, lny, lNx, lNy))
    pad_mode, pv_re, pv_im = _pad_mode(padding, padval, lny, lnx, lNy, lNx)
    if weight_angles:
        weights = util.compute_angle_weights_1d(angles)
        if not copy and isinstance(uSin, np.ndarray):
            # the reference weights the caller's array in place (:311-313)
            uSin *= weights.reshape(-1, 1, 1)
            weights = np.ones(A)
    else:
        weights = np.ones(A)
    d_usin, was_tensor = _core.device_input(uSin, cdt)
    fields, params = _plan_fields(
        angles, res, nm, lD, lny, lnx, padyl, padxl, lNy, lNx, pad_mode,
        pv_re, pv_im, intp_order, onlyreal, prefilter, precision,
        d_usin.device)
    if count is not None:
        count.value += 2
    out = _core.run(fields, weights, params, d_usin, (ln, lny, lnx),
                    bool(onlyreal), precision, was_tensor)
    if count is not None:
        count.value += A
    return out
