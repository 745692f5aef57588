"""3D backpropagation with a tilted rotation axis on the GPU (drop-in for
odtbrain.backpropagate_3d_tilted, odtbrain/_alg3d_bppt.py:374-949)."""
import multiprocessing as mp
import warnings

import numpy as np

from . import _core
from . import _native as N
from . import util
from ._alg3d_bpp import _as_input, _pad_mode, _pads, _precision, \
    _validate_common
from ._geometry import norm_vec, rotate_points_to_axis, \
    rotation_matrix_from_point, rotation_matrix_from_point_planerot, \
    sphere_points_from_angles_and_tilt  # noqa: F401 (public helpers)

_ncores = mp.cpu_count()


def affine_params(points, angz, shape):
    """Per-angle inverse rotation and offset of the affine resampling,
    computed as in _alg3d_bppt.py:907-919: ``Dinv`` from
    rotation_matrix_from_point_planerot, ``c = 0.5*shape - 0.5``,
    ``offset = c - Dinv @ c``.  Rows: Dinv (row-major, 9) then offset (3)."""
    out = np.empty((len(points), 12), dtype=np.float64)
    c = 0.5 * np.array(shape) - .5
    for i, p in enumerate(points):
        _d, dinv = rotation_matrix_from_point_planerot(p, plane_angle=angz,
                                                       ret_inv=True)
        offset = c - np.dot(dinv, c)
        out[i, :9] = np.asarray(dinv, dtype=np.float64).reshape(-1)
        out[i, 9:] = offset
    return out


def backpropagate_3d_tilted(uSin, angles, res, nm, lD=0,
                            tilted_axis=[0, 1, 0],
                            coords=None, weight_angles=True, onlyreal=False,
                            padding=(True, True), padfac=1.75, padval="edge",
                            intp_order=2, dtype=None,
                            num_cores=_ncores,
                            save_memory=False,
                            copy=True,
                            count=None, max_count=None,
                            verbose=0, *, prefilter="fused"):
    r"""3D backpropagation with a tilted axis of rotation (GPU).

    Same parameters, defaults, validation, warnings, errors and return value
    as odtbrain.backpropagate_3d_tilted (odtbrain/_alg3d_bppt.py:374-561):
    ``angles`` is (A,) in radians (points on the great circle about
    ``tilted_axis``, referenced to ``angles[0]``) or (A, 3) unit-sphere
    points; the ramp filter is |kx*v + ky*u| with (u, v, w) the normalised
    ``tilted_axis``.  Returns f of shape (Nx, Ny, Nx).  ``num_cores`` and
    ``save_memory`` are validated but have no effect.  Extension keyword
    ``prefilter`` as in backpropagate_3d.
    """
    A = angles.shape[0]
    if angles.shape not in [(A,), (A, 1), (A, 3)]:
        raise ValueError("`angles` must have shape (A,) or (A,3)!")
    _validate_common(uSin, A, padding, coords, num_cores, _ncores)
    dtype, cdt = _precision(dtype)
    precision = N.ODT_FP32 if dtype == np.float32 else N.ODT_FP64
    if max_count is not None:
        max_count.value += A + 2
    uSin = _as_input(uSin, copy)
    angles = np.array(angles, copy=copy)
    angles = np.squeeze(angles)
    lny, lnx = uSin.shape[1], uSin.shape[2]
    ln = lnx
    padyl, padxl, lNy, lNx = _pads(lny, lnx, padding, padfac)
    if verbose > 0:
        print("......Image size (x,y): {}x{}, padded: {}x{}".format(
            lnx, lny, lNx, lNy))
    pad_mode, pv_re, pv_im = _pad_mode(padding, padval, lny, lnx, lNy, lNx)

    tilted_axis = norm_vec(tilted_axis)
    angz = np.arctan2(tilted_axis[0], tilted_axis[1])
    rotmat = np.array([
        [np.cos(angz), -np.sin(angz), 0],
        [np.sin(angz), np.cos(angz), 0],
        [0, 0, 1],
    ])
    axis_yz = norm_vec(np.dot(rotmat, tilted_axis))
    weights = np.ones(A)
    if len(angles.shape) == 1:
        if weight_angles:
            weights = util.compute_angle_weights_1d(angles)
        angles = sphere_points_from_angles_and_tilt(angles, axis_yz)
    else:
        if weight_angles:
            warnings.warn("3D angular weighting not yet supported!")
        for i in range(angles.shape[0]):
            angles[i] = norm_vec(np.dot(rotmat, angles[i]))
    if weight_angles and not copy and isinstance(uSin, np.ndarray):
        uSin *= weights.reshape(-1, 1, 1)
        weights = np.ones(A)

    km = (2 * np.pi * nm) / res
    u, v, _w = tilted_axis
    points = rotate_points_to_axis(points=angles, axis=axis_yz)
    params = affine_params(points, angz, (lnx, lny, ln))
    d_usin, was_tensor = _core.device_input(uSin, cdt)
    fields = dict(
        kind=N.ODT_KIND_3D_TILTED, precision=precision, A=A, Ny=lny, Nx=lnx,
        lNy=lNy, lNx=lNx, padyl=padyl, padxl=padxl, pad_mode=pad_mode,
        intp_order=int(intp_order), onlyreal=int(bool(onlyreal)),
        flags=N.ODT_FLAG_IIR_PREFILTER if prefilter == "iir" else 0,
        angle_param_stride=12, padval_re=pv_re, padval_im=pv_im,
        km=float(km), lD=float(lD), z0=-ln / 2.0, ramp_u=float(u),
        ramp_v=float(v))
    esz = (4 if precision == N.ODT_FP32 else 8) * (1 if onlyreal else 2)
    fields["slab_budget_bytes"] = _core.slab_budget(
        ln * lny * lnx * esz, A, d_usin.device)
    if count is not None:
        count.value += 2
    out = _core.run(fields, weights, params, d_usin, (ln, lny, lnx),
                    bool(onlyreal), precision, was_tensor)
    if count is not None:
        count.value += A
    return out
