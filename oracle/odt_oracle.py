"""CPU restatement of ODTbrain's filtered backpropagation (TEST INFRASTRUCTURE).

Checker and CPU baseline only -- never imported by the product package.
Each function cites the reference lines it restates.  FFTW is replaced by
``scipy.fft`` (unnormalised backward transform via ``norm="forward"``),
numexpr by numpy; scipy.ndimage is called with the reference's arguments.
"""
import ctypes
import multiprocessing as mp

import numpy as np
import scipy.fft
import scipy.ndimage

# --------------------------------------------------------------------------
# util.compute_angle_weights_1d  (odtbrain/util.py:4-67)
# --------------------------------------------------------------------------


def compute_angle_weights_1d(angles):
    """Angular weights, coverage wrapped at pi, duplicates share a weight.

    Restates odtbrain/util.py:40-67: shift to min and wrap mod pi (:42),
    unique angles with counts (:45-46), neighbour distance of the sorted
    unique angles wrapped mod pi (:54), divide by multiplicity (:58) and
    normalise so that the mean weight is one (:59).
    """
    wrapped = (np.asarray(angles).flatten() - np.min(angles)) % np.pi
    uniq, inverse, counts = np.unique(wrapped, return_inverse=True,
                                      return_counts=True)
    order = np.argsort(uniq)
    su = uniq[order]
    gap = (np.roll(su, -1) - np.roll(su, 1)) % np.pi
    total = np.sum(gap)
    gap = gap / counts[order]
    w_sorted = gap / total * wrapped.size
    w_uniq = np.zeros_like(w_sorted)
    w_uniq[order] = w_sorted
    return w_uniq[inverse.reshape(-1)]


# --------------------------------------------------------------------------
# padding sizes (odtbrain/_alg3d_bpp.py:285-304, _alg2d_bpp.py:174-182)
# --------------------------------------------------------------------------


def padded_order(n, padfac):
    """``int(max(64, 2**ceil(log(n*padfac)/log(2))))`` (_alg3d_bpp.py:286)."""
    return int(max(64., 2 ** np.ceil(np.log(n * padfac) / np.log(2))))


def _pad_projection(img, padding, padval, pads):
    """Pad one projection as in _alg3d_bpp.py:493-506."""
    padyl, padyr, padxl, padxr = pads
    if not (padding[0] and padding[1]):
        return img
    if isinstance(padval, str) and padval == "edge":
        return np.pad(img, ((padyl, padyr), (padxl, padxr)), mode="edge")
    return np.pad(img, ((padyl, padyr), (padxl, padxr)),
                  mode="linear_ramp", end_values=(padval,))


def _dtypes(dtype):
    """dtype handling of _alg3d_bpp.py:254-260."""
    if dtype is None:
        dtype = np.float64
    dtype = np.dtype(dtype)
    if dtype.name not in ["float32", "float64"]:
        raise ValueError("dtype must be float32 or float64!")
    return dtype, np.dtype("complex{}".format(2 * int(dtype.name[5:])))


def _filters(lNx, lNy, km, A, lD, tilted_uv=None):
    """Frequency-domain factors of _alg3d_bpp.py:371-399 / _alg3d_bppt.py:737-771.

    Returns (prefactor, f2_exp_fac) where prefactor already carries the
    ramp (|kx| or the tilted |kx*v + ky*u|) and the low-pass mask.
    """
    kx = (2 * np.pi * np.fft.fftfreq(lNx)).reshape(1, -1)
    ky = (2 * np.pi * np.fft.fftfreq(lNy)).reshape(-1, 1)
    dphi0 = 2 * np.pi / A
    klp = (kx ** 2 + ky ** 2 < km ** 2)
    M = 1. / km * np.sqrt((km ** 2 - kx ** 2 - ky ** 2) * klp)
    pref = -1j * km / (2 * np.pi)
    pref *= dphi0
    if tilted_uv is None:
        pref = pref * (np.abs(kx) * klp)
    else:
        u, v = tilted_uv
        pref = pref * (np.abs(kx * v + ky * u) * klp)
    pref = pref * np.exp(-1j * km * (M - 1) * lD)
    return pref, 1j * km * (M - 1)


# --------------------------------------------------------------------------
# y-chunk rotation through a process pool (odtbrain/_alg3d_bpp.py:17-81)
# --------------------------------------------------------------------------

_shared = {}


def _pool_init(buf, shape, dtype):
    _shared["buf"] = buf
    _shared["shape"] = shape
    _shared["dtype"] = dtype


def _rotate_chunk(task):
    """One worker task: rotate y-planes [y0, y1) in place (_alg3d_bpp.py:68-81)."""
    y0, y1, ang_deg, order = task
    arr = np.frombuffer(_shared["buf"], dtype=_shared["dtype"]).reshape(
        _shared["shape"])
    scipy.ndimage.rotate(arr[:, y0:y1, :], angle=-ang_deg, axes=(0, 2),
                         reshape=False, output=arr[:, y0:y1, :], order=order,
                         mode="constant", cval=0)


def mprotate(ang_deg, lny, pool, order, ncores):
    """Split [0, lny) into ``ncores`` chunks and rotate them (_alg3d_bpp.py:35-65)."""
    size = int(np.floor(lny / ncores))
    tasks = []
    for t in range(ncores):
        y0 = t * size
        y1 = lny if t == ncores - 1 else (t + 1) * size
        tasks.append((y0, y1, ang_deg, order))
    pool.map(_rotate_chunk, tasks)


def rotate_inplace(arr, ang_deg, y0, y1, order):
    """Single-process equivalent of one mprotate call on planes [y0, y1)."""
    scipy.ndimage.rotate(arr[:, y0:y1, :], angle=-ang_deg, axes=(0, 2),
                         reshape=False, output=arr[:, y0:y1, :], order=order,
                         mode="constant", cval=0)


# --------------------------------------------------------------------------
# backpropagate_3d (odtbrain/_alg3d_bpp.py:84-555)
# --------------------------------------------------------------------------


def backpropagate_3d(uSin, angles, res, nm, lD=0, weight_angles=True,
                     onlyreal=False, padding=(True, True), padfac=1.75,
                     padval="edge", intp_order=2, dtype=None, num_cores=1,
                     save_memory=False, angle_range=None, time_limit=None,
                     stats=None):
    """Restatement of backpropagate_3d (_alg3d_bpp.py:84-555).

    ``num_cores > 1`` rotates y-chunks in a ``multiprocessing.Pool`` over a
    shared ``RawArray`` and threads the FFTs, like the reference.
    ``angle_range=(a0, a1)`` restricts the angle loop to a contiguous subset
    and ``time_limit`` (seconds) stops the loop once exceeded (both only to
    time a bounded CPU-baseline sample; the filters are still built for the
    full ``A``).  ``stats`` (a dict) receives setup/loop seconds and the
    number of angles processed.
    """
    import time
    t_start = time.perf_counter()
    A = angles.size
    if len(uSin.shape) != 3:
        raise ValueError("Input data `uSin` must have shape (A,Ny,Nx).")
    if len(uSin) != A:
        raise ValueError("`len(angles)` must be  equal to `len(uSin)`.")
    dtype, dtype_complex = _dtypes(dtype)
    uSin = np.array(uSin, copy=True)
    lny, lnx = uSin.shape[1], uSin.shape[2]
    ln = lnx
    padx = padded_order(lnx, padfac) - lnx if padding[0] else 0
    pady = padded_order(lny, padfac) - lny if padding[1] else 0
    padyl = int(np.ceil(pady / 2))
    padxl = int(np.ceil(padx / 2))
    pads = (padyl, pady - padyl, padxl, padx - padxl)
    lNx, lNy = lnx + padx, lny + pady
    if weight_angles:
        uSin *= compute_angle_weights_1d(angles).reshape(-1, 1, 1)
    km = (2 * np.pi * nm) / res
    prefactor, f2_exp_fac = _filters(lNx, lNy, km, A, lD)
    z = np.linspace(-ln / 2.0, ln / 2.0, ln, endpoint=False)
    if not save_memory:
        filter2 = np.exp(f2_exp_fac * z.reshape(-1, 1, 1))
    outarr = np.zeros((ln, lny, lnx),
                      dtype=dtype if onlyreal else dtype_complex)
    workers = max(1, int(num_cores))
    if workers > 1:
        shared = mp.RawArray(ctypes.c_float if dtype == np.float32
                             else ctypes.c_double, ln * lny * lnx)
        ctx = mp.get_context("fork")
        pool = ctx.Pool(processes=workers, initializer=_pool_init,
                        initargs=(shared, (ln, lny, lnx), dtype))
        arr = np.frombuffer(shared, dtype=dtype).reshape(ln, lny, lnx)
    else:
        pool = None
        arr = np.zeros((ln, lny, lnx), dtype=dtype)
    filtered = np.zeros((ln, lny, lnx), dtype=dtype_complex)
    a0, a1 = (0, A) if angle_range is None else angle_range
    t_loop = time.perf_counter()
    done = 0
    try:
        for aa in range(a0, a1):
            if time_limit is not None and done >= 1 and \
                    time.perf_counter() - t_loop > time_limit:
                break
            done += 1
            one = np.asarray(_pad_projection(uSin[aa], padding, padval, pads),
                             dtype=dtype_complex)
            one = scipy.fft.fft2(one, workers=workers)
            one *= prefactor / (lNx * lNy)
            for p in range(ln):
                if save_memory:
                    spec = np.exp(f2_exp_fac * z[p]) * one
                else:
                    spec = filter2[p] * one
                spec = np.asarray(spec, dtype=dtype_complex)
                inv = scipy.fft.ifft2(spec, norm="forward", workers=workers)
                filtered[p] = inv[padyl:padyl + lny, padxl:padxl + lnx]
            phi0 = np.rad2deg(angles[aa])
            for part in ((0,) if onlyreal else (0, 1)):
                arr[:] = filtered.real if part == 0 else filtered.imag
                if pool is not None:
                    mprotate(phi0, lny, pool, intp_order, workers)
                else:
                    rotate_inplace(arr, phi0, 0, lny, intp_order)
                if onlyreal:
                    outarr += arr
                elif part == 0:
                    outarr.real += arr
                else:
                    outarr.imag += arr
    finally:
        if pool is not None:
            pool.terminate()
            pool.join()
    if stats is not None:
        t_end = time.perf_counter()
        stats.update(setup_s=t_loop - t_start, loop_s=t_end - t_loop,
                     angles=done, workers=workers)
    return outarr


# --------------------------------------------------------------------------
# tilted-axis geometry (odtbrain/_alg3d_bppt.py:26-371)
# --------------------------------------------------------------------------


def norm_vec(vector):
    """Unit vector (_alg3d_bppt.py:26-30)."""
    assert len(vector) == 3
    v = np.array(vector)
    return v / np.sqrt(np.sum(v ** 2))


def rotate_points_to_axis(points, axis):
    """Rotate points so that ``axis`` maps to [0,1,0] (_alg3d_bppt.py:33-120)."""
    u, v, w = norm_vec(axis)
    pts = np.array(points)
    phi = np.arctan2(u, w)
    theta = np.arccos(v)
    r_phi = np.array([[np.cos(phi), 0, -np.sin(phi)],
                      [0, 1, 0],
                      [np.sin(phi), 0, np.cos(phi)]])
    r_theta = np.array([[1, 0, 0],
                        [0, np.cos(theta), np.sin(theta)],
                        [0, -np.sin(theta), np.cos(theta)]])
    dr = np.dot(r_phi.T, np.dot(r_theta, r_phi))
    out = np.zeros((len(pts), 3))
    for i, p in enumerate(pts):
        out[i] = np.dot(dr, p)
    return out


def rotation_matrix_from_point(point, ret_inv=False):
    """Matrix taking [0,0,1] to ``point`` (_alg3d_bppt.py:123-175)."""
    x, y, z = point
    phi = np.arctan2(x, z)
    theta = -np.arctan2(y, np.sqrt(x ** 2 + z ** 2))
    r_theta = np.array([[1, 0, 0],
                        [0, np.cos(theta), -np.sin(theta)],
                        [0, np.sin(theta), np.cos(theta)]])
    r_phi = np.array([[np.cos(phi), 0, -np.sin(phi)],
                      [0, 1, 0],
                      [np.sin(phi), 0, np.cos(phi)]])
    d = np.dot(r_phi, r_theta)
    dinv = np.dot(r_theta.T, r_phi.T)
    return (d, dinv) if ret_inv else d


def rotation_matrix_from_point_planerot(point, plane_angle, ret_inv=False):
    """Same with an extra in-plane rotation about z (_alg3d_bppt.py:178-227)."""
    d, dinv = rotation_matrix_from_point(point, ret_inv=True)
    rz = np.array([[np.cos(plane_angle), -np.sin(plane_angle), 0],
                   [np.sin(plane_angle), np.cos(plane_angle), 0],
                   [0, 0, 1]])
    dr = np.dot(d, rz)
    drinv = np.dot(rz.T, dinv)
    return (dr, drinv) if ret_inv else dr


def sphere_points_from_angles_and_tilt(angles, tilted_axis):
    """Points on the great circle about a tilted axis (_alg3d_bppt.py:230-371)."""
    assert len(angles.shape) == 1
    u, v, w = norm_vec(tilted_axis)
    pts = np.zeros((angles.shape[0], 3), dtype=float)
    pts[:, 0] = np.sin(angles - angles[0])
    pts[:, 2] = np.cos(angles - angles[0])
    theta = np.arccos(v)
    if np.allclose(u, 0) and np.allclose(w, 0):
        phi = 0
    else:
        phi = np.arctan2(u, w)
    pts *= np.cos(theta)
    rx = np.array([[1, 0, 0],
                   [0, np.cos(theta), -np.sin(theta)],
                   [0, np.sin(theta), np.cos(theta)]])
    for i in range(pts.shape[0]):
        pts[i] = np.dot(rx, pts[i])
    pts = pts - (pts[0] - np.array([0, 0, 1])).reshape(1, 3)
    ry = np.array([[+np.cos(phi), 0, np.sin(phi)],
                   [0, 1, 0],
                   [-np.sin(phi), 0, np.cos(phi)]])
    for i in range(pts.shape[0]):
        pts[i] = np.dot(ry, pts[i])
    return pts


def backpropagate_3d_tilted(uSin, angles, res, nm, lD=0,
                            tilted_axis=(0, 1, 0), weight_angles=True,
                            onlyreal=False, padding=(True, True),
                            padfac=1.75, padval="edge", intp_order=2,
                            dtype=None, num_cores=1, save_memory=False,
                            angle_range=None):
    """Restatement of backpropagate_3d_tilted (_alg3d_bppt.py:374-949)."""
    import warnings
    A = angles.shape[0]
    if angles.shape not in [(A,), (A, 1), (A, 3)]:
        raise ValueError("`angles` must have shape (A,) or (A,3)!")
    if len(uSin.shape) != 3:
        raise ValueError("Input data `uSin` must have shape (A,Ny,Nx).")
    if len(uSin) != A:
        raise ValueError("`len(angles)` must be  equal to `len(uSin)`.")
    dtype, dtype_complex = _dtypes(dtype)
    uSin = np.array(uSin, copy=True)
    angles = np.squeeze(np.array(angles, copy=True))
    lny, lnx = uSin.shape[1], uSin.shape[2]
    ln = lnx
    padx = padded_order(lnx, padfac) - lnx if padding[0] else 0
    pady = padded_order(lny, padfac) - lny if padding[1] else 0
    padyl = int(np.ceil(pady / 2))
    padxl = int(np.ceil(padx / 2))
    pads = (padyl, pady - padyl, padxl, padx - padxl)
    lNx, lNy = lnx + padx, lny + pady
    tilted_axis = norm_vec(tilted_axis)
    angz = np.arctan2(tilted_axis[0], tilted_axis[1])
    rotmat = np.array([[np.cos(angz), -np.sin(angz), 0],
                       [np.sin(angz), np.cos(angz), 0],
                       [0, 0, 1]])
    axis_yz = norm_vec(np.dot(rotmat, tilted_axis))
    if len(angles.shape) == 1:
        if weight_angles:
            weights = compute_angle_weights_1d(angles).reshape(-1, 1, 1)
        angles = sphere_points_from_angles_and_tilt(angles, axis_yz)
    else:
        if weight_angles:
            warnings.warn("3D angular weighting not yet supported!")
            weights = 1
        for i in range(angles.shape[0]):
            angles[i] = norm_vec(np.dot(rotmat, angles[i]))
    if weight_angles:
        uSin *= weights
    km = (2 * np.pi * nm) / res
    u, v, _w = tilted_axis
    prefactor, f2_exp_fac = _filters(lNx, lNy, km, A, lD, tilted_uv=(u, v))
    z = np.linspace(-ln / 2.0, ln / 2.0, ln, endpoint=False)
    if not save_memory:
        filter2 = np.exp(f2_exp_fac * z.reshape(-1, 1, 1))
    outarr = np.zeros((ln, lny, lnx),
                      dtype=dtype if onlyreal else dtype_complex)
    workers = max(1, int(num_cores))
    filtered = np.zeros((ln, lny, lnx), dtype=dtype_complex)
    points = rotate_points_to_axis(points=angles, axis=axis_yz)
    a0, a1 = (0, A) if angle_range is None else angle_range
    for aa in range(a0, a1):
        one = np.asarray(_pad_projection(uSin[aa], padding, padval, pads),
                         dtype=dtype_complex)
        one = scipy.fft.fft2(one, workers=workers)
        one *= prefactor / (lNx * lNy)
        for p in range(ln):
            if save_memory:
                spec = np.exp(f2_exp_fac * z[p]) * one
            else:
                spec = filter2[p] * one
            spec = np.asarray(spec, dtype=dtype_complex)
            inv = scipy.fft.ifft2(spec, norm="forward", workers=workers)
            filtered[p] = inv[padyl:padyl + lny, padxl:padxl + lnx]
        fpt = filtered.transpose(2, 1, 0)[:, ::-1, :]
        _d, dinv = rotation_matrix_from_point_planerot(points[aa], angz,
                                                       ret_inv=True)
        c = 0.5 * np.array(fpt.shape) - .5
        offset = c - np.dot(dinv, c)
        parts = (("real",) if onlyreal else ("real", "imag"))
        for part in parts:
            rot = scipy.ndimage.affine_transform(
                getattr(fpt, part), dinv, offset=offset, mode="constant",
                cval=0, order=intp_order).transpose(2, 1, 0)[:, ::-1, :]
            if onlyreal:
                outarr += rot
            elif part == "real":
                outarr.real += rot
            else:
                outarr.imag += rot
    return outarr


# --------------------------------------------------------------------------
# backpropagate_2d (odtbrain/_alg2d_bpp.py:8-326)
# --------------------------------------------------------------------------


def backpropagate_2d(uSin, angles, res, nm, lD=0, weight_angles=True,
                     onlyreal=False, padding=True, padval=0,
                     angle_range=None):
    """Restatement of backpropagate_2d (_alg2d_bpp.py:132-326).

    The reference runs in float64/complex128 throughout; so does this.
    """
    A = angles.shape[0]
    assert len(uSin.shape) == 2, "Input data `uB` must have shape (A,N)!"
    assert len(uSin) == A, "`len(angles)` must be  equal to `len(uSin)`!"
    km = (2 * np.pi * nm) / res
    if weight_angles:
        sinogram = uSin * compute_angle_weights_1d(angles).reshape(-1, 1)
    else:
        sinogram = uSin
    ln = sinogram.shape[1]
    order = max(64., 2 ** np.ceil(np.log(ln * 2.1) / np.log(2)))
    pad = order - ln if padding else 0
    padl = int(np.ceil(pad / 2))
    padr = int(pad - padl)
    if padval is None:
        sino = np.pad(sinogram, ((0, 0), (padl, padr)), mode="edge")
    else:
        sino = np.pad(sinogram, ((0, 0), (padl, padr)), mode="linear_ramp",
                      end_values=(padval,))
    lN = sino.shape[1]
    kx = (2 * np.pi * np.fft.fftfreq(lN)).reshape(1, -1)
    dphi0 = 2 * np.pi / A
    klp = (kx ** 2 < km ** 2)
    M = 1. / km * np.sqrt((km ** 2 - kx ** 2) * klp)
    pref = -1j * km / (2 * np.pi)
    pref *= dphi0
    pref = pref * (np.abs(kx) * klp)
    pref = pref * np.exp(-1j * km * (M - 1) * lD)
    projection = np.fft.fft(sino, axis=-1) * pref
    x = np.arange(lN) - ln / 2.0 + .5
    filter2 = np.exp(1j * x.reshape(-1, 1) * km * (M.reshape(1, -1) - 1))
    projection = projection.reshape(A, 1, lN)
    outarr = np.zeros((ln, ln)) if onlyreal else \
        np.zeros((ln, ln), dtype=np.dtype(complex))
    a0, a1 = (0, A) if angle_range is None else angle_range
    for i in range(a0, a1):
        filt = np.fft.ifft(projection[i] * filter2, axis=-1)
        sino_i = filt[:ln, padl:padl + ln]
        outarr += scipy.ndimage.rotate(sino_i.real, -angles[i] * 180 / np.pi,
                                       reshape=False, mode="constant", cval=0)
        if not onlyreal:
            outarr += 1j * scipy.ndimage.rotate(
                sino_i.imag, -angles[i] * 180 / np.pi, reshape=False,
                mode="constant", cval=0)
    return outarr


# --------------------------------------------------------------------------
# angle-parallel evaluation (test infrastructure, no reference counterpart)
# --------------------------------------------------------------------------

_ap_job = {}


def _ap_worker(rng):
    func, args, kw = _ap_job["job"]
    return func(*args, angle_range=rng, **kw)


def angle_parallel(func, uSin, angles, *args, workers=1, **kw):
    """Evaluate one of the three restated functions with its angle loop split
    into ``workers`` contiguous blocks run in a fork process pool, and return
    the sum of the partial volumes in block order.

    Every function above builds its filters and angle weights from all
    angles before the loop and then adds one rotated slab per angle, so the
    partial volume of ``angle_range=(a0, a1)`` is exactly the sum of those
    angles' contributions; only the order of the floating-point additions
    differs from the sequential loop (rounding level).  Inner pools are
    disabled (``num_cores=1`` where the function takes it)."""
    A = len(angles)
    workers = max(1, min(int(workers), A))
    bounds = np.linspace(0, A, workers + 1).round().astype(int)
    ranges = [(int(a), int(b)) for a, b in zip(bounds[:-1], bounds[1:])
              if b > a]
    if func is not backpropagate_2d:
        kw = dict(kw, num_cores=1)
    if workers == 1:
        return func(uSin, angles, *args, **kw)
    _ap_job["job"] = (func, (uSin, angles) + tuple(args), kw)
    ctx = mp.get_context("fork")
    try:
        with ctx.Pool(processes=len(ranges)) as pool:
            parts = pool.map(_ap_worker, ranges, chunksize=1)
    finally:
        _ap_job.clear()
    out = parts[0]
    for p in parts[1:]:
        out += p
    return out
