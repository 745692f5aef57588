"""Restatement of the B-spline resampling that ODTbrain calls (TEST INFRASTRUCTURE).

The reference rotates with ``scipy.ndimage.rotate`` (odtbrain/_alg3d_bpp.py:73-81,
_alg2d_bpp.py:311-321) and ``scipy.ndimage.affine_transform``
(odtbrain/_alg3d_bppt.py:931-944), pinned by ``scipy>=1.4,<2``
(pyproject.toml:37; 1.15.3 installed here).  This module restates their
published algorithm in plain numpy so that the semantics the GPU kernels
implement are written down and checked:

* prefilter: exact inverse of the mirror-extended B-spline sampling operator,
  computed as the causal/anti-causal recursive filter of Unser (1993) and
  Thevenaz, Blu & Unser (2000) with the mirror-boundary initialisations
  (causal: closed sum over the period 2n-2; anti-causal: z/(z^2-1) rule);
* resampling: ``in = M @ out + offset`` in float64 with the sum order
  ``((0 + o0*m0) + o1*m1 [+ o2*m2]) + offset``; a sample with any input
  coordinate < 0 or > n-1 is exactly ``cval`` (no tolerance); otherwise the
  order-k B-spline is evaluated on mirror-extended coefficients, first tap
  ``floor(c + 0.5) - k//2`` (even k) or ``floor(c) - k//2`` (odd k).

Checked against scipy.ndimage on random inputs and against the reference
fixture ``alg3d_bpp__test_3d_mprotate`` (tests/test_oracle_bspline.py).
"""
import math

import numpy as np

POLES = {
    2: [math.sqrt(8.0) - 3.0],
    3: [math.sqrt(3.0) - 2.0],
    4: [math.sqrt(664.0 - math.sqrt(438976.0)) + math.sqrt(304.0) - 19.0,
        math.sqrt(664.0 + math.sqrt(438976.0)) - math.sqrt(304.0) - 19.0],
    5: [math.sqrt(67.5 - math.sqrt(4436.25)) + math.sqrt(26.25) - 6.5,
        math.sqrt(67.5 + math.sqrt(4436.25)) - math.sqrt(26.25) - 6.5],
}


def bspline(x, order):
    """Centred B-spline of degree ``order`` at ``x`` (truncated-power form)."""
    x = np.asarray(x, dtype=np.float64)
    n = order
    if n == 0:
        return np.where((x >= -0.5) & (x < 0.5), 1.0, 0.0)
    out = np.zeros_like(x)
    for k in range(n + 2):
        t = x + (n + 1) / 2.0 - k
        out += (-1) ** k * math.comb(n + 1, k) * np.where(t > 0, t, 0.0) ** n
    return out / math.factorial(n)


def _filter_line(c, order):
    """In-place mirror-boundary prefilter of one float64 line."""
    n = c.size
    if n == 1:
        return c
    poles = POLES[order]
    gain = 1.0
    for z in poles:
        gain *= (1.0 - z) * (1.0 - 1.0 / z)
    c *= gain
    for z in poles:
        # causal initialisation for the mirror extension (period 2n-2)
        zn1 = z ** (n - 1)
        c0 = c[0] + zn1 * c[n - 1]
        zi = z
        for i in range(1, n - 1):
            c0 += zi * (c[i] + zn1 * c[n - 1 - i])
            zi *= z
        c[0] = c0 / (1.0 - zn1 * zn1)
        for i in range(1, n):
            c[i] += z * c[i - 1]
        # anti-causal initialisation
        c[n - 1] = (z * c[n - 2] + c[n - 1]) * z / (z * z - 1.0)
        for i in range(n - 2, -1, -1):
            c[i] = z * (c[i + 1] - c[i])
    return c


def spline_filter1d(arr, order, axis):
    """Prefilter along one axis (float64 result), mirror boundary."""
    a = np.moveaxis(np.array(arr, dtype=np.float64), axis, -1)
    flat = a.reshape(-1, a.shape[-1])
    for row in flat:
        _filter_line(row, order)
    return np.moveaxis(flat.reshape(a.shape), -1, axis)


def spline_filter(arr, order):
    out = np.array(arr, dtype=np.float64)
    if order < 2:
        return out
    for ax in range(out.ndim):
        out = spline_filter1d(out, order, ax)
    return out


def mirror_index(idx, n):
    """Whole-sample-symmetric index extension onto [0, n-1]."""
    if n <= 1:
        return 0
    s2 = 2 * n - 2
    if idx < 0:
        idx = s2 * int(-idx / s2) + idx
        idx = idx + s2 if idx <= 1 - n else -idx
    elif idx >= n:
        idx -= s2 * int(idx / s2)
        if idx >= n:
            idx = s2 - idx
    return idx


def tap_start(c, order):
    return (math.floor(c + 0.5) if order % 2 == 0 else math.floor(c)) \
        - order // 2


def weights(c, order):
    """Tap weights w[j] for taps start..start+order at coordinate ``c``."""
    s = tap_start(c, order)
    return bspline(c - (s + np.arange(order + 1)), order)


def map_coordinates_affine(coef, matrix, offset, out_shape, order, cval=0.0):
    """Resample prefiltered coefficients ``coef`` at ``matrix @ o + offset``."""
    matrix = np.asarray(matrix, dtype=np.float64)
    offset = np.asarray(offset, dtype=np.float64)
    nd = coef.ndim
    out = np.empty(out_shape, dtype=np.float64)
    for o in np.ndindex(*out_shape):
        cc = []
        for h in range(nd):
            t = 0.0
            for ll in range(len(o)):
                t += float(o[ll]) * matrix[h, ll]
            cc.append(t + offset[h])
        if any(c < 0 or c > coef.shape[h] - 1 for h, c in enumerate(cc)):
            out[o] = cval
            continue
        idxs, wts = [], []
        for h, c in enumerate(cc):
            s = tap_start(c, order)
            idxs.append([mirror_index(s + j, coef.shape[h])
                         for j in range(order + 1)])
            wts.append(weights(c, order))
        val = 0.0
        for taps in np.ndindex(*([order + 1] * nd)):
            w = 1.0
            for h in range(nd):
                w *= wts[h][taps[h]]
            val += coef[tuple(idxs[h][taps[h]] for h in range(nd))] * w
        out[o] = val
    return out


def rotate_plane(img, angle_deg, order, cosdg, sindg):
    """scipy.ndimage.rotate(img, angle_deg, reshape=False) for a 2D array.

    ``cosdg``/``sindg`` are passed in (scipy.special's degree-exact trig).
    """
    c, s = cosdg(angle_deg), sindg(angle_deg)
    rot = np.array([[c, s], [-s, c]])
    shape = np.asarray(img.shape)
    offset = (shape - 1) / 2 - rot @ ((shape - 1) / 2)
    coef = spline_filter(img, order) if order > 1 else np.asarray(
        img, dtype=np.float64)
    return map_coordinates_affine(coef, rot, offset, img.shape, order)
