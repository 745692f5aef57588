"""Host-side float64 geometry of the tilted-axis reconstruction.

Same functions and conventions as odtbrain/_alg3d_bppt.py:26-371.  The
matrix products are evaluated in the same order as the reference
(numpy.dot), so the per-angle inverse rotation and offset handed to the
affine gather kernel carry the reference's rounding; the kernel's
out-of-bounds decisions (exact, scipy constant mode) depend on those bits.
"""
import numpy as np


def norm_vec(vector):
    """Unit vector along ``vector`` (_alg3d_bppt.py:26-30)."""
    assert len(vector) == 3
    v = np.array(vector)
    return v / np.sqrt(np.sum(v ** 2))


def _ry(phi, sign):
    """Rotation in the x-z plane; sign=-1 gives [[c,0,-s],[0,1,0],[s,0,c]]."""
    c, s = np.cos(phi), np.sin(phi)
    return np.array([[c, 0, sign * s], [0, 1, 0], [-sign * s, 0, c]])


def _rx(theta, sign):
    """Rotation in the y-z plane; sign=-1 gives [[1,0,0],[0,c,-s],[0,s,c]]."""
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[1, 0, 0], [0, c, sign * s], [0, -sign * s, c]])


def rotate_points_to_axis(points, axis):
    """Rotate ``points`` so that ``axis`` becomes [0,1,0] (_alg3d_bppt.py:33-120).

    Rotation by phi in the x-z plane into the y-z plane, by theta in the
    y-z plane up to the y axis, then back by -phi in the x-z plane.
    """
    u, v, w = norm_vec(axis)
    pts = np.array(points)
    phi = np.arctan2(u, w)
    theta = np.arccos(v)
    r_phi = _ry(phi, -1)
    r_theta = _rx(theta, +1)
    tilt = np.dot(r_theta, r_phi)
    full = np.dot(r_phi.T, tilt)
    out = np.zeros((len(pts), 3))
    for i, p in enumerate(pts):
        out[i] = np.dot(full, p)
    return out


def rotation_matrix_from_point(point, ret_inv=False):
    """Matrix that rotates [0,0,1] onto ``point`` (_alg3d_bppt.py:123-175)."""
    x, y, z = point
    phi = np.arctan2(x, z)
    theta = -np.arctan2(y, np.sqrt(x ** 2 + z ** 2))
    r_theta = _rx(theta, -1)
    r_phi = _ry(phi, -1)
    d = np.dot(r_phi, r_theta)
    if not ret_inv:
        return d
    return d, np.dot(r_theta.T, r_phi.T)


def rotation_matrix_from_point_planerot(point, plane_angle, ret_inv=False):
    """As above followed by a rotation about z by ``plane_angle``
    (_alg3d_bppt.py:178-227)."""
    d, dinv = rotation_matrix_from_point(point, ret_inv=True)
    c, s = np.cos(plane_angle), np.sin(plane_angle)
    rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    dr = np.dot(d, rz)
    if not ret_inv:
        return dr
    return dr, np.dot(rz.T, dinv)


def sphere_points_from_angles_and_tilt(angles, tilted_axis):
    """Unit-sphere points of ``angles`` on the great circle about a tilted
    axis (_alg3d_bppt.py:230-371).  Point 0 is [0,0,1]; angles are taken
    relative to ``angles[0]``.
    """
    assert len(angles.shape) == 1
    u, v, w = norm_vec(tilted_axis)
    pts = np.zeros((angles.shape[0], 3), dtype=float)
    rel = angles - angles[0]
    pts[:, 0] = np.sin(rel)
    pts[:, 2] = np.cos(rel)
    theta = np.arccos(v)
    if np.allclose(u, 0) and np.allclose(w, 0):
        phi = 0
    else:
        phi = np.arctan2(u, w)
    # (a1) shrink the circle to radius cos(theta)
    pts *= np.cos(theta)
    # (a2) tilt about x by theta
    rx = _rx(theta, -1)
    for i in range(pts.shape[0]):
        pts[i] = np.dot(rx, pts[i])
    # (a3) shift so that point 0 lies at (0, 0, 1)
    pts = pts - (pts[0] - np.array([0, 0, 1])).reshape(1, 3)
    # (b) rotate about y by phi
    ry = _ry(phi, +1)
    for i in range(pts.shape[0]):
        pts[i] = np.dot(ry, pts[i])
    return pts
