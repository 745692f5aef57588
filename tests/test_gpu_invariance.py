"""Invariance tests of the reference mirrored for the GPU path:
tests/test_angle_weights.py:9-60 (here for backpropagate_3d as well),
tests/test_save_memory.py:9-47 and tests/test_alg3d_bppt.py:53-126."""
import numpy as np
import pytest

import odtbrain_amd as G

from common import (create_test_sino_3d, create_test_sino_3d_tilted, cutout,
                    get_test_parameter_set)

pytestmark = pytest.mark.gpu


def test_3d_angle_offset_invariance():
    """Whole turns added to every other angle change nothing, with or
    without angle weights (uniform angles)."""
    sino, angles = create_test_sino_3d(A=10)
    params = get_test_parameter_set(2)
    r1 = [G.backpropagate_3d(sino, angles, weight_angles=False, padval=0,
                             dtype=np.float64, **p) for p in params]
    ang2 = angles.copy()
    ang2[::2] += 2 * np.pi * np.arange(ang2[::2].shape[0])
    r2 = [G.backpropagate_3d(sino, ang2, weight_angles=False, padval=0,
                             dtype=np.float64, **p) for p in params]
    r3 = [G.backpropagate_3d(sino, ang2, weight_angles=True, padval=0,
                             dtype=np.float64, **p) for p in params]
    assert np.allclose(np.array(r1).view(float), np.array(r2).view(float))
    assert np.allclose(np.array(r2).view(float), np.array(r3).view(float))


def test_3d_angle_swap_invariance():
    """Permuting the angles (and their projections) changes nothing for
    non-uniform angles with weights."""
    sino, angles = create_test_sino_3d(A=11)
    sino, angles = sino[:-2], angles[:-2]
    params = get_test_parameter_set(2)
    r1 = [G.backpropagate_3d(sino, angles, weight_angles=True,
                             dtype=np.float64, **p) for p in params]
    order = np.argsort(angles % .5)
    r2 = [G.backpropagate_3d(sino[order], angles[order], weight_angles=True,
                             dtype=np.float64, **p) for p in params]
    assert np.allclose(np.array(r1).view(float), np.array(r2).view(float))


@pytest.mark.parametrize("fn", ["3d", "tilted"])
def test_save_memory_and_num_cores_have_no_effect(fn):
    sino, angles = create_test_sino_3d(A=8, Nx=18, Ny=14)
    p = get_test_parameter_set(1)[0]
    f = G.backpropagate_3d if fn == "3d" else G.backpropagate_3d_tilted
    a = f(sino, angles, padval="edge", dtype=np.float64, **p)
    b = f(sino, angles, padval="edge", dtype=np.float64, save_memory=True,
          num_cores=1, **p)
    assert np.array_equal(a, b)


def _tilted_runs(tilts, onlyreal, params):
    out = []
    for angz in tilts:
        sino, angles = create_test_sino_3d_tilted(tilt_plane=angz, A=21)
        rotmat = np.array([[np.cos(angz), -np.sin(angz), 0],
                           [np.sin(angz), np.cos(angz), 0],
                           [0, 0, 1]])
        axis = np.dot(rotmat, [0, 1, 0])
        res = []
        for p in params:
            f = G.backpropagate_3d_tilted(sino, angles, padval="edge",
                                          tilted_axis=axis,
                                          padding=(False, False), dtype=float,
                                          onlyreal=onlyreal, **p)
            res.append(cutout(f) if not onlyreal else f)
        out.append(np.array(res).flatten().view(float) if not onlyreal
                   else np.array(res).flatten())
    return out


def test_tilted_plane_rotation_soft():
    """tests/test_alg3d_bppt.py:53-89: in-plane rotated sinograms with the
    axis rotated alike reconstruct alike (soft tolerance 0.2)."""
    res = _tilted_runs([1.1, 0.0, 0.234, 2.80922, -.29, 9.87], False,
                       get_test_parameter_set(1))
    for ii in range(len(res)):
        assert np.allclose(res[ii], res[ii - 1], atol=.2, rtol=.2)


def test_tilted_alignment_along_axes_soft():
    """tests/test_alg3d_bppt.py:92-126: the reconstruction follows the axis
    (not its antiparallel) for quarter-turn tilts of the plane."""
    res = _tilted_runs([0, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi], True,
                       get_test_parameter_set(1)[:1])
    for ii in range(len(res)):
        assert np.allclose(res[ii], res[ii - 1], atol=.2, rtol=.2)
