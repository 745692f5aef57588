"""Host logic of odtbrain_amd and the C ABI library: CPU only (no compute
call reaches the GPU)."""
import ctypes
import json
import multiprocessing as mp
import re

import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.special

from common import ROOT_DIR, create_test_sino_3d
import odtbrain_amd
from odtbrain_amd import _core, _geometry, util
from odtbrain_amd import _native as N
from odtbrain_amd._alg3d_bppt import affine_params
from oracle import odt_oracle as O


def header_functions():
    text = (ROOT_DIR / "include" / "odtbrain_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(odt_\w+)\(",
                                 text, flags=re.M)))


def test_library_exports_every_header_symbol():
    names = header_functions()
    assert len(names) >= 10
    lib = ctypes.CDLL(str(N.LIB_PATH))
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.EXPORTED)
    assert N.lib().odt_version() == 1


def test_desc_struct_layout():
    # 14 int32, 7 double, 2 pointers, 1 int64
    assert ctypes.sizeof(N.OdtDesc) == 14 * 4 + 7 * 8 + 2 * 8 + 8


def test_weights_closed_forms():
    """tests/test_util.py:6-32 of the reference."""
    w = util.compute_angle_weights_1d(np.linspace(0, np.pi, 18,
                                                  endpoint=False))
    assert np.allclose(w, 1, rtol=0, atol=1e-14)
    w = util.compute_angle_weights_1d(np.linspace(0, np.pi, 18,
                                                  endpoint=True))
    assert np.allclose(w[1:-1], 18 / 17, rtol=0, atol=1e-14)
    assert np.allclose(w[[0, -1]], 18 / 17 / 2, rtol=0, atol=1e-14)
    a0 = np.linspace(0, np.pi, 14, endpoint=False)
    a1 = np.roll(a0, -3)
    a = np.roll(np.concatenate((np.ones(4) * a1[0], a1)), 3)
    w = util.compute_angle_weights_1d(a)
    assert np.allclose(w[3:8], 18 / 14 / 5, rtol=0, atol=1e-14)
    assert np.allclose(np.mean(w), 1, rtol=0, atol=1e-14)


def test_weights_match_oracle_random():
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = rng.uniform(-7, 7, size=rng.integers(2, 50))
        a[::5] = a[0]
        assert np.allclose(util.compute_angle_weights_1d(a),
                           O.compute_angle_weights_1d(a), rtol=1e-13)


def test_geometry_bit_identical_to_oracle():
    rng = np.random.default_rng(4)
    for _ in range(10):
        axis = rng.standard_normal(3)
        angles = np.sort(rng.uniform(0, 6, 17))
        pg = _geometry.sphere_points_from_angles_and_tilt(angles, axis)
        po = O.sphere_points_from_angles_and_tilt(angles, axis)
        assert np.array_equal(pg, po)
        assert np.array_equal(_geometry.rotate_points_to_axis(pg, axis),
                              O.rotate_points_to_axis(po, axis))
        for p in pg[:3]:
            a = _geometry.rotation_matrix_from_point_planerot(p, 0.3, True)
            b = O.rotation_matrix_from_point_planerot(p, 0.3, True)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_planar_params_reproduce_scipy_rotate():
    """The (matrix, offset) handed to the kernel resample like scipy's
    rotate: affine_transform with them equals rotate, bit for bit."""
    rng = np.random.default_rng(5)
    img = rng.standard_normal((13, 13))
    for ang in (-2.0, 90.0, -270.0, 123.4, -359.0):
        p = _core.planar_params([ang], 13, 13)[0]
        m = np.array([[p[0], p[1]], [p[2], p[3]]])
        a = ndi.affine_transform(img, m, offset=p[4:6], order=2,
                                 mode="constant", cval=0)
        b = ndi.rotate(img, ang, reshape=False, order=2, mode="constant",
                       cval=0)
        assert np.array_equal(a, b)


def test_cosdg_exact_at_quadrants():
    p = _core.planar_params([-270.0, 90.0, 180.0], 8, 8)
    assert p[0][0] == 0.0 and p[1][0] == 0.0 and p[2][1] == 0.0


def test_affine_params_match_reference_expression():
    pts = _geometry.sphere_points_from_angles_and_tilt(
        np.linspace(0, 3, 5), [0, 1, 0.2])
    prm = affine_params(pts, 0.1, (7, 8, 7))
    c = 0.5 * np.array((7, 8, 7)) - .5
    for i, p in enumerate(pts):
        _d, dinv = O.rotation_matrix_from_point_planerot(p, 0.1, True)
        assert np.array_equal(prm[i, :9], dinv.reshape(-1))
        assert np.array_equal(prm[i, 9:], c - np.dot(dinv, c))


def test_padded_order():
    assert _core.padded_order(144, 1.75) == 256
    assert _core.padded_order(150, 1.75) == 512
    assert _core.padded_order(22, 1.75) == 64
    assert _core.padded_order(256, 1.75) == 512


def test_validation_errors_before_gpu():
    """Exceptions of _alg3d_bpp.py:239-258 / _alg3d_bppt.py:564-578 /
    _alg2d_bpp.py:136-141, raised before any device work."""
    sino, angles = create_test_sino_3d(A=4, Nx=8, Ny=8)
    f3 = odtbrain_amd.backpropagate_3d
    with pytest.raises(ValueError):
        f3(sino[0], angles, 2.0, 1.333)
    with pytest.raises(ValueError):
        f3(sino[:3], angles, 2.0, 1.333)
    with pytest.raises(ValueError):
        f3(sino, angles, 2.0, 1.333, padding=(True,))
    with pytest.raises(ValueError):
        f3(sino, angles, 2.0, 1.333, padding=(1, 0))
    with pytest.raises(NotImplementedError):
        f3(sino, angles, 2.0, 1.333, coords=np.zeros((3, 2)))
    with pytest.raises(ValueError):
        f3(sino, angles, 2.0, 1.333, num_cores=mp.cpu_count() + 1)
    with pytest.raises(ValueError):
        f3(sino, angles, 2.0, 1.333, dtype=np.int32)
    with pytest.raises(ValueError):
        f3(sino, angles, 2.0, 1.333, padding=(True, False))
    ft = odtbrain_amd.backpropagate_3d_tilted
    with pytest.raises(ValueError):
        ft(sino, np.zeros((4, 2)), 2.0, 1.333)
    f2 = odtbrain_amd.backpropagate_2d
    with pytest.raises(AssertionError):
        f2(sino, angles, 2.0, 1.333)
    with pytest.raises(TypeError):
        f2(sino[:, 0, :], angles, 2.0, 1.333, coords=np.zeros((2, 2)))


def test_plan_create_rejects_bad_descriptions():
    """C ABI validation runs on the host before any device call."""
    w = np.ones(3)
    prm = np.zeros(18)
    base = dict(kind=0, precision=0, A=3, Ny=4, Nx=4, lNy=4, lNx=4,
                padyl=0, padxl=0, pad_mode=0, intp_order=2, onlyreal=0,
                flags=0, angle_param_stride=6, km=1.0, z0=-2.0, ramp_v=1.0)
    bad = [dict(base, precision=7), dict(base, intp_order=6),
           dict(base, angle_param_stride=12), dict(base, lNx=3),
           dict(base, pad_mode=0, lNx=8), dict(base, km=0.0),
           dict(base, kind=2), dict(base, A=0)]
    for fields in bad:
        d = N.OdtDesc()
        for k, v in fields.items():
            setattr(d, k, v)
        d.weights = N.dptr(w)
        d.angle_params = N.dptr(prm)
        h = ctypes.c_void_p()
        rc = N.lib().odt_plan_create(ctypes.byref(d), ctypes.byref(h))
        assert rc == 1, fields
        assert N.lib().odt_last_error()
        assert not h.value


def test_rotate_accumulate_rejects_non_rotation():
    """The standalone rotate entry point validates its matrices on the host
    before any device work, so this runs without a GPU."""
    params = np.array([[2.0, 0.0, 0.0, 1.0, 0.0, 0.0]])
    dummy = ctypes.c_void_p(1)
    rc = N.lib().odt_rotate_accumulate(N.ODT_FP32, 2, 1, 1, 8, 8, dummy,
                                       N.dptr(params), dummy, 0, None)
    assert rc == N.ODT_ERR_INVALID
    assert b"rotation" in N.lib().odt_last_error()


def test_tuning_cache_format_and_lookup(tmp_path, monkeypatch):
    """The autotuner's cache: entries round-trip through the JSON format,
    the user cache overrides, files of another format are ignored, and
    ODTBRAIN_AMD_NO_TUNING disables lookups."""
    from odtbrain_amd import tune
    path = tmp_path / "t.json"
    monkeypatch.setenv("ODTBRAIN_AMD_TUNE_CACHE", str(path))
    tune.reload()
    desc = dict(kind=0, precision=0, A=40, Ny=48, Nx=48, lNy=128, lNx=128,
                intp_order=2, onlyreal=0, flags=0)
    assert tune.lookup(desc) is None
    tune.save(tune.shape_key(desc), {"rotate_group": 36, "unknown": 5})
    data = json.loads(path.read_text())
    assert data["format"] == "odtbrain_amd-tuning" and data["version"] == 1
    assert tune.lookup(desc) == {"rotate_group": 36}
    assert tune.lookup(dict(desc, A=41)) is None
    monkeypatch.setenv("ODTBRAIN_AMD_NO_TUNING", "1")
    assert tune.lookup(desc) is None
    monkeypatch.delenv("ODTBRAIN_AMD_NO_TUNING")
    path.write_text(json.dumps({"format": "other", "entries": {}}))
    tune.reload()
    assert tune.lookup(desc) is None
    tune.reload()


def test_tuning_struct_layout():
    from odtbrain_amd import _native as N
    assert ctypes.sizeof(N.OdtTuning) == 24
    assert N.OdtTuning.t_budget_bytes.offset == 8
