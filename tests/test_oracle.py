"""Pins of the CPU oracle (oracle/) against the reference's own evidence:
golden fixtures of ODTbrain's tests, closed forms and known answers, and
scipy.ndimage for the B-spline restatement.  CPU only."""
import ctypes
import multiprocessing as mp

import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.special

from common import (create_test_sino_2d, create_test_sino_3d,
                    create_test_sino_3d_tilted, cutout,
                    get_test_parameter_set, load_golden, quarter_turn_case)
from oracle import apple as AP
from oracle import bspline as B
from oracle import odt_oracle as O


def test_oracle_3d_backprop_phase_fixture():
    """tests/test_alg3d_bpp.py:30-42 (padval=0 ramp, complex)."""
    sino, angles = create_test_sino_3d()
    r = [cutout(O.backpropagate_3d(sino, angles, padval=0, dtype=float, **p))
         for p in get_test_parameter_set(2)]
    data = np.array(r).flatten().view(float)
    assert np.allclose(data, load_golden("alg3d_bpp__test_3d_backprop_phase"))


def test_oracle_3d_backprop_nopadreal_fixture():
    """tests/test_alg3d_bpp.py:45-62 (no padding, onlyreal)."""
    sino, angles = create_test_sino_3d()
    r = [cutout(O.backpropagate_3d(sino, angles, padding=(False, False),
                                   dtype=float, onlyreal=True, **p))
         for p in get_test_parameter_set(2)]
    data = np.array(r).flatten().view(float)
    assert np.allclose(data,
                       load_golden("alg3d_bpp__test_3d_backprop_nopadreal"))


def test_oracle_mprotate_fixture_pool():
    """tests/test_alg3d_bpp.py:116-134: arange(20^3) rotated by 2 degrees at
    order 2 over y in [0, 10) through a process pool on a shared array."""
    ln2, ln = 20, 10
    shared = mp.RawArray(ctypes.c_double, ln2 ** 3)
    arr = np.frombuffer(shared).reshape(ln2, ln2, ln2)
    arr[:] = np.arange(ln2 ** 3).reshape(ln2, ln2, ln2)
    pool = mp.get_context("fork").Pool(
        3, initializer=O._pool_init,
        initargs=(shared, (ln2, ln2, ln2), np.dtype(float)))
    try:
        O.mprotate(2, ln, pool, 2, 3)
    finally:
        pool.terminate()
        pool.join()
    assert np.allclose(arr.flatten(), load_golden("alg3d_bpp__test_3d_mprotate"))


@pytest.mark.parametrize("name,ampl", [("alg2d_bpp__test_2d_backprop_phase",
                                        (1.0, 1.0)),
                                       ("alg2d_bpp__test_2d_backprop_full",
                                        (0.9, 1.1))])
def test_oracle_2d_fixtures(name, ampl):
    """tests/test_alg2d_bpp.py:13-36."""
    sino, angles = create_test_sino_2d(ampl_range=ampl)
    r = [cutout(O.backpropagate_2d(sino, angles, **p))
         for p in get_test_parameter_set(2)]
    assert np.allclose(np.array(r).flatten().view(float), load_golden(name))


def test_oracle_tilted_equals_plain_for_y_axis():
    """tests/test_alg3d_bppt.py:10-50 (both pinned through the fixtures)."""
    sino, angles = create_test_sino_3d()
    for p in get_test_parameter_set(2):
        a = cutout(O.backpropagate_3d(sino, angles, padval=0, dtype=float,
                                      onlyreal=True, **p))
        b = cutout(O.backpropagate_3d_tilted(sino, angles, padval=0,
                                             dtype=float, onlyreal=True, **p))
        assert np.allclose(a, b)


def test_oracle_weights_closed_forms():
    """tests/test_util.py:6-32."""
    w = O.compute_angle_weights_1d(np.linspace(0, np.pi, 18, endpoint=False))
    assert np.allclose(w, 1, rtol=0, atol=1e-14)
    w = O.compute_angle_weights_1d(np.linspace(0, np.pi, 18, endpoint=True))
    assert np.allclose(w[1:-1], 18 / 17, rtol=0, atol=1e-14)
    assert np.allclose(w[[0, -1]], 18 / 17 / 2, rtol=0, atol=1e-14)
    a0 = np.linspace(0, np.pi, 14, endpoint=False)
    a1 = np.roll(a0, -3)
    a = np.roll(np.concatenate((np.ones(4) * a1[0], a1)), 3)
    w = O.compute_angle_weights_1d(a)
    assert np.allclose(np.mean(w), 1, rtol=0, atol=1e-14)
    assert np.allclose(w[:3], 18 / 14, rtol=0, atol=1e-14)
    assert np.allclose(w[8:], 18 / 14, rtol=0, atol=1e-14)
    assert np.allclose(w[3:8], 18 / 14 / 5, rtol=0, atol=1e-14)


def test_oracle_sphere_points_known_answers():
    """tests/test_spherecoords_from_angles_and_axis.py:8-32."""
    angles = np.array([0, np.pi / 2, np.pi])
    axes = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 1, 1], [1, 0, 1], [1, 1, 1]]
    res = [O.sphere_points_from_angles_and_tilt(angles, ax) for ax in axes]
    s2 = 1 / np.sqrt(2)
    correct = np.array([[[1, 0, 0], [1, 0, 0], [1, 0, 0]],
                        [[0, 0, 1], [1, 0, 0], [0, 0, -1]],
                        [[0, 0, 1], [0, 0, 1], [0, 0, 1]],
                        [[0, 0, 1], [s2, .5, .5], [0, 1, 0]],
                        [[s2, 0, s2], [s2, 0, s2], [s2, 0, s2]],
                        [[s2, 0, s2],
                         [0.87965281125489458, s2 / 3 * 2,
                          0.063156230327168605],
                         [s2 / 3, s2 / 3 * 4, s2 / 3]]])
    assert np.allclose(correct, np.array(res))


def test_oracle_rotation_matrices_known_answers():
    """tests/test_rotation_matrices.py:8-102."""
    sq2 = np.sqrt(2)
    r = O.rotate_points_to_axis([[1, 2, 3]], [1, 2, 3])
    assert r[0][0] < 1e-14 and r[0][2] < 1e-14
    r = O.rotate_points_to_axis([[0, 0, 1], [1, 0, 0], [1, 1, 0]], [0, 1, 1])
    assert np.allclose(r, [[0, 1 / sq2, 1 / sq2], [1, 0, 0],
                           [1, 1 / sq2, -1 / sq2]])
    r = O.rotate_points_to_axis([[0, 0, 1], [1, 0, 0], [0, -1, 0]], [1, 0, 1])
    assert np.allclose(r, [[-.5, 1 / sq2, .5], [.5, 1 / sq2, -.5],
                           [1 / sq2, 0, 1 / sq2]])
    r = O.rotate_points_to_axis([[0, 0, 1], [1, 0, 1], [1, -1, 0]], [1, 0, 0])
    assert np.allclose(r, [[0, 0, 1], [0, 1, 1], [1, 1, 0]])
    r = O.rotate_points_to_axis([[0, 0, 1], [1, 0, 0], [1, -1, 0]], [0, -1, 0])
    assert np.allclose(r, [[0, 0, -1], [1, 0, 0], [1, 1, 0]])
    m = O.rotation_matrix_from_point([0, 1, 1])
    assert np.allclose(m @ [0, 0, 1], [0, 1 / sq2, 1 / sq2])
    assert np.allclose(m @ [1, 0, 0], [1, 0, 0])
    m = O.rotation_matrix_from_point([-1, 1, 0])
    assert np.allclose(m @ [1, 0, 0], [0, 0, -1])
    assert np.allclose(m @ [0, 1 / sq2, -1 / sq2], [-1, 0, 0])


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_bspline_prefilter_matches_scipy(order):
    rng = np.random.default_rng(order)
    x = rng.standard_normal((6, 11))
    for ax in (0, 1):
        ref = ndi.spline_filter1d(x, order, axis=ax, mode="mirror")
        assert np.allclose(B.spline_filter1d(x, order, ax), ref,
                           rtol=0, atol=1e-12)


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
def test_bspline_rotate_matches_scipy(order):
    """Including the exact out-of-bounds set at 90-degree multiples."""
    rng = np.random.default_rng(10 + order)
    for ang in (2.0, 90.0, -270.0, 33.3, 180.0, -17.5):
        img = rng.standard_normal((9, 12))
        got = B.rotate_plane(img, ang, order, scipy.special.cosdg,
                             scipy.special.sindg)
        ref = ndi.rotate(img, ang, reshape=False, order=order,
                         mode="constant", cval=0)
        assert np.array_equal(got == 0, ref == 0)
        assert np.allclose(got, ref, rtol=0, atol=1e-11)


def test_bspline_restatement_reproduces_mprotate_fixture():
    ln2 = 20
    arr = np.arange(ln2 ** 3, dtype=float).reshape(ln2, ln2, ln2)
    for y in range(10):
        arr[:, y, :] = B.rotate_plane(arr[:, y, :], -2.0, 2,
                                      scipy.special.cosdg,
                                      scipy.special.sindg)
    assert np.allclose(arr.flatten(), load_golden("alg3d_bpp__test_3d_mprotate"))


def test_oracle_planted_mistake_is_caught():
    """A dropped term in the prefactor (the lD phase) must break the fixture
    check, showing the fixture pins more than the formula's shape."""
    sino, angles = create_test_sino_3d()
    p = get_test_parameter_set(2)[1]           # lD = pi
    good = cutout(O.backpropagate_3d(sino, angles, padval=0, dtype=float, **p))
    bad = cutout(O.backpropagate_3d(sino, angles, padval=0, dtype=float,
                                    **dict(p, lD=0)))
    ref = load_golden("alg3d_bpp__test_3d_backprop_phase").view(complex)
    ref = ref[ref.size // 2:].reshape(good.shape)
    assert np.allclose(good, ref)
    assert not np.allclose(bad, ref)


# --------------------------------------------------------------------------
# odt_to_ri / opt_to_ri (odtbrain/_translate_ri.py)
# --------------------------------------------------------------------------
def test_translate_oracle_fixtures():
    from oracle import translate as T
    f, res, nm = T.processing_test_data()
    for name, fn in (("processing__test_odt_to_ri", T.odt_to_ri),
                     ("processing__test_opt_to_ri", T.opt_to_ri)):
        ri = fn(f, res, nm)
        assert np.allclose(np.array(ri).flatten().view(float),
                           load_golden(name))
        assert np.allclose(fn(f[0], res, nm), ri[0])


def test_odt_to_ri_oracle_inverts_the_object_function():
    """f = km^2 ((n/nm)^2 - 1) maps back to n for any n with positive real
    part (a dropped km^2 or nm, or the other root, fails this)."""
    from oracle import translate as T
    rng = np.random.default_rng(3)
    res, nm = 6.5, 1.333
    km = 2 * np.pi * nm / res
    n = rng.uniform(1.0, 1.6, 500) + 1j * rng.uniform(-0.2, 0.2, 500)
    f = km ** 2 * ((n / nm) ** 2 - 1)
    assert np.allclose(T.odt_to_ri(f, res, nm), n, rtol=1e-13, atol=0)
    # opt_to_ri is affine with slope res / (2 pi) and value nm at 0
    g = rng.standard_normal(50)
    assert np.allclose(T.opt_to_ri(g, res, nm),
                       nm + g * res / (2 * np.pi))
    assert T.opt_to_ri(np.zeros(1), res, nm)[0] == nm


# --------------------------------------------------------------------------
# sinogram_as_rytov / sinogram_as_radon (odtbrain/_prepare_sino.py)
# --------------------------------------------------------------------------
def _mod_rest(a, b):
    """Remainder closest to zero (tests/test_processing.py:35-48)."""
    q = np.array(a / b, dtype=int)
    r = a - b * q
    wrong = np.where(np.abs(r) > b / 2)
    r[wrong] -= b * np.sign(r[wrong])
    return r


def test_prepare_oracle_fixtures():
    """tests/test_processing.py:79-137: the fixtures (with the reference
    test's own 2 pi allowance) and the 2D-vs-3D consistency."""
    from oracle import prepare as P
    twopi = 2 * np.pi
    sino = P.processing_test_sino()
    rad = P.sinogram_as_radon(sino)
    if rad.min() > twopi:
        rad -= twopi
    assert np.allclose(np.array(rad).flatten().view(float),
                       load_golden("processing__test_sino_radon"))
    for sl in ((slice(None), slice(None), 0), (slice(None), 0, slice(None))):
        r2 = P.sinogram_as_radon(sino[sl])
        assert np.allclose(0, _mod_rest(r2 - rad[sl], twopi), atol=1e-6)
    sino = P.processing_test_sino(rytov=True)
    ryt = P.sinogram_as_rytov(sino)
    if ryt.imag.min() > twopi:
        ryt.imag -= twopi
    assert np.allclose(np.array(ryt).flatten().view(float),
                       load_golden("processing__test_sino_rytov"))
    for sl in ((slice(None), slice(None), 0), (slice(None), 0, slice(None))):
        r2 = P.sinogram_as_rytov(sino[sl])
        d = r2 - ryt[sl]
        assert np.allclose(0, (d.real + 1j * _mod_rest(d.imag, twopi))
                           .view(float), atol=1e-6)


def test_unwrap_2d_residue_free_fields_are_recovered():
    """Smooth phase fields spanning many turns: wrapping then unwrapping
    gives the field back up to one whole-turn offset (a wrong increment sign,
    a dropped merge shift or a swapped neighbour fails this)."""
    from oracle import prepare as P
    rng = np.random.default_rng(5)
    for H, W in ((9, 11), (16, 13)):
        y, x = np.mgrid[0:H, 0:W]
        phi = np.zeros((H, W))
        for _ in range(3):
            cy, cx = rng.uniform(0, H), rng.uniform(0, W)
            s = rng.uniform(2, 6)
            phi += rng.uniform(-12, 12) * np.exp(-((y - cy) ** 2 +
                                                  (x - cx) ** 2) / (2 * s * s))
        phi += 0.7 * x - 0.4 * y
        # residue free: neighbouring differences scaled to at most 2.5 rad
        md = max(np.abs(np.diff(phi, axis=0)).max(),
                 np.abs(np.diff(phi, axis=1)).max())
        phi *= 2.5 / md
        assert np.ptp(phi) > 3 * np.pi
        wrapped = np.angle(np.exp(1j * phi))
        un = P.unwrap_2d(wrapped)
        k = (un - phi) / (2 * np.pi)
        assert np.allclose(k, np.round(k[0, 0]), atol=1e-9)


def test_unwrap_2d_merge_rules_known_answer():
    """A 1x4 strip has three horizontal edges of equal reliability (all
    border pixels): they merge in creation order, each singleton joining
    the group of its left neighbour, so the first pixel keeps its value."""
    from oracle import prepare as P
    wrapped = np.array([[3.0, -3.0, 3.1, -3.1]])
    un = P.unwrap_2d(wrapped)
    assert np.allclose(un, [[3.0, -3.0 + 2 * np.pi, 3.1, -3.1 + 2 * np.pi]])


def test_align_and_divmod_neg_known_answers():
    from oracle import prepare as P
    twopi = 2 * np.pi
    assert P.divmod_neg(5.0, twopi) == (1.0, 5.0 - twopi)
    assert P.divmod_neg(-4.0, twopi) == (-1.0, -4.0 + twopi)
    assert P.divmod_neg(3.0, twopi) == (0.0, 3.0)
    # a ramp over the angles with whole-turn jumps at every corner sample:
    # alignment removes the jumps and the turns of the minimum
    A, N = 6, 8
    base = np.linspace(0.1, 0.6, A)[:, None] + np.zeros((A, N))
    jumps = twopi * np.array([0, 1, 1, 2, 2, 2])[:, None]
    sino = base + jumps + 3 * twopi
    P.align_unwrapped(sino)
    assert np.allclose(sino, base)


# --------------------------------------------------------------------------
# apple.correct (odtbrain/apple.py)
# --------------------------------------------------------------------------
def test_apple_oracle_fixture():
    """tests/test_apple.py:48-69 through the oracle's own preparation and
    reconstruction (atol 2e-6, the reference test's bar)."""
    from oracle import apple as AP
    from oracle import prepare as P
    from oracle import odt_oracle as O
    sino, angles = create_test_sino_3d(Nx=10, Ny=10)
    p = get_test_parameter_set(1)[0]
    sryt = P.sinogram_as_rytov(uSin=sino, u0=1, align=False)
    f = O.backpropagate_3d(sryt, angles, padval=0, dtype=float, **p)
    fc = AP.correct(f=f, res=p["res"], nm=p["nm"], enforce_envelope=.95,
                    max_iter=100, min_diff=0.01)
    fo = np.array(cutout(fc), dtype=np.complex128)
    assert np.allclose(fo.flatten().view(float),
                       load_golden("apple__test_correct_reproduce"),
                       atol=2e-6)


def test_apple_masks_closed_forms():
    """apple_core_3d: the axis of symmetry is y; on the ky axis (kx = kz =
    0) every frequency with |ky| > 0 inside the sqrt(2) km sphere is core,
    nothing outside it; spillover_region is the complement of the inscribed
    ellipsoid (centre excluded, corners included)."""
    from oracle import apple as AP
    shape = (16, 12, 16)
    res, nm = 3.0, 1.333
    core = AP.apple_core_3d(shape, res, nm)
    km = 2 * np.pi * nm / res
    ky = 2 * np.pi * np.fft.fftfreq(12)
    expect = (np.abs(ky) > 0) & (np.abs(ky) <= np.sqrt(2) * km)
    assert np.array_equal(core[0, :, 0], expect)
    assert not core[0, 0, 0]
    with pytest.raises(ValueError):
        AP.apple_core_3d((10, 10, 5), .1, 1)
    spov = AP.spillover_region((8, 6, 8))
    assert not spov[4, 3, 4] and spov[0, 0, 0] and spov[7, 5, 7]


@pytest.mark.parametrize("which", ["3d", "tilted", "2d"])
def test_angle_parallel_equals_sequential(which):
    """The angle-split evaluation used for the full-size GPU checks adds the
    same per-angle contributions as the sequential loop: equal to rounding,
    and an empty split (one worker) is the sequential call itself."""
    if which == "2d":
        s, a = create_test_sino_2d(A=11, N=20)
        f, args, kw = O.backpropagate_2d, (2.0, 1.333), {}
    else:
        s, a = create_test_sino_3d(A=9, Nx=12, Ny=10)
        f, args = ((O.backpropagate_3d, (2.0, 1.333)) if which == "3d" else
                   (O.backpropagate_3d_tilted, (2.0, 1.333)))
        kw = {} if which == "3d" else {"tilted_axis": [0.2, 1, 0.1]}
    seq = f(s, a, *args, **kw)
    par = O.angle_parallel(f, s, a, *args, workers=4, **kw)
    assert np.abs(par - seq).max() <= 1e-14 * np.abs(seq).max()
    assert np.array_equal(O.angle_parallel(f, s, a, *args, workers=1, **kw),
                          seq)
    # a dropped block would show: the first block alone differs
    part = f(s, a, *args, angle_range=(0, 3), **kw)
    assert np.abs(part - seq).max() > 1e-3 * np.abs(seq).max()


def test_oracle_tilted_x_axis_equals_quarter_turned_y_axis():
    """Turning every projection a quarter turn (np.rot90, k=1, detector
    y -> x) turns the rotation axis from +y to -x; the tilted path rotates
    that axis back onto the y-z plane (angz) and builds its plane rotation
    into the resampling, so its reconstruction must equal the y-axis one to
    float64 rounding.  This pins the x component of the tilted ramp
    |kx v + ky u| and the angz handling (_alg3d_bppt.py:736-771, 904-936).
    The antiparallel axis (+x) must not match."""
    sino, angles = quarter_turn_case()
    p = get_test_parameter_set(1)[0]
    fy = O.backpropagate_3d_tilted(sino, angles, padval="edge",
                                   tilted_axis=[0, 1, 0], **p)
    turned = np.rot90(sino, k=1, axes=(1, 2)).copy()
    fx = O.backpropagate_3d_tilted(turned, angles, padval="edge",
                                   tilted_axis=[-1, 0, 0], **p)
    assert np.linalg.norm(fx - fy) <= 1e-12 * np.linalg.norm(fy)
    anti = O.backpropagate_3d_tilted(turned, angles, padval="edge",
                                     tilted_axis=[1, 0, 0], **p)
    assert np.linalg.norm(anti - fy) > 1e-2 * np.linalg.norm(fy)


def test_oracle_tilted_soft_plane_rotation_and_alignment():
    """tests/test_alg3d_bppt.py:53-126 (ODTbrain's soft tests, atol/rtol
    0.2) run on the oracle."""
    p = get_test_parameter_set(1)[0]

    def rec(angz, onlyreal):
        sino, angles = create_test_sino_3d_tilted(tilt_plane=angz, A=21)
        rotmat = np.array([[np.cos(angz), -np.sin(angz), 0],
                           [np.sin(angz), np.cos(angz), 0],
                           [0, 0, 1]])
        return O.backpropagate_3d_tilted(
            sino, angles, padval="edge", tilted_axis=rotmat @ [0, 1, 0],
            padding=(False, False), dtype=float, onlyreal=onlyreal, **p)

    res = [cutout(rec(a, False)).flatten().view(float)
           for a in [1.1, 0.0, 0.234, 2.80922, -.29, 9.87]]
    for i in range(len(res)):
        assert np.allclose(res[i], res[i - 1], atol=.2, rtol=.2)
    res = [rec(a, True) for a in [0, np.pi / 2, np.pi, 3 * np.pi / 2,
                                   2 * np.pi]]
    for i in range(len(res)):
        assert np.allclose(res[i], res[i - 1], atol=.2, rtol=.2)


def test_apple_constraint_sh_known_answer():
    """constraint_sh (apple.py:74-121) on a volume whose histogram is known
    by construction.  Spill-over region (outside the inscribed ellipsoid):
    values -1 (and one 1), so fmin = -1, fmax = 1 and the 100 bins are 0.02
    wide, bin i = [-1 + 0.02 i, -0.98 + 0.02 i).  Inside: the peak bin 50
    [0, 0.02) (0.005 and 0.009), 2 values in bin 49 (-0.015), 50 in bin 48
    (-0.035), 10 in bin 47 (-0.055), single values at 0.5 and -0.7.
    count_to_half on the bins below the peak in reverse measures against
    bin 49 (2 values): bins 48 and 47 are not below 1, bin 46 (empty) is,
    so the band starts at bin 48 and its lower end is that bin's centre
    -0.03; above the peak bin 51 is empty, so the upper end is the peak's
    centre 0.01.  Values strictly between fmin and -0.03 or strictly
    between 0.01 and fmax are halved: -0.035, -0.055, -0.7 and 0.5.  The
    value -0.015 is kept only because of that reference bin: measured
    against the peak (858 values), bin 49 would already end the band."""
    shape = (12, 12, 12)
    spov = AP.spillover_region(shape)
    inner = np.argwhere(~spov)
    data = np.zeros(shape, dtype=np.complex128)
    data.real[spov] = -1.0
    corner = tuple(np.argwhere(spov)[0])
    data.real[corner] = 1.0
    vals = ([0.005] * 60 + [-0.015] * 2 + [-0.035] * 50 + [-0.055] * 10 +
            [0.5, -0.7])
    assert len(vals) <= len(inner)
    rest = [0.009] * (len(inner) - len(vals))  # also in the peak bin
    for p, v in zip(inner, vals + rest):
        data.real[tuple(p)] = v
    data.imag[:] = 0.3
    before = data.real.copy()
    AP.constraint_sh(data)
    assert np.all(data.imag == 0)
    want = before.copy()
    for v in (-0.035, -0.055, -0.7, 0.5):
        want[before == v] = v / 2
    assert np.array_equal(data.real, want)
    assert np.all(data.real[before == -0.015] == -0.015)
    assert np.all(data.real[np.abs(before) == 1] == before[np.abs(before) == 1])
