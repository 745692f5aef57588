"""GPU parity of the drop-in functions against the reference fixtures and
the CPU oracle on the same seeded inputs.

Tolerances: the float64 path must reproduce the reference fixtures with the
reference's own check (np.allclose defaults) and the oracle to relative L2
1e-9; the float32 path must stay within relative L2 1e-5 of the complex128
oracle (the accuracy bar of this project) and within the reference's own
FP32 bar atol 6e-7 on its fixture inputs (tests/test_alg3d_bpp.py:101-113).
"""
import multiprocessing as mp
import os

import numpy as np
import pytest

import odtbrain_amd as G
from common import (create_test_sino_2d, create_test_sino_3d, cutout,
                    get_test_parameter_set, load_golden, quarter_turn_case,
                    rel_l2)
from odtbrain_amd import phantoms
from oracle import odt_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FP32_REL_L2 = 1e-5
FP64_REL_L2 = 1e-9


def test_3d_backprop_phase_fixture():
    sino, angles = create_test_sino_3d()
    r = [cutout(G.backpropagate_3d(sino, angles, padval=0, dtype=float, **p))
         for p in get_test_parameter_set(2)]
    data = np.array(r).flatten().view(float)
    assert np.allclose(data, load_golden("alg3d_bpp__test_3d_backprop_phase"))


def test_3d_backprop_nopadreal_fixture():
    sino, angles = create_test_sino_3d()
    r = [cutout(G.backpropagate_3d(sino, angles, padding=(False, False),
                                   dtype=float, onlyreal=True, **p))
         for p in get_test_parameter_set(2)]
    data = np.array(r).flatten().view(float)
    assert np.allclose(data,
                       load_golden("alg3d_bpp__test_3d_backprop_nopadreal"))


@pytest.mark.parametrize("name,ampl", [("alg2d_bpp__test_2d_backprop_phase",
                                        (1.0, 1.0)),
                                       ("alg2d_bpp__test_2d_backprop_full",
                                        (0.9, 1.1))])
def test_2d_fixtures(name, ampl):
    sino, angles = create_test_sino_2d(ampl_range=ampl)
    r = [cutout(G.backpropagate_2d(sino, angles, **p))
         for p in get_test_parameter_set(2)]
    assert np.allclose(np.array(r).flatten().view(float), load_golden(name))


def test_3d_fp32_against_fp64_fixture_inputs():
    """The project's rel-L2 bar and the reference's FP32 bar (atol 6e-7 on
    the cut-out volume, tests/test_alg3d_bpp.py:101-113)."""
    sino, angles = create_test_sino_3d()
    for p in get_test_parameter_set(2):
        f32 = G.backpropagate_3d(sino, angles, padval=0, dtype=np.float32,
                                 **p)
        ref = O.backpropagate_3d(sino, angles, padval=0, dtype=float, **p)
        assert f32.dtype == np.complex64
        err = rel_l2(f32, ref)
        d = cutout(f32.astype(np.complex128)) - cutout(ref.copy())
        maxabs = float(np.abs(d.view(float)).max())
        print(f"fp32 fixture inputs {p}: rel_l2={err:.2e} maxabs={maxabs:.2e}")
        assert err < FP32_REL_L2
        assert maxabs < 6e-7, maxabs


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("padval", ["edge", 0.5])
def test_3d_orders_and_padding_vs_oracle(order, padval):
    sino, angles = create_test_sino_3d(A=13, Nx=18, Ny=14)
    kw = dict(res=3.0, nm=1.335, lD=1.5, padval=padval, intp_order=order)
    ref = O.backpropagate_3d(sino, angles, **kw)
    got = G.backpropagate_3d(sino, angles, **kw)
    assert got.shape == ref.shape == (18, 14, 18)
    assert rel_l2(got, ref) < FP64_REL_L2
    got32 = G.backpropagate_3d(sino, angles, dtype=np.float32, **kw)
    assert rel_l2(got32, ref) < FP32_REL_L2


@pytest.mark.parametrize("nx,ny", [(32, 15), (34, 7), (36, 4), (32, 3)])
def test_3d_rotate_plane_chunks_vs_oracle(nx, ny):
    """y sizes around the four-plane chunks of the complex rotate kernel:
    a last chunk of 3 planes (its upper half holds one), exactly one chunk,
    and fewer than four planes (the two-plane kernel); x and z >= 32 take
    the 16-byte window copies."""
    sino, angles = create_test_sino_3d(A=9, Nx=nx, Ny=ny)
    kw = dict(res=3.0, nm=1.335, lD=0.5)
    ref = O.backpropagate_3d(sino, angles, **kw)
    got = G.backpropagate_3d(sino, angles, **kw)
    assert got.shape == ref.shape == (nx, ny, nx)
    assert rel_l2(got, ref) < FP64_REL_L2
    got32 = G.backpropagate_3d(sino, angles, dtype=np.float32, **kw)
    assert rel_l2(got32, ref) < FP32_REL_L2


@pytest.mark.parametrize("nx,ny", [(37, 23), (45, 31), (64, 29)])
def test_3d_unpadded_non_power_of_two_vs_oracle(nx, ny):
    """padding=(False, False): transforms of the detector lengths
    themselves (primes and composites that are not powers of two), as the
    reference's FFTW plans take them (_alg3d_bpp.py:460-478)."""
    sino, angles = create_test_sino_3d(A=11, Nx=nx, Ny=ny)
    kw = dict(res=3.0, nm=1.335, lD=0.5, padding=(False, False))
    ref = O.backpropagate_3d(sino, angles, **kw)
    got = G.backpropagate_3d(sino, angles, **kw)
    assert got.shape == ref.shape == (nx, ny, nx)
    assert rel_l2(got, ref) < FP64_REL_L2
    got32 = G.backpropagate_3d(sino, angles, dtype=np.float32, **kw)
    assert rel_l2(got32, ref) < FP32_REL_L2


@pytest.mark.parametrize("n", [101, 250])
def test_2d_unpadded_non_power_of_two_vs_oracle(n):
    sino, angles = create_test_sino_2d(N=n, A=17)
    kw = dict(res=3.0, nm=1.335, lD=0.5, padding=False)
    ref = O.backpropagate_2d(sino, angles, **kw)
    got = G.backpropagate_2d(sino, angles, **kw)
    assert got.shape == ref.shape == (n, n)
    assert rel_l2(got, ref) < FP64_REL_L2
    got32 = G.backpropagate_2d(sino, angles, dtype=np.float32, **kw)
    assert rel_l2(got32, ref) < FP32_REL_L2


def test_3d_iir_prefilter_equals_fused():
    sino, angles = create_test_sino_3d(A=11, Nx=20, Ny=16)
    for order in (2, 3):
        a = G.backpropagate_3d(sino, angles, 2.5, 1.333, intp_order=order)
        b = G.backpropagate_3d(sino, angles, 2.5, 1.333, intp_order=order,
                               prefilter="iir")
        assert rel_l2(a, b) < 1e-12


def test_3d_onlyreal_is_real_part():
    """tests/test_alg3d_bpp.py:77-98."""
    sino, angles = create_test_sino_3d(Nx=10, Ny=10)
    for p in get_test_parameter_set(2):
        c = G.backpropagate_3d(sino, angles, padval=0, dtype=float, **p)
        r = G.backpropagate_3d(sino, angles, padval=0, dtype=float,
                               onlyreal=True, **p)
        assert r.dtype == np.float64
        assert np.allclose(c.real, r)


def test_3d_multibatch_accumulation_equals_single(monkeypatch):
    """Slab budget below one batch forces several accumulate launches with
    the compensated sum carried in HBM."""
    sino, angles = create_test_sino_3d(A=17, Nx=16, Ny=12)
    one = G.backpropagate_3d(sino, angles, 2.2, 1.333, dtype=np.float32)
    monkeypatch.setenv("ODTBRAIN_AMD_SLAB_BUDGET", str(16 * 12 * 16 * 8 * 3))
    many = G.backpropagate_3d(sino, angles, 2.2, 1.333, dtype=np.float32)
    assert rel_l2(many, one) < 1e-6
    ref = O.backpropagate_3d(sino, angles, 2.2, 1.333)
    assert rel_l2(many, ref) < FP32_REL_L2


def test_3d_weights_even_and_copy_semantics():
    """tests/test_weighting.py:10-22 and tests/test_copy.py:9-21."""
    sino, angles = create_test_sino_3d()
    p = get_test_parameter_set(1)[0]
    f1 = G.backpropagate_3d(sino, angles, weight_angles=False, **p)
    f2 = G.backpropagate_3d(sino, angles, weight_angles=True, **p)
    assert np.allclose(f1, f2)
    s2 = sino.copy()
    f3 = G.backpropagate_3d(s2, angles, padval=0, copy=False, **p)
    f4 = G.backpropagate_3d(sino, angles, padval=0, copy=True, **p)
    assert np.allclose(f3, f4)


def test_3d_counters():
    """tests/test_counters.py:48-62."""
    sino, angles = create_test_sino_3d(Nx=10, Ny=10)
    jmc, jmm = mp.Value("i", 0), mp.Value("i", 0)
    G.backpropagate_3d(sino, angles, padval=0, dtype=float, count=jmc,
                       max_count=jmm, **get_test_parameter_set(1)[0])
    assert jmc.value == jmm.value != 0


def test_3d_device_tensor_in_and_out():
    sino, angles = create_test_sino_3d(A=7, Nx=12, Ny=12)
    ref = G.backpropagate_3d(sino, angles, 2.0, 1.333, dtype=np.float32)
    t = torch.from_numpy(sino.astype(np.complex64)).cuda()
    out = G.backpropagate_3d(t, angles, 2.0, 1.333, dtype=np.float32)
    assert isinstance(out, torch.Tensor) and out.is_cuda
    assert np.array_equal(out.cpu().numpy(), ref)


def test_3d_deterministic():
    sino, angles = create_test_sino_3d(A=9, Nx=16, Ny=16)
    a = G.backpropagate_3d(sino, angles, 2.0, 1.333, dtype=np.float32)
    b = G.backpropagate_3d(sino, angles, 2.0, 1.333, dtype=np.float32)
    assert np.array_equal(a, b)


def test_config1_sphere_vs_oracle():
    """64^2 detector, 100 angles, res 6, nm 1.333, sphere r=16 n=1.343,
    Rytov data with noise sigma 0.01, edge padding, order 2."""
    uB, angles = phantoms.sphere_sinogram_3d()
    kw = dict(res=6.0, nm=1.333, lD=0, padding=(True, True), padfac=1.75,
              padval="edge", intp_order=2)
    ref = O.backpropagate_3d(uB, angles, **kw)
    g64 = G.backpropagate_3d(uB, angles, **kw)
    g32 = G.backpropagate_3d(uB, angles, dtype=np.float32, **kw)
    e64, e32 = rel_l2(g64, ref), rel_l2(g32, ref)
    print(f"config1 rel_l2 fp64={e64:.3e} fp32={e32:.3e} "
          f"maxabs32={np.abs(g32 - ref).max():.3e}")
    assert e64 < FP64_REL_L2
    assert e32 < FP32_REL_L2


@pytest.mark.parametrize("axis", [[0, 1, 0], [0, np.cos(0.2), np.sin(0.2)],
                                  [0.3, 1, 0.2]])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_tilted_vs_oracle(axis, order):
    sino, angles = create_test_sino_3d(A=12, Nx=14, Ny=12)
    kw = dict(res=2.5, nm=1.333, tilted_axis=axis, intp_order=order,
              padval=0)
    ref = O.backpropagate_3d_tilted(sino, angles, **kw)
    got = G.backpropagate_3d_tilted(sino, angles, **kw)
    assert rel_l2(got, ref) < FP64_REL_L2
    got32 = G.backpropagate_3d_tilted(sino, angles, dtype=np.float32, **kw)
    assert rel_l2(got32, ref) < FP32_REL_L2


def test_tilted_y_axis_equals_plain():
    """tests/test_alg3d_bppt.py:10-50."""
    sino, angles = create_test_sino_3d()
    for p in get_test_parameter_set(2):
        a = cutout(G.backpropagate_3d(sino, angles, padval=0, dtype=float,
                                      onlyreal=True, **p))
        b = cutout(G.backpropagate_3d_tilted(sino, angles, padval=0,
                                             dtype=float, onlyreal=True, **p))
        assert np.allclose(a, b)


def test_tilted_points_input_and_iir():
    sino, angles = create_test_sino_3d(A=10, Nx=12, Ny=12)
    pts = O.sphere_points_from_angles_and_tilt(angles, [0, 1, 0.1])
    with pytest.warns(UserWarning):
        ref = O.backpropagate_3d_tilted(sino, pts.copy(), 2.5, 1.333,
                                        tilted_axis=[0, 1, 0.1])
    with pytest.warns(UserWarning):
        got = G.backpropagate_3d_tilted(sino, pts.copy(), 2.5, 1.333,
                                        tilted_axis=[0, 1, 0.1])
    assert rel_l2(got, ref) < FP64_REL_L2
    with pytest.warns(UserWarning):
        got2 = G.backpropagate_3d_tilted(sino, pts.copy(), 2.5, 1.333,
                                         tilted_axis=[0, 1, 0.1],
                                         prefilter="iir")
    assert rel_l2(got2, ref) < FP64_REL_L2


@pytest.mark.parametrize("A,N", [(43, 64), (16, 40)])
def test_2d_angle_groups_vs_oracle(A, N):
    """2D plans with enough angles run the rotate-accumulate in concurrent
    angle groups (43 angles on a 64^2 image: 5 groups of 9, the last of 7;
    16 angles: 2 groups of 8) folded in group order."""
    sino, angles = create_test_sino_2d(A=A, N=N, ampl_range=(0.9, 1.1))
    kw = dict(res=3.3, nm=1.34, lD=1.0)
    ref = O.backpropagate_2d(sino, angles, **kw)
    got = G.backpropagate_2d(sino, angles, **kw)
    assert rel_l2(got, ref) < FP64_REL_L2
    got32 = G.backpropagate_2d(sino, angles, dtype=np.float32, **kw)
    assert rel_l2(got32, ref) < FP32_REL_L2
    r = G.backpropagate_2d(sino, angles, onlyreal=True, **kw)
    assert np.allclose(r, got.real, rtol=0, atol=1e-12 * np.abs(got).max())


@pytest.mark.parametrize("padval", [0, None, 1.5])
@pytest.mark.parametrize("padding", [True, False])
def test_2d_vs_oracle(padval, padding):
    sino, angles = create_test_sino_2d(A=15, N=30, ampl_range=(0.9, 1.1))
    kw = dict(res=3.3, nm=1.34, lD=2.0, padding=padding, padval=padval)
    ref = O.backpropagate_2d(sino, angles, **kw)
    got = G.backpropagate_2d(sino, angles, **kw)
    assert got.shape == (30, 30) and got.dtype == np.complex128
    assert rel_l2(got, ref) < FP64_REL_L2
    got32 = G.backpropagate_2d(sino, angles, dtype=np.float32, **kw)
    assert rel_l2(got32, ref) < FP32_REL_L2
    r = G.backpropagate_2d(sino, angles, onlyreal=True, **kw)
    assert r.dtype == np.float64 and np.allclose(r, got.real)


def test_2d_angle_offset_and_swap_invariance():
    """tests/test_angle_weights.py:9-60."""
    sino, angles = create_test_sino_2d()
    p = get_test_parameter_set(1)[0]
    r1 = G.backpropagate_2d(sino, angles, weight_angles=False, **p)
    a2 = angles.copy()
    a2[::2] += 2 * np.pi * np.arange(a2[::2].shape[0])
    r2 = G.backpropagate_2d(sino, a2, weight_angles=False, **p)
    r3 = G.backpropagate_2d(sino, a2, weight_angles=True, **p)
    assert np.allclose(r1, r2) and np.allclose(r2, r3)
    s, a = sino[:-2], angles[:-2]
    o = np.argsort(a % .5)
    assert np.allclose(G.backpropagate_2d(s, a, **p),
                       G.backpropagate_2d(s[o], a[o], **p))


def test_2d_counters():
    sino, angles = create_test_sino_2d(N=10)
    jmc, jmm = mp.Value("i", 0), mp.Value("i", 0)
    G.backpropagate_2d(sino, angles, padval=0, count=jmc, max_count=jmm,
                       **get_test_parameter_set(1)[0])
    assert jmc.value == jmm.value != 0


def test_large_linearity_and_symmetry():
    """Size-independent properties at a 256^2 detector: linearity of the
    whole reconstruction (f(a u + b v) = a f(u) + b f(v)) in float32."""
    rng = np.random.default_rng(11)
    A, Nd = 24, 256
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    u = (rng.standard_normal((A, Nd, Nd)) * 0.01).astype(np.complex64)
    v = (1j * rng.standard_normal((A, Nd, Nd)) * 0.01).astype(np.complex64)
    kw = dict(res=6.0, nm=1.333, dtype=np.float32, padval="edge")
    fu = G.backpropagate_3d(u, angles, **kw)
    fv = G.backpropagate_3d(v, angles, **kw)
    fw = G.backpropagate_3d((2.0 * u - 3.0j * v).astype(np.complex64),
                            angles, **kw)
    assert rel_l2(fw, 2.0 * fu - 3.0j * fv) < 1e-5


def test_tilted_x_axis_equals_quarter_turned_y_axis():
    """Quarter-turned projections with the axis along -x reconstruct as the
    original projections with the y axis (the identity the oracle is pinned
    by), on the GPU in FP64 to rounding and in FP32 to the FP32 bar, and
    against the oracle."""
    sino, angles = quarter_turn_case()
    p = get_test_parameter_set(1)[0]
    turned = np.rot90(sino, k=1, axes=(1, 2)).copy()
    fy = G.backpropagate_3d_tilted(sino, angles, padval="edge",
                                   tilted_axis=[0, 1, 0], **p)
    fx = G.backpropagate_3d_tilted(turned, angles, padval="edge",
                                   tilted_axis=[-1, 0, 0], **p)
    assert rel_l2(fx, fy) < 1e-12
    ref = O.backpropagate_3d_tilted(turned, angles, padval="edge",
                                    tilted_axis=[-1, 0, 0], **p)
    assert rel_l2(fx, ref) < FP64_REL_L2
    fx32 = G.backpropagate_3d_tilted(turned, angles, padval="edge",
                                     tilted_axis=[-1, 0, 0],
                                     dtype=np.float32, **p)
    assert rel_l2(fx32, fy) < FP32_REL_L2


def test_tilted_counters():
    """tests/test_counters.py:80-93."""
    sino, angles = create_test_sino_3d(Nx=10, Ny=10)
    jmc, jmm = mp.Value("i", 0), mp.Value("i", 0)
    G.backpropagate_3d_tilted(sino, angles, padval=0, dtype=float,
                              count=jmc, max_count=jmm,
                              **get_test_parameter_set(1)[0])
    assert jmc.value == jmm.value != 0


@pytest.mark.parametrize("batch_angles,groups", [(20, None), (21, None),
                                                 (42, 3), (22, 2)])
def test_2d_angle_groups_across_slab_batches(monkeypatch, batch_angles,
                                             groups):
    """2D plans whose slab budget splits the 43 angles into batches: 20
    angles per batch (20, 20, 3), 21 (21, 21 and a single-angle last batch,
    which takes the ungrouped kernel), the group count overridden
    (ODTBRAIN_AMD_ROTATE_2D_GROUPS, read when a plan is built: each case
    has its own budget, hence its own plan).  Between batches the group fold reads
    and writes the running sum and its compensation; FP64 and FP32 against
    the oracle and against the single-batch run."""
    sino, angles = create_test_sino_2d(A=43, N=64, ampl_range=(0.9, 1.1))
    kw = dict(res=3.3, nm=1.34, lD=1.0)
    ref = O.backpropagate_2d(sino, angles, **kw)
    whole64 = G.backpropagate_2d(sino, angles, **kw)
    whole32 = G.backpropagate_2d(sino, angles, dtype=np.float32, **kw)
    for dt, esz in ((np.float64, 16), (np.float32, 8)):
        monkeypatch.setenv("ODTBRAIN_AMD_SLAB_BUDGET",
                           str(batch_angles * 64 * 64 * esz))
        if groups is not None:
            monkeypatch.setenv("ODTBRAIN_AMD_ROTATE_2D_GROUPS", str(groups))
        got = G.backpropagate_2d(sino, angles, dtype=dt, **kw)
        monkeypatch.delenv("ODTBRAIN_AMD_SLAB_BUDGET")
        monkeypatch.delenv("ODTBRAIN_AMD_ROTATE_2D_GROUPS", raising=False)
        if dt is np.float64:
            assert rel_l2(got, ref) < FP64_REL_L2
            assert rel_l2(got, whole64) < 1e-13
        else:
            assert rel_l2(got, ref) < FP32_REL_L2
            assert rel_l2(got, whole32) < 1e-6


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-7), (np.float64, 1e-14)])
def test_3d_overlapped_filter_chunks_equal_one_stream(monkeypatch, dtype, tol):
    """A T budget of a few angles splits the filter stage into chunks; with
    ODTBRAIN_AMD_OVERLAP=1 the row transforms of chunk c then run on the
    plan's second stream while the column transforms of chunk c + 1 fill
    the other T buffer.  The per-angle arithmetic does not depend on the
    chunking or the stream, so the result equals the one-stream plan's and
    the oracle's.  The env variables are read at plan creation: the plan
    cache is cleared around each variant."""
    from odtbrain_amd import _core
    sino, angles = create_test_sino_3d(A=23, Nx=24, Ny=18)
    kw = dict(res=3.0, nm=1.335, dtype=dtype)
    monkeypatch.setenv("ODTBRAIN_AMD_T_BUDGET", str(1 << 20))
    monkeypatch.setenv("ODTBRAIN_AMD_OVERLAP", "1")
    _core._PLAN_CACHE.clear()
    ov = G.backpropagate_3d(sino, angles, **kw)
    monkeypatch.delenv("ODTBRAIN_AMD_OVERLAP")
    _core._PLAN_CACHE.clear()
    one = G.backpropagate_3d(sino, angles, **kw)
    _core._PLAN_CACHE.clear()
    assert rel_l2(ov, one) < tol
    ref = O.backpropagate_3d(sino, angles, 3.0, 1.335)
    assert rel_l2(ov, ref) < (FP32_REL_L2 if dtype == np.float32 else 1e-9)
