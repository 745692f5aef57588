"""The four GPU workloads at their full sizes, checked against the CPU
oracle (the restatement of the reference in complex128) on the same seeded
inputs.

* 256^2 detector, 360 angles (3-ellipsoid Rytov phantom, complex64 input,
  512^2 padding, 256^3 output): FP32 and FP64 against the oracle run in
  full with the reference's y-chunk process pool.
* tilted axis (0.2 rad about x), 192^2 detector, 240 jittered angles over
  1.4 pi, weight_angles, order 2 (5 seeded spheres): FP32 and FP64 against
  the oracle, its angle loop split over the host cores.
* the 48-sinogram batch (32 backpropagate_2d, N 128..1024, 90-720 angles;
  16 backpropagate_3d, N 64-160, 60-300 angles) as one HIP graph: every
  replay equals the eager calls bit for bit, every instance's FP32 result is
  within 1e-5 of its FP64 result, and one instance per size class matches
  the oracle.
* the differentiable reconstruction, 128^2 detector, 180 angles: forward
  against the oracle, the dot-product test <g, M u> = <M^H g, u>, and FP64
  central differences (eps 1e-6) of ||BP(u) - f_t||^2 + 1e-3 TV(BP(u)) on
  64 seeded directions.

Tolerances: FP32 relative L2 <= 1e-5 against the complex128 oracle (the
accuracy bar of this project); FP64 relative L2 <= 1e-9.  Each case prints
its relative L2 and max abs error.
"""
import os
import time

import numpy as np
import pytest

import odtbrain_amd as G
from common import rel_l2
from odtbrain_amd import autograd as AG
from odtbrain_amd import phantoms
from oracle import odt_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FP32_REL_L2 = 1e-5
FP64_REL_L2 = 1e-9
RES, NM = 6.0, 1.333
# the host's CPU share (16 cores per GPU on the MI355X hosts)
WORKERS = max(1, min(16, os.cpu_count() or 1))


def _report(name, got, ref):
    got = np.asarray(got)
    err = rel_l2(got, ref)
    mx = float(np.abs(got.astype(ref.dtype) - ref).max())
    print(f"{name}: rel_l2={err:.3e} max_abs={mx:.3e} "
          f"(max |ref| {np.abs(ref).max():.3e})", flush=True)
    return err


@pytest.mark.timeout(900)
def test_256_detector_360_angles_vs_oracle():
    uB, angles = phantoms.ellipsoid_sinogram_3d(N=256, A=360)
    kw = dict(padval="edge", intp_order=2)
    d_in = torch.from_numpy(uB).cuda()
    g32 = G.backpropagate_3d(d_in, angles, RES, NM, dtype=np.float32, **kw)
    g64 = G.backpropagate_3d(d_in.to(torch.complex128), angles, RES, NM,
                             dtype=np.float64, **kw)
    g32, g64 = g32.cpu().numpy(), g64.cpu().numpy()
    t0 = time.perf_counter()
    ref = O.backpropagate_3d(uB.astype(np.complex128), angles, RES, NM,
                             num_cores=WORKERS, **kw)
    print(f"oracle 256^2 x 360: {time.perf_counter() - t0:.1f} s on "
          f"{WORKERS} cores", flush=True)
    assert g32.shape == ref.shape == (256, 256, 256)
    assert _report("256^2x360 fp64", g64, ref) < FP64_REL_L2
    assert _report("256^2x360 fp32", g32, ref) < FP32_REL_L2


@pytest.mark.timeout(900)
def test_tilted_192_detector_240_angles_vs_oracle():
    uB, angles, axis = phantoms.tilted_sphere_sinogram()
    kw = dict(tilted_axis=axis, weight_angles=True, intp_order=2)
    d_in = torch.from_numpy(uB).cuda()
    g32 = G.backpropagate_3d_tilted(d_in, angles, RES, NM, dtype=np.float32,
                                    **kw).cpu().numpy()
    g64 = G.backpropagate_3d_tilted(d_in.to(torch.complex128), angles, RES,
                                    NM, dtype=np.float64,
                                    **kw).cpu().numpy()
    t0 = time.perf_counter()
    ref = O.angle_parallel(O.backpropagate_3d_tilted,
                           uB.astype(np.complex128), angles, RES, NM,
                           workers=WORKERS, **kw)
    print(f"tilted oracle 192^2 x 240: {time.perf_counter() - t0:.1f} s on "
          f"{WORKERS} processes", flush=True)
    assert g32.shape == ref.shape == (192, 192, 192)
    assert _report("tilted fp64", g64, ref) < FP64_REL_L2
    assert _report("tilted fp32", g32, ref) < FP32_REL_L2


def _size_class_representatives(raw):
    """Per size class (2D: each detector size; 3D: the smallest and the
    largest detector) the instance with the fewest angles."""
    best = {}
    for i, (kind, s, ang, _r, _n) in enumerate(raw):
        N = s.shape[-1]
        if kind == "2d":
            key = ("2d", N)
        else:
            key = ("3d", "small" if N <= 100 else "large")
        if key not in best or len(ang) < len(raw[best[key]][2]):
            best[key] = i
    return sorted(best.values())


@pytest.mark.timeout(900)
def test_batch_48_sinograms_graph_vs_eager_fp64_and_oracle():
    raw = phantoms.batch_jobs()
    fns = {"2d": G.backpropagate_2d, "3d": G.backpropagate_3d}
    jobs, ins = [], []
    for kind, s, ang, res, nm in raw:
        d = torch.from_numpy(s).cuda()
        ins.append(d)
        jobs.append((fns[kind], (d, ang, res, nm), {"dtype": np.float32}))
    eager = [fn(*a, **kw).cpu().numpy() for fn, a, kw in jobs]
    b = G.Batch(jobs)
    for rep in range(2):
        outs = b.run()
        torch.cuda.synchronize()
        for i, (o, e) in enumerate(zip(outs, eager)):
            assert np.array_equal(o.cpu().numpy(), e), (rep, i)
        # eager work of the program between replays
        G.backpropagate_2d(raw[0][1], raw[0][2], RES, NM, dtype=np.float32)
        torch.cuda.synchronize()
    worst = 0.0
    f64 = []
    for (fn, a, _kw), e, d in zip(jobs, eager, ins):
        r = fn(d.to(torch.complex128), *a[1:], dtype=np.float64)
        r = r.cpu().numpy()
        f64.append(r)
        err = rel_l2(e, r)
        worst = max(worst, err)
        assert err < FP32_REL_L2
    print(f"batch: max FP32-vs-FP64 rel_l2 over 48 instances {worst:.3e}",
          flush=True)
    for i in _size_class_representatives(raw):
        kind, s, ang, res, nm = raw[i]
        f = O.backpropagate_2d if kind == "2d" else O.backpropagate_3d
        ref = O.angle_parallel(f, s.astype(np.complex128), ang, res, nm,
                               workers=WORKERS)
        name = f"batch #{i} {kind} N={s.shape[-1]} A={len(ang)}"
        assert _report(name + " fp64", f64[i], ref) < FP64_REL_L2
        assert _report(name + " fp32", eager[i], ref) < FP32_REL_L2


def _tv(f, eps=1e-3):
    """Smoothed isotropic total variation (forward differences)."""
    d0 = f[1:, :-1, :-1] - f[:-1, :-1, :-1]
    d1 = f[:-1, 1:, :-1] - f[:-1, :-1, :-1]
    d2 = f[:-1, :-1, 1:] - f[:-1, :-1, :-1]
    return torch.sum(torch.sqrt(d0.abs() ** 2 + d1.abs() ** 2 +
                                d2.abs() ** 2 + eps))


@pytest.mark.timeout(900)
def test_differentiable_128_detector_180_angles():
    N, A, lam = 128, 180, 1e-3
    sino, angles = phantoms.ellipsoid_sinogram_3d(
        N=N, A=A, count=4, semi=(10, 30), spread=15.0, seed=5)
    d_sino = torch.from_numpy(sino).cuda()
    # forward against the oracle
    g32 = G.backpropagate_3d(d_sino, angles, RES, NM, dtype=np.float32)
    g64 = G.backpropagate_3d(d_sino.to(torch.complex128), angles, RES, NM,
                             dtype=np.float64)
    ref = O.angle_parallel(O.backpropagate_3d, sino.astype(np.complex128),
                           angles, RES, NM, workers=WORKERS)
    assert _report("128^2x180 fp64", g64.cpu().numpy(), ref) < FP64_REL_L2
    assert _report("128^2x180 fp32", g32.cpu().numpy(), ref) < FP32_REL_L2

    gen = torch.Generator(device="cuda").manual_seed(7)
    # dot-product test of the adjoint
    for ct, dt, tol in ((torch.complex128, np.float64, 1e-12),
                        (torch.complex64, np.float32, 2e-5)):
        a = torch.randn(A, N, N, dtype=ct, device="cuda", generator=gen)
        bv = torch.randn(N, N, N, dtype=ct, device="cuda", generator=gen)
        Ma = G.backpropagate_3d(a, angles, RES, NM, dtype=dt)
        MHb = AG.backpropagate_3d_adjoint(bv, angles, RES, NM, dtype=dt)
        lhs = complex(torch.sum(torch.conj(bv).to(torch.complex128) *
                                Ma.to(torch.complex128)))
        rhs = complex(torch.sum(torch.conj(MHb).to(torch.complex128) *
                                a.to(torch.complex128)))
        rel = abs(lhs - rhs) / abs(lhs)
        print(f"dot-product test {dt.__name__}: {rel:.3e}", flush=True)
        assert rel < tol

    # FP64 central differences of the regularised loss on 64 directions
    f_t = g64

    def loss(u):
        f = AG.backpropagate_3d_torch(u, angles, RES, NM, dtype=np.float64)
        return torch.sum((f - f_t).abs() ** 2) + lam * _tv(f)

    u0 = (0.5 * d_sino.to(torch.complex128)).requires_grad_(True)
    loss(u0).backward()
    grad = u0.grad.detach()
    u0 = u0.detach()
    eps, errs = 1e-6, []
    with torch.no_grad():
        for _ in range(64):
            d = torch.complex(
                torch.randn(A, N, N, dtype=torch.float64, device="cuda",
                            generator=gen),
                torch.randn(A, N, N, dtype=torch.float64, device="cuda",
                            generator=gen))
            fd = (float(loss(u0 + eps * d)) - float(loss(u0 - eps * d))) / \
                (2 * eps)
            an = float(torch.sum(torch.real(torch.conj(grad) * d)))
            errs.append(abs(fd - an) / max(abs(an), 1e-300))
    print(f"central differences on 64 directions: median rel err "
          f"{np.median(errs):.3e}, max {max(errs):.3e}", flush=True)
    assert np.median(errs) < 1e-6
    assert max(errs) < 1e-4
