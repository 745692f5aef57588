"""Adjoint of backpropagate_3d (odt_plan_adjoint) and the autograd
Function: dot-product test <g, M u> = <M^H g, u> in FP64 and FP32 for all
supported spline orders and paddings, central finite differences of a
TV-regularised loss in FP64, and a few gradient-descent steps decreasing
that loss."""
import numpy as np
import pytest

import odtbrain_amd as G
from odtbrain_amd import autograd as AG

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _rand(shape, dt, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return torch.from_numpy(x.astype(dt)).cuda()


def _dot(a, b):
    return complex(torch.sum(torch.conj(a.to(torch.complex128)) *
                             b.to(torch.complex128)).item())


@pytest.mark.parametrize("order", [2, 3, 4, 5])
@pytest.mark.parametrize("case", ["edge", "nopad", "noweight"])
def test_adjoint_dot_product_fp64(order, case):
    if case == "nopad":
        A, ny, nx = 7, 64, 64
        kw = dict(padding=(False, False))
    else:
        A, ny, nx = 11, 20, 24
        kw = dict()
    if case == "noweight":
        kw["weight_angles"] = False
    rng = np.random.default_rng(order)
    angles = np.sort(rng.uniform(0, 2 * np.pi, A))
    u = _rand((A, ny, nx), np.complex128, 10 + order)
    g = _rand((nx, ny, nx), np.complex128, 20 + order)
    Mu = G.backpropagate_3d(u, angles, 6.0, 1.333, intp_order=order,
                            dtype=np.float64, **kw)
    MHg = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333,
                                      intp_order=order, dtype=np.float64,
                                      **kw)
    lhs, rhs = _dot(g, Mu), _dot(MHg, u)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("pad", [True, False])
def test_fused_adjoint_equals_explicit_prefilter_adjoint(order, pad):
    """The transposed fused prefilter (boundary terms in the x transform,
    conjugate fused propagator in the z sum) against the transposed
    recursive prefilter (FIR convolutions): the same operator."""
    A, ny, nx = (9, 20, 24) if pad else (7, 64, 64)
    kw = {} if pad else dict(padding=(False, False))
    angles = np.linspace(0.3, 2 * np.pi + 0.3, A, endpoint=False)
    g = _rand((nx, ny, nx), np.complex128, 30 + order)
    a = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, intp_order=order,
                                    dtype=np.float64, prefilter="fused", **kw)
    b = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, intp_order=order,
                                    dtype=np.float64, prefilter="iir", **kw)
    err = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    assert err < 1e-10, err


def test_adjoint_dot_product_fp32():
    A, ny, nx = 16, 30, 32
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    u = _rand((A, ny, nx), np.complex64, 1)
    g = _rand((nx, ny, nx), np.complex64, 2)
    Mu = G.backpropagate_3d(u, angles, 6.0, 1.333, dtype=np.float32)
    MHg = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333,
                                      dtype=np.float32)
    lhs, rhs = _dot(g, Mu), _dot(MHg, u)
    assert abs(lhs - rhs) <= 2e-5 * abs(lhs), (lhs, rhs)


def _tv(f, eps=1e-3):
    dz = f[1:, :-1, :-1] - f[:-1, :-1, :-1]
    dy = f[:-1, 1:, :-1] - f[:-1, :-1, :-1]
    dx = f[:-1, :-1, 1:] - f[:-1, :-1, :-1]
    return torch.sum(torch.sqrt(dz.abs() ** 2 + dy.abs() ** 2 +
                                dx.abs() ** 2 + eps))


def _loss(u, angles, f_t, lam, dtype):
    f = AG.backpropagate_3d_torch(u, angles, 6.0, 1.333, dtype=dtype)
    return torch.sum((f - f_t).abs() ** 2) + lam * _tv(f)


def test_gradient_central_differences_fp64():
    A, ny, nx = 10, 16, 18
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    u0 = _rand((A, ny, nx), np.complex128, 5) * 0.1
    f_t = _rand((nx, ny, nx), np.complex128, 6) * 0.01
    u = u0.clone().requires_grad_(True)
    L = _loss(u, angles, f_t, 1e-3, np.float64)
    L.backward()
    grad = u.grad.detach()
    eps = 1e-6
    worst = 0.0
    for k in range(8):
        d = _rand((A, ny, nx), np.complex128, 100 + k)
        with torch.no_grad():
            lp = _loss(u0 + eps * d, angles, f_t, 1e-3, np.float64).item()
            lm = _loss(u0 - eps * d, angles, f_t, 1e-3, np.float64).item()
        fd = (lp - lm) / (2 * eps)
        an = float(torch.sum(torch.real(torch.conj(grad) * d)).item())
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-30))
    assert worst < 1e-6, worst


def test_gradient_descent_decreases_loss_fp32():
    """Steepest descent with an Armijo backtracking step: every step must
    satisfy the sufficient-decrease condition with a step no smaller than
    2^-30 (a wrong gradient direction fails that), and eight steps must
    remove most of the data misfit of a zero start."""
    from odtbrain_amd import phantoms
    sino, angles = phantoms.sphere_sinogram_3d(N=32, A=24, radius=8.0)
    d_sino = torch.from_numpy(sino.astype(np.complex64)).cuda()
    f_t = G.backpropagate_3d(d_sino, angles, 6.0, 1.333, dtype=np.float32)
    u = torch.zeros_like(d_sino).requires_grad_(True)
    losses = []
    t = 1.0
    for _ in range(8):
        L = _loss(u, angles, f_t, 1e-4, np.float32)
        losses.append(L.item())
        u.grad = None
        L.backward()
        g = u.grad.detach()
        gg = float(torch.sum(g.abs() ** 2).item())
        with torch.no_grad():
            for _ in range(30):
                lt = _loss(u - t * g, angles, f_t, 1e-4, np.float32).item()
                if lt <= losses[-1] - 1e-4 * t * gg:
                    break
                t *= 0.5
            else:
                raise AssertionError("no descent step along -grad")
            u -= t * g
        t *= 2.0
    assert all(b < a for a, b in zip(losses, losses[1:])), losses
    assert losses[-1] < 0.5 * losses[0], losses


@pytest.mark.parametrize("budget_angles", [1, 3, 5])
def test_adjoint_slab_budgets_agree(monkeypatch, budget_angles):
    """A slab budget of one angle runs the transposed prefilter as
    recursions in place; larger ones as convolutions over half batches
    (several chunks here).  All give the same adjoint (FP64 to 1e-12) and
    pass the dot-product test."""
    A, ny, nx = 9, 20, 24
    rng = np.random.default_rng(11)
    angles = np.sort(rng.uniform(0, 2 * np.pi, A))
    u = _rand((A, ny, nx), np.complex128, 12)
    g = _rand((nx, ny, nx), np.complex128, 13)
    full = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, dtype=np.float64)
    monkeypatch.setenv("ODTBRAIN_AMD_SLAB_BUDGET",
                       str(budget_angles * nx * ny * nx * 16))
    part = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333,
                                       dtype=np.float64)
    err = float(torch.linalg.vector_norm(part - full) /
                torch.linalg.vector_norm(full))
    assert err < 1e-12, err
    Mu = G.backpropagate_3d(u, angles, 6.0, 1.333, dtype=np.float64)
    lhs, rhs = _dot(g, Mu), _dot(part, u)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("dtype,orders,tol", [(np.float64, (2, 3), 1e-13),
                                              (np.float32, (2, 3, 4, 5), 2e-6)])
@pytest.mark.parametrize("ny", [20, 18])
def test_adjoint_transposed_gather_equals_direct(monkeypatch, dtype, orders,
                                                 tol, ny):
    """The rotate transpose from the planes-innermost copy of the gradient
    with a per-block table of forward taps (k_adj_rotate_t) and the direct
    gather (k_adj_rotate, ODTBRAIN_AMD_ADJ_DIRECT) agree (their spline
    weights come from two formulas of the same polynomials), and the
    transposed one passes the dot-product test.  Ny = 18 is not a multiple
    of 4: both runs take the direct gather."""
    A, nx = 11, 24
    rng = np.random.default_rng(21)
    angles = np.sort(rng.uniform(0, 2 * np.pi, A))
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    u = _rand((A, ny, nx), cdt, 22)
    g = _rand((nx, ny, nx), cdt, 23)
    for order in orders:
        kw = dict(dtype=dtype, intp_order=order)
        t = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, **kw)
        monkeypatch.setenv("ODTBRAIN_AMD_ADJ_DIRECT", "1")
        d = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, **kw)
        monkeypatch.delenv("ODTBRAIN_AMD_ADJ_DIRECT")
        err = float(torch.linalg.vector_norm((t - d).to(torch.complex128)) /
                    torch.linalg.vector_norm(d.to(torch.complex128)))
        assert err < tol, (order, err)
        Mu = G.backpropagate_3d(u, angles, 6.0, 1.333, **kw)
        lhs, rhs = _dot(g, Mu), _dot(t, u)
        bar = 1e-12 if dtype == np.float64 else 1e-5
        assert abs(lhs - rhs) <= bar * abs(lhs), (order, lhs, rhs)
