"""Differentiable 3D backpropagation: the adjoint of backpropagate_3d and a
torch.autograd.Function around it.

backpropagate_3d is linear in the sinogram (edge or no padding): f = M u
with a complex matrix M.  ``backpropagate_3d_adjoint`` applies M^H on the
GPU (odt_plan_adjoint: transposed rotation gather, transposed B-spline
prefilter, forward transforms of every z-slice times the conjugate
propagator summed over z, conjugate prefactor, inverse transform, padding
folded back); ``Backpropagate3D`` uses the fast forward and this adjoint as
its backward (for a complex-linear map torch's gradient convention gives
grad_u = M^H grad_f).  The reference has no gradient; the checks are the
dot-product test <M u, g> = <u, M^H g> and finite differences.
"""
import numpy as np

from . import _core
from . import _native as N
from . import util
from ._alg3d_bpp import _pad_mode, _pads, _plan_fields, _precision


def _setup(A, lny, lnx, angles, res, nm, lD, weight_angles, padding, padfac,
           padval, intp_order, dtype, device, prefilter):
    dtype, cdt = _precision(dtype)
    precision = N.ODT_FP32 if dtype == np.float32 else N.ODT_FP64
    padyl, padxl, lNy, lNx = _pads(lny, lnx, padding, padfac)
    pad_mode, pv_re, pv_im = _pad_mode(padding, padval, lny, lnx, lNy, lNx)
    weights = util.compute_angle_weights_1d(angles) if weight_angles \
        else np.ones(A)
    fields, params = _plan_fields(
        angles, res, nm, lD, lny, lnx, padyl, padxl, lNy, lNx, pad_mode,
        pv_re, pv_im, intp_order, False, prefilter, precision, device)
    return fields, weights, params, cdt, pad_mode


def backpropagate_3d_adjoint(grad_f, angles, res, nm, lD=0,
                             weight_angles=True, padding=(True, True),
                             padfac=1.75, padval="edge", intp_order=2,
                             dtype=None, shape=None, prefilter=None):
    """M^H grad_f for M = backpropagate_3d(., angles, res, nm, ...).

    ``grad_f``: (Nx, Ny, Nx) complex volume (numpy or GPU tensor);
    ``shape``: sinogram shape (A, Ny, Nx) (default from grad_f and angles).
    Returns the (A, Ny, Nx) complex sinogram gradient (numpy for numpy in,
    GPU tensor for a GPU tensor in).  Edge or no padding only (the linear
    ramp padding is affine, not linear).  ``prefilter`` as for
    backpropagate_3d: None takes the forward's default ("fused" for spline
    orders 2 and 3, which shares the forward's plan and transposes its
    fused prefilter exactly; "iir" otherwise)."""
    torch = _core._torch()
    angles = np.asarray(angles)
    A = angles.size
    d_g, was_tensor = _core.device_input(grad_f, _precision(dtype)[1])
    if d_g.dim() != 3:
        raise ValueError("`grad_f` must have shape (Nx, Ny, Nx)")
    ln, lny, lnx = d_g.shape
    if shape is not None and tuple(shape) != (A, lny, lnx):
        raise ValueError("`shape` does not match `grad_f` and `angles`")
    if not (isinstance(padval, str) and padval == "edge") and \
            padding[0] and padding[1]:
        raise NotImplementedError("adjoint: padval='edge' or no padding")
    if prefilter is None:
        prefilter = "fused" if intp_order in (2, 3) else "iir"
    fields, weights, params, cdt, _ = _setup(
        A, lny, lnx, angles, res, nm, lD, weight_angles, padding, padfac,
        padval, intp_order, dtype, d_g.device, prefilter)
    plan = _core.get_plan(fields, weights, params, d_g.device)
    out = torch.empty((A, lny, lnx), dtype=d_g.dtype, device=d_g.device)
    plan.adjoint(d_g, out)
    return out if was_tensor else out.cpu().numpy()


def _function():
    torch = _core._torch()

    class Backpropagate3D(torch.autograd.Function):
        """f = backpropagate_3d(u, ...) with grad_u = M^H grad_f."""

        @staticmethod
        def forward(ctx, u, angles, res, nm, opts):
            from ._alg3d_bpp import backpropagate_3d
            ctx.args = (angles, res, nm, opts)
            return backpropagate_3d(u.detach(), angles, res, nm, **opts)

        @staticmethod
        def backward(ctx, grad_f):
            angles, res, nm, opts = ctx.args
            keep = ("lD", "weight_angles", "padding", "padfac", "padval",
                    "intp_order", "dtype")
            kw = {k: v for k, v in opts.items() if k in keep}
            g = backpropagate_3d_adjoint(grad_f.contiguous(), angles, res, nm,
                                         **kw)
            return g, None, None, None, None

    return Backpropagate3D


_FN = None


def backpropagate_3d_torch(u, angles, res, nm, **opts):
    """Differentiable backpropagate_3d for a complex GPU tensor ``u``
    (A, Ny, Nx); ``opts`` as for backpropagate_3d (no ``onlyreal``)."""
    global _FN
    if _FN is None:
        _FN = _function()
    if opts.get("onlyreal"):
        raise NotImplementedError("the differentiable path is complex")
    return _FN.apply(u, np.asarray(angles), res, nm, dict(opts))
