"""Object function to refractive index (odtbrain/_translate_ri.py), computed
by the HIP kernels of odt_translate.hip through odt_odt_to_ri /
odt_opt_to_ri.

Same signatures and results as the reference: numpy in -> numpy out (real
in -> real out, a negative radicand gives NaN as numpy's real sqrt; complex
in -> the principal root, whose real part is never negative, so the
reference's sign fix has nothing to do); a torch tensor on the GPU in -> a
tensor on the same device out, computed on the current stream.
"""
import ctypes

import numpy as np

from . import _core
from . import _native as N


def _prepare(f):
    """(device tensor, precision, is_complex, was_tensor, numpy dtype)."""
    torch = _core._torch()
    if isinstance(f, torch.Tensor):
        t, was = f, True
        if not t.is_cuda:
            t, was = t.cuda(), False
        if t.dtype in (torch.complex64, torch.complex128):
            cplx = True
        elif t.dtype in (torch.float32, torch.float64):
            cplx = False
        else:
            t, cplx = t.to(torch.float64), False
        t = t.contiguous()
        prec = N.ODT_FP32 if t.dtype in (torch.float32, torch.complex64) \
            else N.ODT_FP64
        return t, prec, cplx, was, None
    arr = np.asarray(f)
    if arr.dtype.kind == "c":
        dt = np.complex64 if arr.dtype == np.complex64 else np.complex128
    elif arr.dtype == np.float32:
        dt = np.float32
    else:
        dt = np.float64
    arr = np.ascontiguousarray(arr, dtype=dt)
    if not torch.cuda.is_available():
        raise N.OdtError("odtbrain_amd needs a ROCm GPU (torch.cuda) to run")
    t = torch.from_numpy(arr).cuda()
    prec = N.ODT_FP32 if dt in (np.float32, np.complex64) else N.ODT_FP64
    return t, prec, arr.dtype.kind == "c", False, dt


def _apply(entry, f, res, nm):
    torch = _core._torch()
    t, prec, cplx, was_tensor, _ = _prepare(f)
    out = torch.empty_like(t)
    stream = torch.cuda.current_stream(t.device).cuda_stream
    with torch.cuda.device(t.device):
        rc = getattr(N.lib(), entry)(prec, int(cplx), t.numel(), float(res),
                                     float(nm), ctypes.c_void_p(t.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()),
                                     ctypes.c_void_p(stream))
    N.check(rc, entry)
    if was_tensor:
        return out
    return out.cpu().numpy()


def odt_to_ri(f, res, nm):
    """Refractive index n = nm * sqrt(f / km^2 + 1), km = 2 pi nm / res, of
    the ODT object function ``f`` (any shape), with the positive root
    (odtbrain/_translate_ri.py:5-57)."""
    return _apply("odt_odt_to_ri", f, res, nm)


def opt_to_ri(f, res, nm):
    """Refractive index n = nm + f * res / (2 pi) of an OPT object function
    (odtbrain/_translate_ri.py:60-90)."""
    return _apply("odt_opt_to_ri", f, res, nm)
