"""Missing-apple-core correction (odtbrain/apple.py), computed on the GPU
by odt_apple.hip through odt_apple_correct / odt_apple_core.

``correct`` has the reference's signature and counter semantics
(odtbrain/apple.py:124-262) and returns complex64 as the reference does
(numpy for numpy input, a GPU tensor for a GPU tensor); internally the
transforms run in complex128.  Both constraints of the reference are
provided: non-negativity (``method="nn"``, the default) and the symmetric
histogram (``method="sh"``).  ``envelope_gauss`` is the reference's
envelope function (odtbrain/apple.py:288-353) on the GPU.  Where the
envelope needs the nearest-sample fill (non-cubic volumes whose core meets
its transpose), equidistant candidates are resolved by the smallest linear
index, which may differ from scipy's choice.
"""
import ctypes

import numpy as np

from . import _core
from . import _native as N


def apple_core_3d(shape, res, nm):
    """Boolean mask of the apple core in Fourier space, y the axis of
    symmetry (odtbrain/apple.py:7-49)."""
    torch = _core._torch()
    n0, n1, n2 = (int(v) for v in shape)
    if n0 != n2:
        raise ValueError("`shape[0]` and `shape[2]` must be identical!")
    if not torch.cuda.is_available():
        raise N.OdtError("odtbrain_amd needs a ROCm GPU (torch.cuda) to run")
    d = torch.empty((n0, n1, n2), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream(d.device).cuda_stream
    N.check(N.lib().odt_apple_core(n0, n1, n2, float(res), float(nm),
                                   ctypes.c_void_p(d.data_ptr()),
                                   ctypes.c_void_p(stream)), "odt_apple_core")
    return d.cpu().numpy().astype(bool)


def correct(f, res, nm, method="nn", mask=None, bg_shell_width=None,
            enforce_envelope=0.95, max_iter=100, min_diff=.01,
            count=None, max_count=None):
    """Fill the missing apple core of the object function ``f`` (3D,
    shape[0] == shape[2]) by iterative projections
    (odtbrain/apple.py:124-262)."""
    torch = _core._torch()
    if enforce_envelope < 0 or enforce_envelope > 1:
        raise ValueError("`enforce_envelope` must be in interval [0, 1]")
    if method not in ("nn", "sh"):
        raise ValueError("`method` must be 'nn' or 'sh'")
    if max_count is not None:
        with max_count.get_lock():
            max_count.value += max_iter + 2
    was_tensor = isinstance(f, torch.Tensor) and f.is_cuda
    if isinstance(f, torch.Tensor):
        t = f.cuda() if not f.is_cuda else f
    else:
        arr = np.asarray(f)
        if not torch.cuda.is_available():
            raise N.OdtError("odtbrain_amd needs a ROCm GPU (torch.cuda) to run")
        t = torch.from_numpy(np.ascontiguousarray(arr))
        t = t.cuda()
    if not t.is_complex():
        t = t.to(torch.complex128 if t.dtype == torch.float64
                 else torch.complex64)
    t = t.contiguous()
    if t.dim() != 3:
        raise ValueError("`f` must be a 3D array")
    n0, n1, n2 = t.shape
    if n0 != n2:
        raise ValueError("`shape[0]` and `shape[2]` must be identical!")
    prec = N.ODT_FP32 if t.dtype == torch.complex64 else N.ODT_FP64
    d_mask = None
    if mask is not None:
        m = torch.as_tensor(np.asarray(mask, dtype=bool)) \
            if not isinstance(mask, torch.Tensor) else mask
        d_mask = torch.broadcast_to(m.to(device=t.device, dtype=torch.uint8),
                                    t.shape).contiguous()
    L = N.lib()
    size = ctypes.c_size_t()
    N.check(L.odt_apple_workspace_size(n0, n1, n2, ctypes.byref(size)),
            "odt_apple_workspace_size")
    ws = torch.empty(size.value, dtype=torch.uint8, device=t.device)
    out = torch.empty(t.shape, dtype=torch.complex64, device=t.device)
    iters = ctypes.c_int32(0)
    stream = torch.cuda.current_stream(t.device).cuda_stream
    if count is not None:
        with count.get_lock():
            count.value += 2
    with torch.cuda.device(t.device):
        rc = L.odt_apple_correct(
            prec, n0, n1, n2, ctypes.c_void_p(t.data_ptr()), float(res),
            float(nm), 0 if method == "nn" else 1,
            ctypes.c_void_p(d_mask.data_ptr() if d_mask is not None else None),
            -1.0 if bg_shell_width is None else float(bg_shell_width),
            float(enforce_envelope), int(max_iter), float(min_diff), N.ODT_FP32,
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
            size.value, ctypes.c_void_p(stream), ctypes.byref(iters))
    N.check(rc, "odt_apple_correct")
    if count is not None:
        with count.get_lock():
            count.value += max_iter  # iterations run plus the skipped ones
    del ws
    return out if was_tensor else out.cpu().numpy()


def envelope_gauss(ftdata, core):
    """Gaussian-filtered envelope of the Fourier amplitudes without the
    apple core (odtbrain/apple.py:288-353): |ftdata| with the core filled
    from the (1, 0, 2)-transposed data and remaining gaps from the nearest
    valid sample, blurred with sigma = max(shape)/100.  ``ftdata`` (not
    shifted) and ``core`` have the shape (n0, n1, n0); returns float64
    (numpy for numpy input, a GPU tensor for a GPU tensor)."""
    torch = _core._torch()
    was_tensor = isinstance(ftdata, torch.Tensor) and ftdata.is_cuda
    if not torch.cuda.is_available():
        raise N.OdtError("odtbrain_amd needs a ROCm GPU (torch.cuda) to run")
    if isinstance(ftdata, torch.Tensor):
        t = ftdata.cuda()
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(ftdata))).cuda()
    t = t.to(torch.complex128).contiguous()
    if t.dim() != 3:
        raise ValueError("`ftdata` must be a 3D array")
    n0, n1, n2 = t.shape
    if n0 != n2:
        raise ValueError("`shape[0]` and `shape[2]` must be identical!")
    c = core if isinstance(core, torch.Tensor) else \
        torch.from_numpy(np.ascontiguousarray(np.asarray(core, dtype=bool)))
    c = c.to(device=t.device, dtype=torch.uint8).contiguous()
    if tuple(c.shape) != tuple(t.shape):
        raise ValueError("`core` must have the shape of `ftdata`")
    L = N.lib()
    size = ctypes.c_size_t()
    N.check(L.odt_apple_workspace_size(n0, n1, n2, ctypes.byref(size)),
            "odt_apple_workspace_size")
    ws = torch.empty(size.value, dtype=torch.uint8, device=t.device)
    env = torch.empty(t.shape, dtype=torch.float64, device=t.device)
    stream = torch.cuda.current_stream(t.device).cuda_stream
    N.check(L.odt_apple_envelope(n0, n1, n2, ctypes.c_void_p(t.data_ptr()),
                                 ctypes.c_void_p(c.data_ptr()),
                                 ctypes.c_void_p(env.data_ptr()),
                                 ctypes.c_void_p(ws.data_ptr()), size.value,
                                 ctypes.c_void_p(stream)),
            "odt_apple_envelope")
    return env if was_tensor else env.cpu().numpy()
