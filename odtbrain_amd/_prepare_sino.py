"""Sinogram preparation (odtbrain/_prepare_sino.py): the Rytov or plain
phase sinogram of a complex field, computed by the HIP kernels of
odt_prepare.hip through odt_sinogram_prepare.

Same signatures and results as the reference.  2D sinograms are unwrapped
along the detector axis with numpy.unwrap's rule; 3D sinograms image by
image with the reliability-sorted algorithm that
skimage.restoration.unwrap_phase implements (Herraez et al., Applied Optics
41 (2002) 7437).  Two details of that algorithm are fixed here rather than
random: border pixels get reliability 1e7 and equal edge reliabilities are
taken in creation order.  For phase fields without residues the result
equals any correct unwrapping up to whole turns per image, which alignment
removes.
"""
import ctypes

import numpy as np

from . import _core
from . import _native as N


def _run(uSin, u0, rytov, align):
    torch = _core._torch()
    was_tensor = isinstance(uSin, torch.Tensor)
    if was_tensor:
        t = uSin if uSin.is_cuda else uSin.cuda()
        cplx = t.dtype in (torch.complex64, torch.complex128)
        tdt = torch.complex64 if t.dtype in (torch.complex64, torch.float32) \
            else torch.complex128
        t = t.to(tdt).contiguous()
        was_tensor = uSin.is_cuda
    else:
        arr = np.asarray(uSin)
        cdt = np.complex64 if arr.dtype in (np.complex64, np.float32) \
            else np.complex128
        if not torch.cuda.is_available():
            raise N.OdtError("odtbrain_amd needs a ROCm GPU (torch.cuda) to run")
        t = torch.from_numpy(np.ascontiguousarray(arr, dtype=cdt)).cuda()
    ndim = t.dim()
    if ndim not in (2, 3):
        raise ValueError("`uSin` must be a 2D or 3D sinogram")
    A = t.shape[0]
    Ny, Nx = (t.shape[1], t.shape[2]) if ndim == 3 else (1, t.shape[1])
    prec = N.ODT_FP32 if t.dtype == torch.complex64 else N.ODT_FP64
    d_u0 = None
    if not (np.isscalar(u0) and u0 == 1):
        if isinstance(u0, torch.Tensor):
            u0t = u0.to(device=t.device, dtype=t.dtype)
        else:
            u0t = torch.as_tensor(np.asarray(u0), device=t.device).to(t.dtype)
        d_u0 = torch.broadcast_to(u0t, t.shape).contiguous()
    L = N.lib()
    size = ctypes.c_size_t()
    N.check(L.odt_prepare_workspace_size(prec, ndim, A, Ny, Nx, int(rytov),
                                         ctypes.byref(size)),
            "odt_prepare_workspace_size")
    ws = torch.empty(max(1, size.value), dtype=torch.uint8, device=t.device)
    if rytov:
        out = torch.empty_like(t)
    else:
        out = torch.empty(t.shape, device=t.device,
                          dtype=torch.float32 if prec == N.ODT_FP32
                          else torch.float64)
    stream = torch.cuda.current_stream(t.device).cuda_stream
    with torch.cuda.device(t.device):
        rc = L.odt_sinogram_prepare(
            prec, ndim, A, Ny, Nx, ctypes.c_void_p(t.data_ptr()),
            ctypes.c_void_p(d_u0.data_ptr() if d_u0 is not None else None),
            int(rytov), int(bool(align)), ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(ws.data_ptr()), size.value,
            ctypes.c_void_p(stream))
    N.check(rc, "odt_sinogram_prepare")
    if was_tensor:
        return out
    return out.cpu().numpy()


def sinogram_as_radon(uSin, align=True):
    """Unwrapped phase of a complex sinogram (A, N) or (A, Ny, Nx)
    (odtbrain/_prepare_sino.py:60-102)."""
    return _run(uSin, 1, False, align)


def sinogram_as_rytov(uSin, u0=1, align=True):
    """Rytov sinogram u0 * (log|uSin/u0| + i * unwrapped angle(uSin/u0))
    (odtbrain/_prepare_sino.py:105-182)."""
    return _run(uSin, u0, True, align)
