"""Host planner and executor shared by the three backpropagation functions.

The planner derives everything the reference derives on the host in
float64 (pad sizes, angle weights, per-angle resampling geometry) and hands
it to libodtbrain_hip through the C ABI.  Device buffers are torch tensors;
torch is used for memory and streams only.
"""
import collections
import ctypes
import os

import numpy as np
import scipy.special

from . import _native as N

_PLAN_CACHE = collections.OrderedDict()
_PLAN_CACHE_SIZE = 8
last_plan = None   # plan of the most recent call (timing queries)


def _torch():
    import torch
    return torch


def padded_order(n, padfac):
    """``int(max(64, 2**ceil(log(n*padfac)/log 2)))`` as in
    odtbrain/_alg3d_bpp.py:286 (same float expression)."""
    return int(max(64., 2 ** np.ceil(np.log(n * padfac) / np.log(2))))


_PARAMS_CACHE = collections.OrderedDict()


def planar_params(angles_deg, nz, nx):
    """Matrix and offset of ``scipy.ndimage.rotate(angle, axes=(0, 1 or 2),
    reshape=False)`` for each angle in degrees: input = M @ (o_z, o_x) + off.

    Built exactly as scipy's rotate builds them (scipy/ndimage/
    _interpolation.py: ``c, s = special.cosdg(angle), special.sindg(angle)``,
    ``rot = [[c, s], [-s, c]]``, ``offset = in_center - rot @ out_center``)
    so the float64 bits agree with the reference call.  Memoised on the
    angles and the plane shape (repeated calls with the same angles, as in
    iterative reconstructions, skip the per-angle loop); the returned array
    is read-only.
    """
    a = np.ascontiguousarray(angles_deg, dtype=np.float64).reshape(-1)
    key = (a.tobytes(), int(nz), int(nx))
    out = _PARAMS_CACHE.get(key)
    if out is None:
        out = _planar_params(a, nz, nx)
        out.flags.writeable = False
        _PARAMS_CACHE[key] = out
        while len(_PARAMS_CACHE) > 64:
            _PARAMS_CACHE.popitem(last=False)
    else:
        _PARAMS_CACHE.move_to_end(key)
    return out


def _planar_params(angles_deg, nz, nx):
    shape = np.asarray([nz, nx])
    out = np.empty((len(angles_deg), 6), dtype=np.float64)
    for i, ang in enumerate(angles_deg):
        c, s = scipy.special.cosdg(ang), scipy.special.sindg(ang)
        rot = np.array([[c, s], [-s, c]])
        out_center = rot @ ((shape - 1) / 2)
        in_center = (shape - 1) / 2
        offset = in_center - out_center
        out[i] = (rot[0, 0], rot[0, 1], rot[1, 0], rot[1, 1],
                  offset[0], offset[1])
    return out


def _key_array(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.tobytes()


class Plan:
    """Owns one odt_plan (device tables) for a fixed description."""

    def __init__(self, desc_fields, weights, params, device, tuning=None):
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        d = N.OdtDesc()
        for k, v in desc_fields.items():
            setattr(d, k, v)
        d.weights = N.dptr(self.weights)
        d.angle_params = N.dptr(self.params)
        self.desc = d
        self.desc_fields = dict(desc_fields)
        self.device = device
        self.tuning = dict(tuning) if tuning else None
        t = None
        if self.tuning:
            t = N.OdtTuning()
            for k, v in self.tuning.items():
                setattr(t, k, int(v))
        h = ctypes.c_void_p()
        N.check(N.lib().odt_plan_create_tuned(
            ctypes.byref(d), ctypes.byref(t) if t is not None else None,
            ctypes.byref(h)), "odt_plan_create_tuned")
        self.handle = h
        size = ctypes.c_size_t()
        N.check(N.lib().odt_plan_workspace_size(h, ctypes.byref(size)),
                "odt_plan_workspace_size")
        self.workspace_bytes = size.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                N.lib().odt_plan_destroy(h)
            except Exception:
                pass

    def enable_timing(self, on=True):
        N.check(N.lib().odt_plan_enable_timing(self.handle, int(on)),
                "odt_plan_enable_timing")

    def last_timing(self):
        r, f, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int()
        N.check(N.lib().odt_plan_last_timing(
            self.handle, ctypes.byref(r), ctypes.byref(f), ctypes.byref(t),
            ctypes.byref(n)), "odt_plan_last_timing")
        return {"rotate_ms": r.value, "filter_ms": f.value,
                "total_ms": t.value, "rotate_launches": n.value}

    def adjoint(self, d_gout, d_gusin, workspace=None, stream=None):
        """d_gusin = M^H d_gout for the sinogram-to-volume map M of this
        plan (odt_plan_adjoint)."""
        torch = _torch()
        if workspace is None:
            workspace = torch.empty(max(1, self.workspace_bytes),
                                    dtype=torch.uint8, device=d_gout.device)
        if stream is None:
            stream = torch.cuda.current_stream(d_gout.device).cuda_stream
        N.check(N.lib().odt_plan_adjoint(
            self.handle, ctypes.c_void_p(d_gout.data_ptr()),
            ctypes.c_void_p(d_gusin.data_ptr()),
            ctypes.c_void_p(workspace.data_ptr()), self.workspace_bytes,
            ctypes.c_void_p(stream)), "odt_plan_adjoint")
        return d_gusin

    def execute(self, d_usin, d_out, workspace=None, stream=None):
        torch = _torch()
        if workspace is None:
            workspace = torch.empty(max(1, self.workspace_bytes),
                                    dtype=torch.uint8, device=d_usin.device)
        if stream is None:
            stream = torch.cuda.current_stream(d_usin.device).cuda_stream
        N.check(N.lib().odt_plan_execute(
            self.handle, ctypes.c_void_p(d_usin.data_ptr()),
            ctypes.c_void_p(d_out.data_ptr()),
            ctypes.c_void_p(workspace.data_ptr()),
            ctypes.c_size_t(workspace.numel()), ctypes.c_void_p(stream)),
            "odt_plan_execute")
        return d_out


    def execute_noalloc(self, d_usin, d_out, stream=None):
        """odt_plan_execute_noalloc: the workspace is the library's static
        per-GPU scratch (N.ODT_SCRATCH_BYTES); nothing is allocated."""
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream(d_usin.device).cuda_stream
        N.check(N.lib().odt_plan_execute_noalloc(
            self.handle, ctypes.c_void_p(d_usin.data_ptr()),
            ctypes.c_void_p(d_out.data_ptr()), ctypes.c_void_p(stream)),
            "odt_plan_execute_noalloc")
        return d_out


def slab_budget(per_angle_bytes, A, device):
    """Bytes allowed for the batch of coefficient slabs: all angles, up to
    three eighths of the device's memory.  Derived from the total memory,
    not the free memory, so that identical calls build identical plans (the
    budget is part of the plan cache key, and the caching allocator's
    holdings change the free memory between calls)."""
    env = os.environ.get("ODTBRAIN_AMD_SLAB_BUDGET")
    if env:
        return int(env)
    torch = _torch()
    total = torch.cuda.get_device_properties(device).total_memory
    return int(max(per_angle_bytes, min(A * per_angle_bytes, 3 * total // 8)))


def get_plan(desc_fields, weights, params, device):
    from . import tune
    tuning = tune.lookup(desc_fields)
    key = (tuple(sorted(desc_fields.items())), _key_array(weights),
           _key_array(params), str(device),
           tuple(sorted(tuning.items())) if tuning else None)
    plan = _PLAN_CACHE.get(key)
    if plan is not None:
        _PLAN_CACHE.move_to_end(key)
        return plan
    plan = Plan(desc_fields, weights, params, device, tuning=tuning)
    _PLAN_CACHE[key] = plan
    while len(_PLAN_CACHE) > _PLAN_CACHE_SIZE:
        _PLAN_CACHE.popitem(last=False)
    return plan


def device_input(uSin, complex_dtype):
    """(tensor on the GPU, was_tensor) for an ndarray or a torch tensor."""
    torch = _torch()
    tdt = torch.complex64 if complex_dtype == np.complex64 else \
        torch.complex128
    if isinstance(uSin, torch.Tensor):
        if not uSin.is_cuda:
            t = uSin.to(dtype=tdt).contiguous().cuda()
            return t, False
        return uSin.to(dtype=tdt).contiguous(), True
    arr = np.ascontiguousarray(uSin, dtype=complex_dtype)
    if not torch.cuda.is_available():
        raise N.OdtError("odtbrain_amd needs a ROCm GPU (torch.cuda) to run")
    return torch.from_numpy(arr).cuda(), False


class Collector:
    """While active (``collect()``), ``run`` records each reconstruction
    instead of executing it; ``odtbrain_amd.batch.Batch`` replays them."""

    def __init__(self):
        self.entries = []


_collector = None


class collect:
    def __enter__(self):
        global _collector
        if _collector is not None:
            raise RuntimeError("nested batch collection")
        _collector = Collector()
        return _collector

    def __exit__(self, *exc):
        global _collector
        _collector = None
        return False


def run(desc_fields, weights, params, d_usin, out_shape, out_real,
        precision, return_tensor, timing=False):
    """Plan (cached), allocate output and workspace, execute."""
    torch = _torch()
    dev = d_usin.device
    if precision == N.ODT_FP64:
        odt = torch.float64 if out_real else torch.complex128
    else:
        odt = torch.float32 if out_real else torch.complex64
    if _collector is not None:
        from . import tune
        plan = Plan(desc_fields, weights, params, dev,
                    tuning=tune.lookup(desc_fields))  # owned by the batch
        out = torch.empty(out_shape, dtype=odt, device=dev)
        _collector.entries.append((plan, d_usin, out))
        return out  # filled by every replay of the batch
    global last_plan
    plan = get_plan(desc_fields, weights, params, dev)
    last_plan = plan
    if timing:
        plan.enable_timing(True)
    out = torch.empty(out_shape, dtype=odt, device=dev)
    plan.execute(d_usin, out)
    if return_tensor:
        return out
    return out.cpu().numpy()
