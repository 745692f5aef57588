"""odt_plan_execute_noalloc, the entry point that allocates nothing (its
workspace is the library's static per-GPU scratch array), called through
the C ABI: bit-identical to odt_plan_execute on consecutive calls of
different sizes (the scratch keeps the previous call's contents), and
plans whose workspace exceeds the scratch are rejected before any launch."""
import ctypes

import numpy as np
import pytest

import odtbrain_amd as G
from odtbrain_amd import _core, phantoms
from odtbrain_amd import _native as N

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _record(monkeypatch, budget, fn, *args, **kw):
    """Plan, device input and output of one reconstruction call, recorded
    without executing it; the slab batch is limited to ``budget`` bytes."""
    monkeypatch.setenv("ODTBRAIN_AMD_SLAB_BUDGET", str(budget))
    monkeypatch.setenv("ODTBRAIN_AMD_T_BUDGET", str(budget))
    with _core.collect() as col:
        fn(*args, **kw)
    (plan, d_in, out), = col.entries
    return plan, d_in, out


def _noalloc(plan, d_in, out):
    stream = torch.cuda.current_stream().cuda_stream
    rc = N.lib().odt_plan_execute_noalloc(
        plan.handle, ctypes.c_void_p(d_in.data_ptr()),
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    return rc


def test_noalloc_equals_execute_on_consecutive_calls(monkeypatch):
    big, ang_big = phantoms.sphere_sinogram_3d(N=64, A=100)
    small, ang_small = phantoms.sphere_sinogram_3d(N=40, A=30, radius=9.0,
                                                   seed=3)
    s2, a2 = phantoms.disk_sinogram_2d(N=256, A=180, seed=4)
    cases = [
        _record(monkeypatch, 32 << 20, G.backpropagate_3d,
                big.astype(np.complex64), ang_big, 6.0, 1.333,
                dtype=np.float32),
        _record(monkeypatch, 8 << 20, G.backpropagate_3d, small, ang_small,
                6.0, 1.333, dtype=np.float64),
        _record(monkeypatch, 16 << 20, G.backpropagate_2d, s2, a2, 6.0,
                1.333, dtype=np.float32),
    ]
    for plan, _d, _o in cases:
        print(f"plan workspace {plan.workspace_bytes} B", flush=True)
        assert 0 < plan.workspace_bytes <= N.ODT_SCRATCH_BYTES
    # the first case again last: its scratch then holds the 2D call's data
    for i in (0, 1, 2, 0):
        plan, d_in, out = cases[i]
        ref = torch.empty_like(out)
        plan.execute(d_in, ref)
        out.fill_(float("nan"))
        assert _noalloc(plan, d_in, out) == N.ODT_OK, N.lib().odt_last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), i
    # the second call of a pair with a smaller input than the first
    assert cases[1][0].workspace_bytes < cases[0][0].workspace_bytes


def test_noalloc_rejects_plans_beyond_the_scratch(monkeypatch):
    sino, ang = phantoms.sphere_sinogram_3d(N=64, A=100)
    plan, d_in, out = _record(monkeypatch, 1 << 30, G.backpropagate_3d, sino,
                              ang, 6.0, 1.333)
    assert plan.workspace_bytes > N.ODT_SCRATCH_BYTES
    out.fill_(1.0)
    assert _noalloc(plan, d_in, out) == N.ODT_ERR_WORKSPACE
    assert b"static scratch" in N.lib().odt_last_error()
    torch.cuda.synchronize()
    assert bool((out == 1.0).all())  # nothing ran
