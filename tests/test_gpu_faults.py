"""Error paths of the C ABI under faults forced from the host
(ODTBRAIN_AMD_FAULT, include/odtbrain_hip.h): a device allocation or a
checked launch that fails part-way through plan creation, execution, the
adjoint, a Bluestein transform or batch creation makes the call return an
error code with a message (OdtError in Python), leaves no device allocation
of the library behind, and the next call succeeds with correct results.
The reference has no failure handling of its own (a worker exception
re-raises from pool.map, _alg3d_bpp.py:65); these tests pin ours.  The
faults are injected on the host: every kernel that was issued runs and ends
normally."""
import ctypes
import gc

import numpy as np
import pytest

import odtbrain_amd as G
from common import create_test_sino_2d, create_test_sino_3d, rel_l2
from odtbrain_amd import _core
from odtbrain_amd import _native as N
from odtbrain_amd import autograd as AG
from oracle import odt_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FAULT = "ODTBRAIN_AMD_FAULT"


def _drop_plans():
    _core._PLAN_CACHE.clear()
    _core.last_plan = None
    gc.collect()
    torch.cuda.synchronize()


def _sweep(monkeypatch, kind, call, limit=80):
    """Run ``call`` with the n-th fault of ``kind`` for n = 1, 2, ... until
    it succeeds (the fault no longer fires); every failing call must raise
    OdtError and leave the library's allocations as they were.  Returns the
    number of faults that fired."""
    fired = 0
    for n in range(1, limit + 1):
        _drop_plans()
        base = N.live_allocations()
        monkeypatch.setenv(FAULT, f"{kind}:{n}")
        msg = None
        try:
            call()
        except N.OdtError as e:
            msg = str(e)
        finally:
            monkeypatch.delenv(FAULT, raising=False)
        if msg is None:
            return fired
        # (outside the except clause: its traceback no longer holds the
        # frames, and with them the plan of a call that failed to execute)
        assert msg, "error without a message"
        torch.cuda.synchronize()
        _drop_plans()
        assert N.live_allocations() == base, (kind, n, msg)
        fired += 1
    raise AssertionError(f"{kind} faults still firing after {limit} tries")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_plan_creation_allocation_and_launch_faults(monkeypatch, dtype):
    sino, angles = create_test_sino_3d(A=9, Nx=20, Ny=16)
    kw = dict(res=3.0, nm=1.335, lD=0.5)
    ref = O.backpropagate_3d(sino, angles, **kw)
    d_in = torch.from_numpy(sino).cuda()

    def call():
        return G.backpropagate_3d(d_in, angles, dtype=dtype, **kw)

    # allocations: the plan's tables (and the float64 forward tables of
    # float32 plans), then the first launches and copies of creation and
    # execution
    assert _sweep(monkeypatch, "alloc", call) >= 9
    assert _sweep(monkeypatch, "launch", call) >= 10
    _drop_plans()
    got = call().cpu().numpy()
    assert rel_l2(got, ref) < (1e-5 if dtype == np.float32 else 1e-9)


def test_execute_launch_faults_then_recovery(monkeypatch):
    """One plan, created before the faults: each failing execute returns an
    error after queuing part of its work; the next execute recomputes the
    whole result."""
    sino, angles = create_test_sino_3d(A=13, Nx=24, Ny=20)
    kw = dict(res=3.0, nm=1.335)
    ref = O.backpropagate_3d(sino, angles, **kw)
    d_in = torch.from_numpy(sino).cuda()
    G.backpropagate_3d(d_in, angles, dtype=np.float64, **kw)
    plan = _core.last_plan
    base = N.live_allocations()
    out = torch.empty((24, 20, 24), dtype=torch.complex128, device="cuda")
    fired = 0
    for n in range(1, 60):
        monkeypatch.setenv(FAULT, f"launch:{n}")
        try:
            plan.execute(d_in, out)
        except N.OdtError as e:
            assert "failure" in str(e) or "k_" in str(e), str(e)
            fired += 1
            continue
        finally:
            monkeypatch.delenv(FAULT, raising=False)
        break
    assert fired >= 4
    torch.cuda.synchronize()
    assert N.live_allocations() == base
    out.fill_(float("nan"))
    plan.execute(d_in, out)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-9


def test_adjoint_launch_faults_then_recovery(monkeypatch):
    A, ny, nx = 9, 20, 24
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    rng = np.random.default_rng(3)
    g = torch.from_numpy(rng.standard_normal((nx, ny, nx)) +
                         1j * rng.standard_normal((nx, ny, nx))).cuda()
    want = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, dtype=np.float64)
    base = N.live_allocations()
    fired = 0
    for n in range(1, 60):
        monkeypatch.setenv(FAULT, f"launch:{n}")
        try:
            AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, dtype=np.float64)
        except N.OdtError:
            fired += 1
            continue
        finally:
            monkeypatch.delenv(FAULT, raising=False)
        break
    assert fired >= 4
    torch.cuda.synchronize()
    assert N.live_allocations() == base
    got = AG.backpropagate_3d_adjoint(g, angles, 6.0, 1.333, dtype=np.float64)
    assert torch.equal(got, want)


@pytest.mark.parametrize("prec", [N.ODT_FP32, N.ODT_FP64])
def test_bluestein_table_faults(monkeypatch, prec):
    """odt_fft_c2c of a length that is not a power of two allocates its
    tables (and a float64 scratch for float32 tables) per call."""
    n, batch = 97, 5
    rng = np.random.default_rng(1)
    x = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
    cdt = torch.complex64 if prec == N.ODT_FP32 else torch.complex128
    d_in = torch.from_numpy(x).to(cdt).cuda()
    d_out = torch.empty_like(d_in)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def call():
        N.check(N.lib().odt_fft_c2c(prec, batch, n, -1,
                                    ctypes.c_void_p(d_in.data_ptr()),
                                    ctypes.c_void_p(d_out.data_ptr()), stream),
                "fft")

    expected_allocs = 2 if prec == N.ODT_FP32 else 1
    assert _sweep(monkeypatch, "alloc", call) == expected_allocs
    assert _sweep(monkeypatch, "launch", call) >= expected_allocs
    call()
    torch.cuda.synchronize()
    ref = np.fft.fft(x, axis=-1)
    tol = 1e-5 if prec == N.ODT_FP32 else 1e-12
    assert rel_l2(d_out.cpu().numpy(), ref) < tol


def test_batch_creation_fault_then_batch(monkeypatch):
    """A launch failing inside batch creation (eager pass or graph capture)
    returns the error and releases the streams, events and graph; a batch of
    the same jobs built afterwards replays correctly."""
    s2, a2 = create_test_sino_2d(A=11, N=24)
    s3, a3 = create_test_sino_3d(A=9, Nx=20, Ny=16)
    kw = dict(res=3.0, nm=1.335)
    d2 = torch.from_numpy(s2).cuda()
    d3 = torch.from_numpy(s3).cuda()
    jobs = [(G.backpropagate_2d, (d2, a2), dict(dtype=np.float64, **kw)),
            (G.backpropagate_3d, (d3, a3), dict(dtype=np.float64, **kw))]
    fired = 0
    for n in range(1, 200, 7):
        monkeypatch.setenv(FAULT, f"launch:{n}")
        try:
            G.Batch(jobs, lanes=2)
        except N.OdtError:
            fired += 1
            continue
        finally:
            monkeypatch.delenv(FAULT, raising=False)
        break
    assert fired >= 2
    torch.cuda.synchronize()
    b = G.Batch(jobs, lanes=2)
    outs = [o.cpu().numpy() for o in b.run()]
    torch.cuda.synchronize()
    assert rel_l2(outs[0], O.backpropagate_2d(s2, a2, **kw)) < 1e-9
    assert rel_l2(outs[1], O.backpropagate_3d(s3, a3, **kw)) < 1e-9
