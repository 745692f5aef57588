"""Batched reconstructions replayed as one HIP graph (odt_batch_*):
every replay equals the eager calls bit for bit, with eager GPU work and
synchronisations between the replays, and the entries match the oracle."""
import numpy as np
import pytest

from common import rel_l2
import odtbrain_amd as G
from odtbrain_amd import phantoms
from oracle import odt_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _jobs():
    s2a, a2a = phantoms.disk_sinogram_2d(N=128, A=40, seed=1)
    s2b, a2b = phantoms.disk_sinogram_2d(N=96, A=33, dn=0.02, seed=2)
    s3a, a3a = phantoms.sphere_sinogram_3d(N=40, A=24, radius=9.0, seed=3)
    s3b, a3b = phantoms.sphere_sinogram_3d(N=33, A=17, radius=7.0, seed=4)
    return [
        (G.backpropagate_2d, (s2a, a2a, 6.0, 1.333), {"dtype": np.float32}),
        (G.backpropagate_2d, (s2b, a2b, 6.0, 1.333), {"dtype": np.float64}),
        (G.backpropagate_3d, (s3a.astype(np.complex64), a3a, 6.0, 1.333),
         {"dtype": np.float32}),
        (G.backpropagate_3d, (s3b, a3b, 6.0, 1.333), {"dtype": np.float64,
                                                     "onlyreal": True}),
    ]


@pytest.mark.parametrize("lanes", [1, 3])
def test_batch_graph_replays_equal_eager_and_oracle(lanes):
    """One lane (the plans in order on one stream) and three concurrent
    lanes (each with its own part of the workspace)."""
    jobs = _jobs()
    eager = [fn(*a, **kw) for fn, a, kw in jobs]
    b = G.Batch(jobs, lanes=lanes)
    other_s, other_a = phantoms.disk_sinogram_2d(N=64, A=20, seed=9)
    for rep in range(3):
        outs = b.run()
        torch.cuda.synchronize()
        for i, (o, e) in enumerate(zip(outs, eager)):
            got = o.cpu().numpy()
            assert got.shape == e.shape and got.dtype == e.dtype, i
            assert np.array_equal(got, e), (rep, i)
        # the program's own eager work between replays
        G.backpropagate_2d(other_s, other_a, 6.0, 1.333, dtype=np.float32)
        torch.cuda.synchronize()
    # entries against the oracle
    for (fn, a, kw), e in zip(jobs, eager):
        kw = dict(kw)
        kw.pop("dtype")
        oracle_fn = O.backpropagate_2d if fn is G.backpropagate_2d else \
            O.backpropagate_3d
        ref = oracle_fn(np.asarray(a[0], dtype=np.complex128), *a[1:], **kw)
        tol = 1e-5 if e.dtype in (np.complex64, np.float32) else 1e-10
        assert rel_l2(e, ref) < tol


def test_batch_rejects_null_stream_and_small_workspace():
    import ctypes
    from odtbrain_amd import _native as N
    b = G.Batch(_jobs()[2:3])
    plan, d_in, out = b.entries[0]
    VP = ctypes.c_void_p * 1
    args = (VP(plan.handle.value), VP(d_in.data_ptr()), VP(out.data_ptr()),
            ctypes.c_void_p(b.workspace.data_ptr()))
    h = ctypes.c_void_p()
    rc = N.lib().odt_batch_create(1, *args, plan.workspace_bytes, None,
                                  ctypes.byref(h))
    assert rc == N.ODT_ERR_INVALID
    assert plan.workspace_bytes > 0
    rc = N.lib().odt_batch_create(1, *args, plan.workspace_bytes - 1,
                                  ctypes.c_void_p(b.stream.cuda_stream),
                                  ctypes.byref(h))
    assert rc == N.ODT_ERR_WORKSPACE
