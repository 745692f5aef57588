"""Many reconstructions of uneven sizes replayed as one HIP graph.

The reference reconstructs one sinogram per call.  ``Batch`` records the
calls of a list of jobs (any of backpropagate_3d, backpropagate_3d_tilted,
backpropagate_2d with their usual arguments), builds one plan per job, and
captures all executions into one HIP graph whose ``lanes`` branches run
concurrently, each on its own part of the workspace (``odt_batch_create_lanes``;
the jobs are spread over the lanes longest first); ``run()`` replays the
graph.  Outputs are GPU tensors owned by the batch and overwritten by every
run.

    jobs = [(odtbrain_amd.backpropagate_2d, (sino, angles, res, nm),
             {"dtype": np.float32}), ...]
    b = Batch(jobs)         # lanes=4 by default
    outs = b.run()          # list of torch tensors, one per job
"""
import ctypes

from . import _core
from . import _native as N


class Batch:
    def __init__(self, jobs, lanes=4):
        torch = _core._torch()
        views = []
        with _core.collect() as col:
            for fn, args, kwargs in jobs:
                views.append(fn(*args, **dict(kwargs or {})))
        self.entries = col.entries
        self._views = views  # the functions' own output shapes
        if not self.entries:
            raise ValueError("empty batch")
        dev = self.entries[0][1].device
        for plan, d_in, out in self.entries:
            if d_in.device != dev:
                raise ValueError("all jobs of a batch must use one device")
        self.device = dev
        self.lanes = max(1, min(int(lanes), len(self.entries)))
        per = max(1, max(p.workspace_bytes for p, _, _ in self.entries))
        per = (per + 255) // 256 * 256
        ws = per * self.lanes
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=dev)
        # the null stream cannot be captured: a dedicated (default-priority)
        # stream runs the batch
        self.stream = torch.cuda.Stream(device=dev)
        n = len(self.entries)
        VP = ctypes.c_void_p * n
        self._plans = VP(*[p.handle.value for p, _, _ in self.entries])
        self._ins = VP(*[d.data_ptr() for _, d, _ in self.entries])
        self._outs = VP(*[o.data_ptr() for _, _, o in self.entries])
        h = ctypes.c_void_p()
        cur = torch.cuda.current_stream(dev)
        self.stream.wait_stream(cur)
        N.check(N.lib().odt_batch_create_lanes(
            n, self._plans, self._ins, self._outs,
            ctypes.c_void_p(self.workspace.data_ptr()), ws, self.lanes,
            ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)),
            "odt_batch_create_lanes")
        self.handle = h

    @property
    def outputs(self):
        return list(self._views)

    def run(self):
        """Replay the graph; the caller's current stream waits for it."""
        torch = _core._torch()
        cur = torch.cuda.current_stream(self.device)
        self.stream.wait_stream(cur)
        N.check(N.lib().odt_batch_launch(
            self.handle, ctypes.c_void_p(self.stream.cuda_stream)),
            "odt_batch_launch")
        cur.wait_stream(self.stream)
        return self.outputs

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                N.lib().odt_batch_destroy(h)
            except Exception:
                pass
