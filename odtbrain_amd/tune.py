"""Autotuner of the plans' launch shapes.

The reference runs on the CPU and has nothing to tune; on the GPU a plan's
launch shapes (angles per rotate-accumulate launch, concurrent angle groups
of the 2D rotate, the chunk of intermediate transforms, the column-stage
variant: ``odt_tuning`` in include/odtbrain_hip.h) trade memory and L2
reuse against launch count, and the best choice depends on the sizes.
``tune(fn, *args, **kwargs)`` times candidates of one reconstruction call
on the device and stores the fastest in a JSON cache keyed by the plan's
shape; plans created afterwards (``_core.get_plan``, batches) use it.

Cache files (JSON)::

    {"format": "odtbrain_amd-tuning", "version": 1,
     "entries": {"<shape key>": {"rotate_group": 48, ...}, ...}}

Lookup order: the user cache (``ODTBRAIN_AMD_TUNE_CACHE``, default
``~/.cache/odtbrain_amd/tuning.json``) over the packaged one measured on
MI355X (``tuning_mi355x.json`` beside this module).
``ODTBRAIN_AMD_NO_TUNING=1`` disables both.
"""
import json
import os
import pathlib
import time

from . import _core

FORMAT, VERSION = "odtbrain_amd-tuning", 1
FIELDS = ("rotate_group", "rotate_groups_2d", "t_budget_bytes",
          "col_per_angle")
PACKAGED = pathlib.Path(__file__).with_name("tuning_mi355x.json")
_entries_cache = None


def cache_path():
    env = os.environ.get("ODTBRAIN_AMD_TUNE_CACHE")
    if env:
        return pathlib.Path(env)
    return pathlib.Path(os.path.expanduser("~/.cache/odtbrain_amd/tuning.json"))


def shape_key(desc_fields):
    """Cache key of a plan description: kind, precision, angle count,
    detector and padded sizes, spline order, real output, flags."""
    f = desc_fields
    return ",".join(f"{k}={int(f[k])}" for k in (
        "kind", "precision", "A", "Ny", "Nx", "lNy", "lNx", "intp_order",
        "onlyreal", "flags"))


def _load(path):
    try:
        data = json.loads(pathlib.Path(path).read_text())
    except (OSError, ValueError):
        return {}
    if not isinstance(data, dict) or data.get("format") != FORMAT or \
            data.get("version") != VERSION:
        return {}
    out = {}
    for key, val in data.get("entries", {}).items():
        if isinstance(val, dict):
            out[key] = {f: int(val[f]) for f in FIELDS if f in val}
    return out


def _entries():
    global _entries_cache
    if _entries_cache is None:
        merged = _load(PACKAGED)
        merged.update(_load(cache_path()))
        _entries_cache = merged
    return _entries_cache


def reload():
    """Forget the loaded caches (they are read again on the next lookup)."""
    global _entries_cache
    _entries_cache = None


def lookup(desc_fields):
    """The cached tuning of a plan description, or None."""
    if os.environ.get("ODTBRAIN_AMD_NO_TUNING"):
        return None
    t = _entries().get(shape_key(desc_fields))
    return dict(t) if t else None


def save(key, tuning, path=None):
    """Add one entry to a cache file (the user cache by default)."""
    path = pathlib.Path(path) if path else cache_path()
    entries = _load(path)
    entries[key] = {f: int(v) for f, v in tuning.items() if f in FIELDS}
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(path.suffix + ".tmp")
    tmp.write_text(json.dumps({"format": FORMAT, "version": VERSION,
                               "entries": entries}, indent=1, sort_keys=True))
    os.replace(tmp, path)
    reload()


def search_space(desc_fields):
    """Candidate values per field, tried one field at a time."""
    kind, order = int(desc_fields["kind"]), int(desc_fields["intp_order"])
    gib = 1 << 30
    space = []
    if kind in (0, 1) and order >= 2:
        space.append(("rotate_group", [24, 36, 72, 96]))
    if kind == 2:
        space.append(("rotate_groups_2d", [1, 2, 4, 8, 16, 32]))
    space.append(("t_budget_bytes", [gib // 4, gib, 4 * gib]))
    if kind in (0, 1):
        space.append(("col_per_angle", [1]))
    return space


def tune(fn, *args, reps=3, path=None, store=True, space=None, **kwargs):
    """Time one reconstruction call ``fn(*args, **kwargs)`` (one of the
    three backpropagation functions) under candidate launch shapes on its
    device: the planner's defaults first, then each field of
    ``search_space`` in turn, keeping a value when it is at least 2 %
    faster.  The fastest tuning is stored in the cache (``path`` or the
    user cache) unless ``store`` is False.  Returns the shape key, the
    tuning and every trial's time (seconds, best of ``reps``)."""
    torch = _core._torch()
    with _core.collect() as col:
        fn(*args, **kwargs)
    if len(col.entries) != 1:
        raise ValueError("tune() takes one reconstruction call")
    plan0, d_in, out = col.entries[0]
    desc, w, prm, dev = plan0.desc_fields, plan0.weights, plan0.params, \
        plan0.device
    del plan0

    def trial(tuning):
        plan = _core.Plan(desc, w, prm, dev, tuning=tuning or None)
        ws = torch.empty(max(1, plan.workspace_bytes), dtype=torch.uint8,
                         device=dev)
        plan.execute(d_in, out, workspace=ws)
        torch.cuda.synchronize(dev)
        best = None
        for _ in range(max(1, int(reps))):
            t0 = time.perf_counter()
            plan.execute(d_in, out, workspace=ws)
            torch.cuda.synchronize(dev)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        del ws, plan
        return best

    best, best_t = {}, trial({})
    trials = [({}, best_t)]
    for field, values in (space if space is not None else search_space(desc)):
        for v in values:
            cand = dict(best)
            cand[field] = int(v)
            t = trial(cand)
            trials.append((cand, t))
            if t < 0.98 * best_t:
                best, best_t = cand, t
    key = shape_key(desc)
    if store:
        save(key, best, path)
    _core._PLAN_CACHE.clear()
    return {"key": key, "tuning": best, "seconds": best_t, "trials": trials}
