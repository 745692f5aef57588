"""The autotuner (odtbrain_amd.tune) and the plans' launch-shape choices
(odt_plan_create_tuned): the tuner stores its fastest choice and later
calls use it; every choice keeps the results -- bit-identical where it only
re-chunks the same kernels' arithmetic (the transform chunk budget; in FP64
also the angles per rotate launch, whose running sums cross launches
unrounded), to rounding where another kernel variant or grouping computes
the same values (the column-stage variant, FP32 launch regrouping, the 2D
angle groups)."""
import numpy as np
import pytest

import odtbrain_amd as G
from common import create_test_sino_2d, create_test_sino_3d, rel_l2
from odtbrain_amd import _core, tune

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _record(fn, *args, **kwargs):
    with _core.collect() as col:
        fn(*args, **kwargs)
    (plan, d_in, out), = col.entries
    return plan, d_in, out


def _run(plan, d_in, out, tuning):
    p = _core.Plan(plan.desc_fields, plan.weights, plan.params, plan.device,
                   tuning=tuning)
    res = torch.empty_like(out)
    p.execute(d_in, res)
    torch.cuda.synchronize()
    return res.cpu().numpy()


def test_tune_stores_and_uses_the_fastest_tuning(tmp_path, monkeypatch):
    monkeypatch.setenv("ODTBRAIN_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    tune.reload()
    sino, angles = create_test_sino_3d(A=30, Nx=40, Ny=40)
    args = (sino, angles, 6.0, 1.333)
    space = [("rotate_group", [8, 16]), ("t_budget_bytes", [1 << 22]),
             ("col_per_angle", [1])]
    res = tune.tune(G.backpropagate_3d, *args, dtype=np.float64, reps=1,
                    space=space)
    assert len(res["trials"]) == 5
    plan, _d, _o = _record(G.backpropagate_3d, *args, dtype=np.float64)
    assert tune.shape_key(plan.desc_fields) == res["key"]
    assert tune.lookup(plan.desc_fields) == res["tuning"]
    assert plan.tuning == (res["tuning"] or None)
    tuned = G.backpropagate_3d(*args, dtype=np.float64)
    monkeypatch.setenv("ODTBRAIN_AMD_NO_TUNING", "1")
    default = G.backpropagate_3d(*args, dtype=np.float64)
    print("tuning", res["tuning"], "rel", rel_l2(tuned, default))
    assert rel_l2(tuned, default) < 1e-15
    tune.reload()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tuning_choices_keep_results_3d(dtype):
    sino, angles = create_test_sino_3d(A=30, Nx=40, Ny=40)
    plan, d_in, out = _record(G.backpropagate_3d, sino, angles, 6.0, 1.333,
                              dtype=dtype)
    ref = _run(plan, d_in, out, None)
    assert np.array_equal(_run(plan, d_in, out, {"t_budget_bytes": 1 << 22}),
                          ref)
    tol = 1e-15 if dtype is np.float64 else 1e-6
    for tuning in ({"col_per_angle": 1}, {"col_per_angle": 2}):
        err = rel_l2(_run(plan, d_in, out, tuning), ref)
        print(tuning, err)
        assert err < tol, tuning
    got = _run(plan, d_in, out, {"rotate_group": 7})
    if dtype is np.float64:
        assert np.array_equal(got, ref)
    else:
        assert rel_l2(got, ref) < 1e-6


def test_tuning_choices_keep_results_2d():
    sino, angles = create_test_sino_2d(A=43, N=64)
    plan, d_in, out = _record(G.backpropagate_2d, sino, angles, 3.3, 1.34)
    ref = _run(plan, d_in, out, None)
    for g in (1, 3, 8):
        assert rel_l2(_run(plan, d_in, out, {"rotate_groups_2d": g}),
                      ref) < 1e-14
    assert np.array_equal(_run(plan, d_in, out, {"t_budget_bytes": 1 << 20}),
                          ref)


def test_tuning_rejects_bad_values():
    from odtbrain_amd import _native as N
    sino, angles = create_test_sino_3d(A=9, Nx=16, Ny=16)
    plan, _d, _o = _record(G.backpropagate_3d, sino, angles, 6.0, 1.333)
    with pytest.raises(N.OdtError):
        _core.Plan(plan.desc_fields, plan.weights, plan.params, plan.device,
                   tuning={"col_per_angle": 3})
