"""GPU parity of the batched selection (snn_select_batch): problems of uneven
n, d and k advanced together, with and without the per-step hipGraph, must
give each problem exactly what a selection of its own gives -- the oracle's
results on small problems, the reference's golden vectors on a problem of the
64-problem workload."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_same_result, expected_arrays, load_case
from oracle import oracle as O
from speedynn_amd import _lib, datasets as ds

pytestmark = pytest.mark.gpu

LOSS_CODES = {"rmse": 0, "rmse_norm": 1, "mad": 2, "mad_norm": 3, "zero_one": 4, "std": 5}

# seed, n, d, k_vals, k_fold, model, method, num_features, auto_feature, losses
MIXED = [
    (71, 300, 20, [3, 5], 300, "classification", "forward", 6, False, ["zero_one"]),
    (72, 500, 37, [5], 5, "regression", "backward", 30, False, ["rmse", "mad"]),
    (73, 257, 70, [1, 2, 10], 257, "regression", "forward", 4, False, ["rmse", "std"]),
    (74, 1000, 64, [20], 7, "classification", "forward", 0, True, ["zero_one"]),
    (75, 150, 5, [24], 150, "classification", "backward", 1, False, ["zero_one"]),
    (76, 777, 130, [8], 777, "classification", "forward", 5, False, ["zero_one"]),
    (77, 64, 3, [2], 64, "regression", "backward", 1, True, ["rmse"]),
]


def _case(seed, n, d, model):
    if model == "classification":
        X, y, _ = ds.classification_argmax(seed, n, d, 3, 3, 0.1)
        if seed % 2:
            X = ds.quantized(X, 0.25)
    else:
        X, y, _ = ds.regression_sum(seed, n, d, 3, 0.5)
    return X, y


def _problems(cases):
    out = []
    for seed, n, d, kv, kf, model, method, nf, auto, losses in cases:
        X, y = _case(seed, n, d, model)
        out.append(dict(X=X, y=y, k_vals=kv, loss_codes=[LOSS_CODES[l] for l in losses],
                        loss_names=losses, method=1 if method == "forward" else 0,
                        method_name=method, model_type=1 if model == "classification" else 0,
                        model_name=model, k_fold=kf, num_features=nf, auto_feature=auto))
    return out


@pytest.fixture(scope="module")
def engine():
    print("engine: creating context on device 0", flush=True)
    return _lib.Engine(0)


@pytest.fixture(scope="module")
def mixed():
    probs = _problems(MIXED)
    expected = [O.select(p["X"], p["y"], p["k_vals"], p["num_features"], p["k_fold"],
                         p["loss_names"], p["auto_feature"], p["model_name"], p["method_name"])
                for p in probs]
    return probs, expected


@pytest.mark.parametrize("use_graph", [False, True], ids=["stream", "graph"])
@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
def test_mixed_batch_matches_oracle(engine, mixed, form, use_graph):
    probs, expected = mixed
    print(f"mixed batch of {len(probs)} form={form} graph={use_graph}", flush=True)
    got, summary = engine.select_batch(probs, form=form, use_graph=use_graph)
    assert summary["rc"] == 0
    for i, (g, e) in enumerate(zip(got, expected)):
        assert g["status"] == 0, i
        assert_same_result(g, e, f"problem {i} ({MIXED[i][0]})")
    # a second call reuses the cached graphs and buffers
    again, _ = engine.select_batch(probs, form=form, use_graph=use_graph)
    for i, (g, e) in enumerate(zip(again, expected)):
        assert_same_result(g, e, f"repeat, problem {i}")


def test_batch_reports_invalid_problems_and_runs_the_rest(engine, mixed):
    probs, expected = mixed
    bad = dict(probs[0])
    bad["k_fold"] = 0
    big_k = dict(probs[1])
    big_k["k_vals"] = [65]
    got, summary = engine.select_batch([bad, probs[0], big_k, probs[2]], form=1,
                                       raise_on_error=False)
    assert summary["rc"] == _lib.SNN_EINVAL
    assert got[0]["status"] == _lib.SNN_EINVAL
    assert got[2]["status"] == _lib.SNN_ENOTSUP
    assert_same_result(got[1], expected[0], "valid problem beside invalid ones")
    assert_same_result(got[3], expected[2], "valid problem beside invalid ones")
    with pytest.raises(_lib.EngineError):
        engine.select_batch([bad], form=1)


@pytest.mark.slow
def test_64_problem_workload(engine):
    """All 64 problems in one batch (precision form, graphs): problem 22 equals
    the reference's golden vector; every problem equals its own plain-form
    selection."""
    print("generating the 64-problem workload", flush=True)
    probs = []
    for p in range(64):
        w = ds.batch_problem(p)
        probs.append(dict(X=w["X"], y=w["y"], k_vals=w["k_vals"], loss_codes=[4], method=1,
                          model_type=1, k_fold=w["k_fold"], num_features=w["num_features"]))
    print("batch run", flush=True)
    got, summary = engine.select_batch(probs, form=1, use_graph=True)
    print(f"batch ms_total={summary['ms_total']:.1f}", flush=True)
    rec, X, y = load_case(os.path.join(GOLDEN, "c3_problem22.json"))
    np.testing.assert_array_equal(X, probs[22]["X"])
    assert_same_result(got[22], expected_arrays(rec), "problem 22 vs reference")
    for p in range(64):
        one = engine.select(probs[p]["X"], probs[p]["y"], probs[p]["k_vals"], [4], method=1,
                            model_type=1, k_fold=probs[p]["k_fold"],
                            num_features=probs[p]["num_features"], form=0)
        assert_same_result(got[p], one, f"problem {p}")
