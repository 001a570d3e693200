"""Every pruned-search kernel against the CPU oracle: SPEEDYNN_PRUNED forces
one kernel (per-wave, shared scan, split shared scan, sorted-column walk, the
shared scan at each workgroup width the autotuner tries) for
every step where it applies, so each is checked on its own rather than only
when the autotuner picks it.  Shapes are chosen so that each kernel applies:
more than four candidate blocks for the split scan, a second forward step with
finite X for the walk, and a long backward run in which the bound matrix M is
recomputed over the remaining candidates (at d/2 and d/4)."""
import os

import numpy as np
import pytest

from conftest import (GOLDEN, assert_same_result, candidate_error_bits, expected_arrays,
                      load_case)
from oracle import oracle as O
from speedynn_amd import _lib, datasets as ds

pytestmark = pytest.mark.gpu

LOSS = {"rmse": 0, "zero_one": 4}

CASES = [
    # seed, n, d, k, method, model, num_features
    (81, 400, 300, 10, "forward", "classification", 4),
    (82, 200, 300, 5, "backward", "regression", 240),
    (83, 300, 150, 5, "backward", "regression", 20),
    (84, 350, 70, 20, "forward", "regression", 5),
]


@pytest.fixture(scope="module")
def engine():
    print("engine: creating context on device 0", flush=True)
    return _lib.Engine(0)


def _data(seed, n, d, model):
    if model == "classification":
        X, y, _ = ds.classification_argmax(seed, n, d, 4, 3, 0.1)
    else:
        X, y, _ = ds.regression_sum(seed, n, d, 4, 0.5)
    return X, y


@pytest.mark.parametrize("kernel", ["wave", "shared", "split", "walk", "shared4", "shared8",
                                    "shared16"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"seed{c[0]}")
def test_forced_pruned_kernel_matches_oracle(engine, monkeypatch, case, kernel):
    seed, n, d, k, method, model, nf = case
    X, y = _data(seed, n, d, model)
    loss = "zero_one" if model == "classification" else "rmse"
    print(f"kernel={kernel} seed={seed} n={n} d={d} {method}", flush=True)
    o = O.select(X, y, [k], nf, n, [loss], False, model, method, record_candidates=True)
    monkeypatch.setenv("SPEEDYNN_PRUNED", kernel)
    g = engine.select(X, y, [k], [LOSS[loss]], method=1 if method == "forward" else 0,
                      model_type=1 if model == "classification" else 0, k_fold=n,
                      num_features=nf, form=1, candidate_errors=True)
    assert_same_result(g, o)
    steps = g["steps"]
    np.testing.assert_array_equal(g["candidate_errors"][:steps].view(np.uint32),
                                  o["candidate_errors"][:steps].view(np.uint32))
    used = g["pruned_kernel"]
    want = {"wave": 1, "shared": 2, "split": 4, "walk": 8, "shared4": 16, "shared8": 32,
            "shared16": 64}[kernel]
    applies = {
        "wave": True,
        "shared": True,
        "split": d > 256,
        "walk": method == "forward",
        "shared4": True,
        "shared8": True,
        "shared16": k <= 20,
    }[kernel]
    if applies:
        assert used & want, (kernel, used)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"seed{c[0]}")
def test_autotuned_selection_matches_oracle(engine, monkeypatch, case):
    """The autotuner on (no kernel forced): the first step of each step range
    runs every applicable kernel, kernel r on the queries r, r + nc, ... of
    that step, so that step's results come from several kernels at once
    (up to five here: per-wave, the shared scan at widths 4, 8 and 16, and the
    walk).  Every step,
    tuning or not, must give the oracle's errors bit for bit.  These shapes
    appear in no other autotuned test, so the tuning steps run here."""
    seed, n, d, k, method, model, nf = case
    X, y = _data(seed, n, d, model)
    loss = "zero_one" if model == "classification" else "rmse"
    print(f"autotuned seed={seed} n={n} d={d} {method}", flush=True)
    o = O.select(X, y, [k], nf, n, [loss], False, model, method, record_candidates=True)
    monkeypatch.delenv("SPEEDYNN_PRUNED", raising=False)
    monkeypatch.delenv("SPEEDYNN_TUNING", raising=False)
    g = engine.select(X, y, [k], [LOSS[loss]], method=1 if method == "forward" else 0,
                      model_type=1 if model == "classification" else 0, k_fold=n,
                      num_features=nf, form=1, candidate_errors=True)
    assert_same_result(g, o)
    steps = g["steps"]
    np.testing.assert_array_equal(g["candidate_errors"][:steps].view(np.uint32),
                                  o["candidate_errors"][:steps].view(np.uint32))
    assert bin(g["pruned_kernel"]).count("1") >= 2, g["pruned_kernel"]  # tuning ran


REDUCED = [os.path.join(GOLDEN, f) for f in ("c2_n3000_d512_oracle.json",
                                             "c4_n1000_d320_oracle.json")]


@pytest.mark.parametrize("kernel", ["wave", "shared", "split", "walk", "shared4", "shared8",
                                    "shared16"])
@pytest.mark.parametrize("path", REDUCED, ids=lambda p: os.path.basename(p)[:13])
def test_forced_kernel_on_workload_recipes_matches_oracle(engine, monkeypatch, path, kernel):
    """Each pruned kernel forced on the n=65536 forward recipe (feature scales
    10^U(-2,2), 32 informative features, k = 20 over eight candidate blocks,
    the last partly filled) and on the n=32768 backward recipe (half the
    columns noisy copies of others, regression, k = 15), at sizes the oracle
    finishes in a minute: every candidate's error at every step."""
    rec, X, y = load_case(path)
    inp = rec["inputs"]
    print(f"kernel={kernel} {rec['name']}", flush=True)
    monkeypatch.setenv("SPEEDYNN_PRUNED", kernel)
    g = engine.select(X, y, inp["k_vals"], [LOSS[inp["loss_functions"][0]]],
                      method=1 if inp["method"] == "forward" else 0,
                      model_type=1 if inp["model_type"] == "classification" else 0,
                      k_fold=inp["k_fold"], num_features=inp["num_features"], form=1,
                      candidate_errors=True)
    assert_same_result(g, expected_arrays(rec), rec["name"])
    steps = g["steps"]
    np.testing.assert_array_equal(g["candidate_errors"][:steps].view(np.uint32),
                                  candidate_error_bits(rec))


_NO_DEFER_SCRIPT = r"""
import json, os, sys
import numpy as np
sys.path.insert(0, os.environ["SNN_ROOT"])
sys.path.insert(0, os.path.join(os.environ["SNN_ROOT"], "tests"))
from conftest import load_case
from speedynn_amd import _lib
rec, X, y = load_case(os.environ["SNN_CASE"])
inp = rec["inputs"]
eng = _lib.Engine(0)
g = eng.select(X, y, inp["k_vals"], [0], method=0, model_type=0, k_fold=inp["k_fold"],
               num_features=inp["num_features"], form=1, candidate_errors=True)
steps = int(g["steps"])
print(json.dumps({"ordering": np.asarray(g["selected_ordering"]).tolist(),
                  "errors": np.asarray(g["selected_errors"], np.float32).view(np.uint32).tolist(),
                  "cand": g["candidate_errors"][:steps].view(np.uint32).ravel().tolist()}))
"""


def test_backward_without_deferred_writes_matches_oracle():
    """SPEEDYNN_DEFER=off (read once per process, so in a process of its own):
    the backward selection writes D after every step instead of every other,
    on the n=32768 backward recipe at reduced size; every candidate's error at
    every step against the oracle fixture, as the deferred default is checked
    above."""
    import json
    import subprocess
    import sys

    path = REDUCED[1]
    rec, _, _ = load_case(path)
    inp = rec["inputs"]
    assert inp["method"] == "backward" and inp["loss_functions"] == ["rmse"]
    assert inp["model_type"] == "regression" and inp["n"] % 4 == 0  # row lists, deferred writes
    env = dict(os.environ, SPEEDYNN_DEFER="off", SNN_CASE=path,
               SNN_ROOT=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    print(f"subprocess: SPEEDYNN_DEFER=off {rec['name']}", flush=True)
    out = subprocess.run([sys.executable, "-c", _NO_DEFER_SCRIPT], env=env, capture_output=True,
                         text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    got = json.loads(out.stdout.strip().splitlines()[-1])
    exp = expected_arrays(rec)
    np.testing.assert_array_equal(got["ordering"], exp["selected_ordering"])
    np.testing.assert_array_equal(np.array(got["errors"], np.uint32),
                                  np.asarray(exp["selected_errors"], np.float32).view(np.uint32))
    bits = candidate_error_bits(rec)
    np.testing.assert_array_equal(np.array(got["cand"], np.uint32).reshape(bits.shape), bits)


_TUNE_FILE_SCRIPT = r"""
import json, os, sys
import numpy as np
sys.path.insert(0, os.environ["SNN_ROOT"])
sys.path.insert(0, os.path.join(os.environ["SNN_ROOT"], "tests"))
from conftest import load_case
from speedynn_amd import _lib
rec, X, y = load_case(os.environ["SNN_CASE"])
inp = rec["inputs"]
eng = _lib.Engine(0)
g = eng.select(X, y, inp["k_vals"], [4], method=1, model_type=1, k_fold=inp["k_fold"],
               num_features=inp["num_features"], form=1, candidate_errors=True)
steps = int(g["steps"])
print(json.dumps({"ordering": np.asarray(g["selected_ordering"]).tolist(),
                  "cand": g["candidate_errors"][:steps].view(np.uint32).ravel().tolist()}))
"""


def test_tune_file_list_cap_and_choices_round_trip(tmp_path):
    """A tune file (SPEEDYNN_TUNE_FILE, read once per process, so in a process
    of its own) that records a row-list cap of n/2 for the n=65536 forward
    recipe's reduced shape: the selection with those longer lists equals the
    oracle fixture in every candidate error, and the file is rewritten with
    the cap kept beside the kernel choices the tuning trials made."""
    import json
    import subprocess
    import sys

    path = REDUCED[0]
    rec, _, _ = load_case(path)
    inp = rec["inputs"]
    assert inp["method"] == "forward" and inp["n"] % 4 == 0 and inp["k_vals"] == [20]
    assert inp["loss_functions"] == ["zero_one"] and inp["model_type"] == "classification"
    tune = tmp_path / "tune.json"
    cap_line = {"n": inp["n"], "d": inp["d"], "list": 20, "mult": 1, "list_cap_div": 2}
    tune.write_text("[\n  " + json.dumps(cap_line) + "\n]\n")
    env = dict(os.environ, SPEEDYNN_TUNE_FILE=str(tune), SNN_CASE=path,
               SNN_ROOT=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    env.pop("SPEEDYNN_TUNING", None)
    env.pop("SPEEDYNN_PRUNED", None)
    print(f"subprocess: tune file with list_cap_div 2, {rec['name']}", flush=True)
    out = subprocess.run([sys.executable, "-c", _TUNE_FILE_SCRIPT], env=env, capture_output=True,
                         text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    got = json.loads(out.stdout.strip().splitlines()[-1])
    np.testing.assert_array_equal(got["ordering"], expected_arrays(rec)["selected_ordering"])
    bits = candidate_error_bits(rec)
    np.testing.assert_array_equal(np.array(got["cand"], np.uint32).reshape(bits.shape), bits)
    entries = json.loads(tune.read_text())
    assert cap_line in entries
    assert any(e.get("kernel") for e in entries if e.get("n") == inp["n"])
