import base64
import glob
import json
import os
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def load_case(path):
    """Golden case -> (record, X, y) with inputs rebuilt and checksum-verified."""
    from speedynn_amd import datasets as ds
    with open(path) as fh:
        rec = json.load(fh)
    inp = rec["inputs"]
    if "X_b64" in inp:
        X = np.frombuffer(base64.b64decode(inp["X_b64"]), np.float32).reshape(inp["n"], inp["d"])
        y = np.frombuffer(base64.b64decode(inp["y_b64"]), np.float32).copy()
        X = X.copy()
    elif inp.get("recipe") == "config":
        w = ds.config(inp["config"])
        X, y = w["X"], w["y"]
    elif inp.get("recipe") == "reduced":
        w = ds.reduced(inp["config"], inp["n"], inp["d"], inp["num_features"])
        X, y = w["X"], w["y"]
    elif inp.get("recipe") == "batch_problem":
        w = ds.batch_problem(inp["problem"])
        X, y = w["X"], w["y"]
    else:
        raise ValueError(path)
    if "X_checksum" in inp:
        assert ds.checksum(X) == inp["X_checksum"], "generator output differs from the fixture"
        assert ds.checksum(y) == inp["y_checksum"], "generator labels differ from the fixture"
    return rec, X, y


def golden_paths(include_large=False):
    paths = sorted(p for p in glob.glob(os.path.join(GOLDEN, "*.json"))
                   if not p.endswith("losses_votes.json"))
    if not include_large:
        paths = [p for p in paths if not p.endswith("_oracle.json")]
    return paths


def candidate_error_bits(rec):
    """Every candidate's error at every step of an oracle fixture, as uint32
    bits [steps][d] (stored as a list or as zlib-compressed base64 of little-endian words)."""
    e, d = rec["expected"], rec["inputs"]["d"]
    if "candidate_errors_zb64" in e:
        raw = zlib.decompress(base64.b64decode(e["candidate_errors_zb64"]))
        a = np.frombuffer(raw, "<u4").astype(np.uint32)
    else:
        a = np.array(e["candidate_errors_bits"], np.uint32)
    return a.reshape(-1, d)


def expected_arrays(rec):
    e = rec["expected"]
    return dict(num_final=e["num_final"],
                selected_dimensions=np.array(e["selected_dimensions"], np.int32),
                selected_ordering=np.array(e["selected_ordering"], np.int32),
                selected_errors=np.array(e["selected_errors_bits"], np.uint32).view(np.float32),
                optimal_k=e["optimal_k"])


def assert_same_result(got, exp, what=""):
    assert got["num_final"] == exp["num_final"], what
    np.testing.assert_array_equal(got["selected_ordering"], exp["selected_ordering"], err_msg=what)
    np.testing.assert_array_equal(got["selected_dimensions"], exp["selected_dimensions"],
                                  err_msg=what)
    np.testing.assert_array_equal(np.asarray(got["selected_errors"], np.float32).view(np.uint32),
                                  np.asarray(exp["selected_errors"], np.float32).view(np.uint32),
                                  err_msg=what)
    assert got["optimal_k"] == exp["optimal_k"], what
