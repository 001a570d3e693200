"""ctypes binding of the CPU oracle (oracle.c) and of the reference's compiled
CPU module (oracle/_ref, built from the reference sources by ``make ref``).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg.  The engine in speedynn_amd/ never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libspeedynn_ref.so")
REF_SRC = "/root/reference/speedynn/lib"

LOSS = {"rmse": 0, "rmse_norm": 1, "mad": 2, "mad_norm": 3, "zero_one": 4, "std": 5,
        "rmse_norm_mad_norm_std": 6, "rmse_norm_mad_norm_galaxies": 7, "rmse_std": 8}
MODEL = {"regression": 0, "classification": 1}

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_lib = None


def build(force: bool = False, with_ref: bool | None = None) -> None:
    """Compile liboracle.so (and _ref/ when the reference sources exist)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B", "liboracle.so"], cwd=HERE)
    if with_ref is None:
        with_ref = os.path.isdir(REF_SRC)
    if with_ref and (force or not os.path.exists(REF_PATH)):
        subprocess.check_call(["make", "-s", "-B", "ref"], cwd=HERE,
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_select.restype = C.c_int
        L.oracle_select.argtypes = [C.c_int, _f32p, C.c_int64, C.c_int, _f32p, _i32p, C.c_int,
                                    C.c_int, C.c_int64, _i32p, C.c_int, C.c_int, C.c_int, C.c_int,
                                    _i32p, _i32p, _f32p, C.POINTER(C.c_int), C.c_void_p, C.c_void_p]
        L.oracle_distance_matrix.argtypes = [_f32p, C.c_int64, C.c_int, _f32p, C.c_int]
        L.oracle_update.argtypes = [_f32p, C.c_int64, C.c_int, _f32p, C.c_int, C.c_int, C.c_int]
        L.oracle_distance_row.argtypes = [_f32p, C.c_int64, C.c_int, C.c_int64, C.c_int, _i32p,
                                          C.c_int, _f32p]
        L.oracle_distance_row_f64.argtypes = [_f32p, C.c_int64, C.c_int, C.c_int64, _i32p, C.c_int,
                                              _f64p]
        L.oracle_candidate_preds.argtypes = [_f32p, C.c_int64, _f32p, C.c_int64, C.c_int, _f32p,
                                             _i32p, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int,
                                             C.c_void_p, C.c_int64, _f32p, C.c_int]
        L.oracle_mode.restype = C.c_float
        L.oracle_mode.argtypes = [_f32p, C.c_int]
        L.oracle_loss.restype = C.c_float
        L.oracle_loss.argtypes = [_f32p, _f32p, C.c_int64, C.c_int]
        L.oracle_max_threads.restype = C.c_int
        _lib = L
    return _lib


def max_threads() -> int:
    return lib().oracle_max_threads()


def _codes(names, table):
    return np.ascontiguousarray([table[x] if isinstance(x, str) else int(x) for x in names],
                                dtype=np.int32)


def select(X, y, k_vals, num_features, k_fold, loss_functions, auto_feature, model_type,
           method, threads: int | None = None, record_candidates: bool = False) -> dict:
    """Run the restated selection; returns the reference getters' values."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n, d = X.shape
    kv = np.ascontiguousarray(k_vals, dtype=np.int32)
    losses = _codes(loss_functions, LOSS)
    mt = MODEL[model_type] if isinstance(model_type, str) else int(model_type)
    meth = {"forward": 1, "backward": 0}[method] if isinstance(method, str) else int(method)
    sel = np.zeros(d, np.int32)
    order = np.zeros(d, np.int32)
    errs = np.zeros(len(losses) * d, np.float32)
    best_k = C.c_int(0)
    steps = d
    cand_err = np.zeros(steps * d, np.float32) if record_candidates else None
    cand_bk = np.zeros(steps * d, np.int32) if record_candidates else None
    nthreads = threads or max_threads()
    nf = lib().oracle_select(meth, X, n, d, y, kv, len(kv), int(num_features), int(k_fold),
                             losses, len(losses), int(bool(auto_feature)), mt, nthreads, sel,
                             order, errs, C.byref(best_k),
                             cand_err.ctypes.data if record_candidates else None,
                             cand_bk.ctypes.data if record_candidates else None)
    if nf < 0:
        raise ValueError(f"oracle_select rejected the arguments (code {nf})")
    out = dict(num_final=nf, selected_dimensions=sel, selected_ordering=order,
               selected_errors=errs, optimal_k=best_k.value)
    if record_candidates:
        out["candidate_errors"] = cand_err.reshape(steps, d)
        out["candidate_best_k"] = cand_bk.reshape(steps, d)
    return out


def distance_matrix(X, threads: int | None = None) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    D = np.empty((n, n), np.float32)
    lib().oracle_distance_matrix(X, n, d, D.reshape(-1), threads or max_threads())
    return D


def update(X, D, f, mult, threads: int | None = None) -> None:
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    lib().oracle_update(X, n, d, D.reshape(-1), int(f), int(mult), threads or max_threads())


def distance_row(X, q, method, order) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    o = np.ascontiguousarray(order, dtype=np.int32)
    row = np.empty(n, np.float32)
    meth = {"forward": 1, "backward": 0}[method] if isinstance(method, str) else int(method)
    lib().oracle_distance_row(X, n, d, int(q), meth, o, len(o), row)
    return row


def distance_row_f64(X, q, feats) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    fs = np.ascontiguousarray(feats, dtype=np.int32)
    row = np.empty(n, np.float64)
    lib().oracle_distance_row_f64(X, n, d, int(q), fs, len(fs), row)
    return row


def candidate_preds(Drows, X, y, k_vals, k_fold, f, mult, model_type, queries=None,
                    threads: int | None = None) -> np.ndarray:
    """Predictions [len(k_vals)][nq] for candidate f.  Drows holds one D row per
    query (nq x n) when ``queries`` is given, else the full n x n matrix."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n, d = X.shape
    kv = np.ascontiguousarray(k_vals, dtype=np.int32)
    Dr = np.ascontiguousarray(Drows, dtype=np.float32)
    mt = MODEL[model_type] if isinstance(model_type, str) else int(model_type)
    if queries is None:
        nq, qp = n, None
    else:
        qa = np.ascontiguousarray(queries, dtype=np.int64)
        nq, qp = len(qa), qa.ctypes.data
    preds = np.empty(len(kv) * nq, np.float32)
    lib().oracle_candidate_preds(Dr.reshape(-1), n, X, n, d, y, kv, len(kv), int(k_fold),
                                 int(f), int(mult), mt, qp, nq, preds, threads or max_threads())
    return preds.reshape(len(kv), nq)


def mode(labels) -> float:
    a = np.ascontiguousarray(labels, dtype=np.float32)
    return float(lib().oracle_mode(a, len(a)))


def loss(labels, preds, loss_function) -> float:
    y = np.ascontiguousarray(labels, dtype=np.float32)
    p = np.ascontiguousarray(preds, dtype=np.float32)
    code = LOSS[loss_function] if isinstance(loss_function, str) else int(loss_function)
    return float(lib().oracle_loss(y, p, len(y), code))


# --------------------------------------------------------------------------
# the reference's own CPU module (only where it was built from the sources)
# --------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_select(X, y, k_vals, num_features, k_fold, loss_functions, auto_feature, model_type,
               method) -> dict:
    """Call the compiled reference (train_*_selection + getters), mirroring
    the call sequence of speedynn/base.py:60-85."""
    R = C.CDLL(REF_PATH)
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n, d = X.shape
    kv = np.ascontiguousarray(k_vals, dtype=np.int32)
    losses = _codes(loss_functions, LOSS)
    mt = MODEL[model_type] if isinstance(model_type, str) else int(model_type)
    fn = R.train_forward_selection if method == "forward" else R.train_backward_selection
    fn.restype = C.c_int
    fn.argtypes = [_f32p, C.c_int, C.c_int, _f32p, C.c_int, _i32p, C.c_int, C.c_int, C.c_int,
                   _i32p, C.c_int, C.c_int, C.c_int]
    nf = fn(X, n, d, y, n, kv, len(kv), int(num_features), int(k_fold), losses, len(losses),
            int(bool(auto_feature)), mt)
    sel = np.zeros(d, np.int32)
    order = np.zeros(d, np.int32)
    errs = np.zeros(len(losses) * d, np.float32)
    R.get_selected_dimensions.argtypes = [_i32p, C.c_int]
    R.get_selected_ordering.argtypes = [_i32p, C.c_int]
    R.get_selected_errors.argtypes = [_f32p, C.c_int]
    R.get_selected_dimensions(sel, d)
    R.get_selected_ordering(order, d)
    R.get_selected_errors(errs, len(errs))
    R.get_final_optimal_model_k.restype = C.c_int
    return dict(num_final=nf, selected_dimensions=sel, selected_ordering=order,
                selected_errors=errs, optimal_k=R.get_final_optimal_model_k())


def ref_loss(labels, preds, loss_function) -> float:
    R = C.CDLL(REF_PATH)
    y = np.ascontiguousarray(labels, dtype=np.float32)
    p = np.ascontiguousarray(preds, dtype=np.float32)
    R.get_loss_all.restype = C.c_float
    R.get_loss_all.argtypes = [_f32p, C.c_int, _f32p, C.c_int, C.c_int]
    code = LOSS[loss_function] if isinstance(loss_function, str) else int(loss_function)
    return float(R.get_loss_all(y, len(y), p, len(p), code))


def ref_mode(labels) -> float:
    R = C.CDLL(REF_PATH)
    a = np.ascontiguousarray(labels, dtype=np.float32)
    R.get_mode_array.restype = C.c_float
    R.get_mode_array.argtypes = [_f32p, C.c_int]
    return float(R.get_mode_array(a, len(a)))
