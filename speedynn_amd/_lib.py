"""ctypes binding of libspeedynn_amd.so (the C ABI declared in include/speedynn.h).

The library is the only compute path: when it is missing or cannot open a
HIP device, every call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPEEDYNN_LIB names another build of the same library (A/B measurements).
LIB_PATH = os.environ.get("SPEEDYNN_LIB") or os.path.join(HERE, "libspeedynn_amd.so")

SNN_OK, SNN_EINVAL, SNN_ENOMEM, SNN_EDEVICE, SNN_ESTATE, SNN_ENOTSUP = 0, -1, -2, -3, -4, -5
FORMS = {"plain": 0, "precision": 1}

# Every C symbol include/speedynn.h declares.
EXPORTS = [
    "train_backward_selection", "train_forward_selection", "get_selected_dimensions",
    "get_selected_ordering", "get_selected_errors", "get_final_optimal_model_k", "get_loss_all",
    "snn_create", "snn_destroy", "snn_last_error", "snn_select", "snn_select_device",
    "snn_distance_rows", "snn_loss", "snn_set_default_form", "snn_select_batch",
    "snn_distance_matrix", "snn_select_device_noalloc",
]
NOALLOC_SCRATCH_BYTES = 64 << 20  # SNN_NOALLOC_SCRATCH_BYTES


class SnnParams(C.Structure):
    _fields_ = [("method", C.c_int), ("model_type", C.c_int), ("k_fold", C.c_int64),
                ("num_features", C.c_int), ("auto_feature", C.c_int), ("form", C.c_int),
                ("use_graph", C.c_int), ("compensated", C.c_int)]


class SnnSummary(C.Structure):
    _fields_ = [("num_final", C.c_int), ("optimal_k", C.c_int), ("steps", C.c_int),
                ("pruned_kernel", C.c_int), ("ms_total", C.c_double),
                ("ms_distance", C.c_double),
                ("ms_search", C.c_double), ("ms_update", C.c_double), ("ms_tally", C.c_double),
                ("evals", C.c_int64),
                ("pruned_point_evals", C.c_int64), ("ms_bound", C.c_double),
                ("ms_search_first", C.c_double)]


class SnnProblem(C.Structure):
    _fields_ = [("X", C.c_void_p), ("y", C.c_void_p), ("n", C.c_int64), ("d", C.c_int),
                ("k_vals", C.c_void_p), ("num_k", C.c_int), ("loss_functions", C.c_void_p),
                ("num_loss", C.c_int), ("params", SnnParams),
                ("selected_dimensions", C.c_void_p), ("selected_ordering", C.c_void_p),
                ("selected_errors", C.c_void_p), ("status", C.c_int), ("num_final", C.c_int),
                ("optimal_k", C.c_int), ("steps", C.c_int), ("inputs_on_device", C.c_int)]


class EngineError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"speedynn_amd error {code}: {message}")
        self.code = code


_lib = None
_lock = threading.RLock()  # default_engine() takes it and then calls lib()
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def lib() -> C.CDLL:
    """Load the engine library, raising if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"speedynn_amd: {LIB_PATH} is missing; build it with "
                "`python -m speedynn_amd.build` (hipcc, gfx950). There is no CPU fallback.")
        L = C.CDLL(LIB_PATH)
        L.snn_last_error.restype = C.c_char_p
        L.snn_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        L.snn_destroy.argtypes = [C.c_void_p]
        L.snn_destroy.restype = None
        L.snn_select.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_int, _f32p, _i32p, C.c_int,
                                 _i32p, C.c_int, C.POINTER(SnnParams), _i32p, _i32p, _f32p,
                                 C.c_void_p, C.POINTER(SnnSummary)]
        L.snn_select_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p,
                                        _i32p, C.c_int, _i32p, C.c_int, C.POINTER(SnnParams),
                                        _i32p, _i32p, _f32p, C.c_void_p, C.POINTER(SnnSummary)]
        L.snn_select_device_noalloc.argtypes = L.snn_select_device.argtypes
        L.snn_distance_rows.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_int, C.c_int, _i32p,
                                        C.c_int, C.c_int, _i64p, C.c_int, _f32p, C.c_void_p]
        L.snn_loss.argtypes = [C.c_void_p, _f32p, _f32p, C.c_int64, C.c_int,
                               C.POINTER(C.c_float)]
        L.snn_set_default_form.argtypes = [C.c_int]
        L.snn_set_default_form.restype = None
        L.snn_distance_matrix.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                          C.c_void_p, C.POINTER(C.c_double)]
        L.snn_select_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(SnnProblem),
                                       C.POINTER(SnnSummary)]
        i = C.c_int
        for name in ("train_backward_selection", "train_forward_selection"):
            fn = getattr(L, name)
            fn.restype = i
            fn.argtypes = [_f32p, i, i, _f32p, i, _i32p, i, i, i, _i32p, i, i, i]
        L.get_selected_dimensions.argtypes = [_i32p, i]
        L.get_selected_ordering.argtypes = [_i32p, i]
        L.get_selected_errors.argtypes = [_f32p, i]
        for name in ("get_selected_dimensions", "get_selected_ordering", "get_selected_errors"):
            getattr(L, name).restype = None
        L.get_final_optimal_model_k.restype = i
        L.get_final_optimal_model_k.argtypes = []
        L.get_loss_all.restype = C.c_float
        L.get_loss_all.argtypes = [_f32p, i, _f32p, i, i]
        _lib = L
        return L


def last_error() -> str:
    return lib().snn_last_error().decode(errors="replace")


def check(code: int) -> int:
    if code < 0:
        raise EngineError(code, last_error())
    return code


class Engine:
    """One device context (stream + cached device buffers)."""

    def __init__(self, device: int = 0):
        self._ctx = C.c_void_p()
        check(lib().snn_create(C.byref(self._ctx), int(device)))
        self.device = device

    def close(self) -> None:
        if self._ctx:
            lib().snn_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def select(self, X, y, k_vals, loss_codes, *, method: int, model_type: int, k_fold: int,
               num_features: int, auto_feature: bool = False, form: int = 1,
               candidate_errors: bool = False, device_inputs=None,
               compensated: bool = False, noalloc: bool = False,
               use_graph: bool = False) -> dict:
        """Run one selection; device_inputs=(X_ptr, y_ptr, n, d) skips the upload.
        noalloc (device inputs only): snn_select_device_noalloc, whose device
        work memory is the library's fixed array.  use_graph: steps replayed
        from hipGraphs kept in the context (no phase times)."""
        kv = np.ascontiguousarray(k_vals, dtype=np.int32)
        lc = np.ascontiguousarray(loss_codes, dtype=np.int32)
        if device_inputs is None:
            X = np.ascontiguousarray(X, dtype=np.float32)
            y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
            n, d = X.shape
        else:
            xp, yp, n, d = device_inputs
        p = SnnParams(method=int(method), model_type=int(model_type), k_fold=int(k_fold),
                      num_features=int(num_features), auto_feature=int(bool(auto_feature)),
                      form=int(form), use_graph=int(bool(use_graph)),
                      compensated=int(bool(compensated)))
        sel = np.zeros(d, np.int32)
        order = np.zeros(d, np.int32)
        errs = np.zeros(len(lc) * d, np.float32)
        cand = np.zeros(d * d, np.float32) if candidate_errors else None
        sm = SnnSummary()
        cptr = cand.ctypes.data if cand is not None else None
        if noalloc and device_inputs is None:
            raise ValueError("noalloc selections take device inputs")
        if device_inputs is None:
            rc = lib().snn_select(self._ctx, X, n, d, y, kv, len(kv), lc, len(lc), C.byref(p),
                                  sel, order, errs, cptr, C.byref(sm))
        else:
            fn = lib().snn_select_device_noalloc if noalloc else lib().snn_select_device
            rc = fn(self._ctx, C.c_void_p(xp), n, d, C.c_void_p(yp), kv, len(kv), lc, len(lc),
                    C.byref(p), sel, order, errs, cptr, C.byref(sm))
        check(rc)
        out = dict(num_final=sm.num_final, optimal_k=sm.optimal_k, steps=sm.steps,
                   selected_dimensions=sel, selected_ordering=order, selected_errors=errs,
                   ms_total=sm.ms_total, ms_distance=sm.ms_distance, ms_search=sm.ms_search,
                   ms_update=sm.ms_update, ms_tally=sm.ms_tally, evals=sm.evals,
                   pruned_point_evals=sm.pruned_point_evals, ms_bound=sm.ms_bound,
                   pruned_kernel=sm.pruned_kernel,
                   ms_search_first=sm.ms_search_first)
        if cand is not None:
            out["candidate_errors"] = cand.reshape(d, d)
        return out

    def select_batch(self, problems, *, form: int = 1, use_graph: bool = True,
                     raise_on_error: bool = True):
        """Run many selections together (snn_select_batch).

        problems: list of dicts with X, y, k_vals, loss_codes, method, model_type,
        k_fold, num_features and optionally auto_feature.  Returns (results,
        summary): one dict per problem (status, num_final, optimal_k, steps,
        selected_dimensions, selected_ordering, selected_errors) and the batch
        totals."""
        keep = []  # arrays referenced by the C structs stay alive for the call
        arr = (SnnProblem * len(problems))()
        outs = []
        for i, pr in enumerate(problems):
            kv = np.ascontiguousarray(pr["k_vals"], dtype=np.int32)
            lc = np.ascontiguousarray(pr["loss_codes"], dtype=np.int32)
            q = arr[i]
            if pr.get("device_inputs") is not None:  # (X_ptr, y_ptr, n, d) in device memory
                xp, yp, n, d = pr["device_inputs"]
                q.X, q.y, q.inputs_on_device = xp, yp, 1
            else:
                X = np.ascontiguousarray(pr["X"], dtype=np.float32)
                y = np.ascontiguousarray(pr["y"], dtype=np.float32).reshape(-1)
                n, d = X.shape
                keep += [X, y]
                q.X, q.y = X.ctypes.data, y.ctypes.data
            sel = np.zeros(d, np.int32)
            order = np.zeros(d, np.int32)
            errs = np.zeros(len(lc) * d, np.float32)
            keep += [kv, lc]
            outs.append((sel, order, errs))
            q.n, q.d = n, d
            q.k_vals, q.num_k = kv.ctypes.data, len(kv)
            q.loss_functions, q.num_loss = lc.ctypes.data, len(lc)
            q.params = SnnParams(method=int(pr["method"]), model_type=int(pr["model_type"]),
                                 k_fold=int(pr["k_fold"]), num_features=int(pr["num_features"]),
                                 auto_feature=int(bool(pr.get("auto_feature", False))),
                                 form=int(form), use_graph=int(bool(use_graph)),
                                 compensated=int(bool(pr.get("compensated", False))))
            q.selected_dimensions = sel.ctypes.data
            q.selected_ordering = order.ctypes.data
            q.selected_errors = errs.ctypes.data
        sm = SnnSummary()
        rc = lib().snn_select_batch(self._ctx, len(problems), arr, C.byref(sm))
        if raise_on_error:
            check(rc)
        results = []
        for i, (sel, order, errs) in enumerate(outs):
            q = arr[i]
            results.append(dict(status=q.status, num_final=q.num_final, optimal_k=q.optimal_k,
                                steps=q.steps, selected_dimensions=sel,
                                selected_ordering=order, selected_errors=errs))
        summary = dict(rc=rc, steps=sm.steps, ms_total=sm.ms_total, ms_distance=sm.ms_distance,
                       ms_search=sm.ms_search, ms_search_first=sm.ms_search_first,
                       ms_bound=sm.ms_bound, evals=sm.evals,
                       pruned_point_evals=sm.pruned_point_evals)
        return results, summary

    def distance_rows(self, X, method: int, order, rows, form: int = 0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, d = X.shape
        o = np.ascontiguousarray(order, dtype=np.int32)
        r = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.zeros((len(r), n), np.float32)
        lo = np.zeros((len(r), n), np.float32) if form == 1 else None
        check(lib().snn_distance_rows(self._ctx, X, n, d, int(method), o, len(o), int(form), r,
                                      len(r), out.reshape(-1),
                                      lo.ctypes.data if lo is not None else None))
        return out, lo

    def distance_matrix(self, X_ptr: int, n: int, d: int, form: int, D_ptr: int) -> float:
        """Full-d distance matrix of device-resident X into device D (n*n
        floats); form 0 = exact FP32 in the reference's order, 2 = BF16x3
        MFMA Gram.  Returns the kernels' device time in ms."""
        ms = C.c_double()
        check(lib().snn_distance_matrix(self._ctx, C.c_void_p(X_ptr), int(n), int(d), int(form),
                                        C.c_void_p(D_ptr), C.byref(ms)))
        return ms.value

    def loss(self, labels, preds, code: int) -> float:
        y = np.ascontiguousarray(labels, dtype=np.float32)
        p = np.ascontiguousarray(preds, dtype=np.float32)
        out = C.c_float()
        check(lib().snn_loss(self._ctx, y, p, len(y), int(code), C.byref(out)))
        return float(out.value)


class DeviceMemory:
    """Minimal libamdhip64 binding for device-resident inputs (snn_select_device)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        if self.L.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        return p.value

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        if self.L.hipMemcpy(C.c_void_p(p), a.ctypes.data, a.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy failed")
        return p

    def copy_d2d(self, dst: int, src: int, nbytes: int) -> None:
        if self.L.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, 3) != 0:
            raise RuntimeError("hipMemcpy failed")

    def free(self, p: int) -> None:
        self.L.hipFree(C.c_void_p(p))

    def sync(self) -> None:
        self.L.hipDeviceSynchronize()


_default_engine = None


def default_engine() -> Engine:
    global _default_engine
    with _lock:
        if _default_engine is None:
            _default_engine = Engine(int(os.environ.get("SPEEDYNN_DEVICE", "0")))
        return _default_engine
