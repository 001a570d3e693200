"""The reference's SWIG module surface (nearest_neighbors_feature_selection_gpu,
speedynn/lib/nearest_neighbors_feature_selection_float_gpu.i:1-24) over the C ABI.

Same free functions and argument conventions as the SWIG bindings: arrays are
borrowed in place and must be C-contiguous with the exact dtype (float32 /
int32), otherwise TypeError, as numpy.i's INPLACE typemaps require; results
are process-global and read back by the getters.  Errors raise instead of
terminating the interpreter.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def _inplace(a, dtype, ndim, name):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or a.ndim != ndim or \
            not a.flags.c_contiguous:
        raise TypeError(f"{name}: expected a C-contiguous {ndim}-d {np.dtype(dtype).name} array")
    return a


def _train(fn, Xtrain, Ytrain, kvals, num_features, kfold, loss_functions, auto_feature,
           model_type):
    X = _inplace(Xtrain, np.float32, 2, "Xtrain")
    Y = _inplace(Ytrain, np.float32, 1, "Ytrain")
    kv = _inplace(kvals, np.int32, 1, "kvals")
    lf = _inplace(loss_functions, np.int32, 1, "loss_functions")
    n, d = X.shape
    return _lib.check(fn(X, n, d, Y, Y.shape[0], kv, kv.shape[0], int(num_features), int(kfold),
                         lf, lf.shape[0], int(auto_feature), int(model_type)))


def train_forward_selection(Xtrain, Ytrain, kvals, num_features, kfold, loss_functions,
                            auto_feature, model_type):
    return _train(_lib.lib().train_forward_selection, Xtrain, Ytrain, kvals, num_features, kfold,
                  loss_functions, auto_feature, model_type)


def train_backward_selection(Xtrain, Ytrain, kvals, num_features, kfold, loss_functions,
                             auto_feature, model_type):
    return _train(_lib.lib().train_backward_selection, Xtrain, Ytrain, kvals, num_features,
                  kfold, loss_functions, auto_feature, model_type)


def get_selected_dimensions(out):
    a = _inplace(out, np.int32, 1, "selected_dimensions")
    _lib.lib().get_selected_dimensions(a, a.shape[0])


def get_selected_ordering(out):
    a = _inplace(out, np.int32, 1, "selected_ordering")
    _lib.lib().get_selected_ordering(a, a.shape[0])


def get_selected_errors(out):
    a = _inplace(out, np.float32, 1, "selected_errors")
    _lib.lib().get_selected_errors(a, a.shape[0])


def get_final_optimal_model_k():
    return int(_lib.lib().get_final_optimal_model_k())


def get_loss_all(labels, predictions, loss_function):
    y = _inplace(labels, np.float32, 1, "labels")
    p = _inplace(predictions, np.float32, 1, "predictions")
    return float(_lib.lib().get_loss_all(y, y.shape[0], p, p.shape[0], int(loss_function)))
