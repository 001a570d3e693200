"""Drop-in selector API: ``IncrementalNNFeatureSelection``.

Same constructor, defaults, ``train()`` and getters as the reference's
speedynn/base.py:9-121 (Python 2 there, Python 3 here); the work runs on the
MI355X through the C ABI in include/speedynn.h.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _lib


class IncrementalNNFeatureSelection:
    """Greedy kNN wrapper feature selection (forward or backward)."""

    # speedynn/base.py:13-23
    loss_dict = {
        "rmse": 0,
        "rmse_norm": 1,
        "mad": 2,
        "mad_norm": 3,
        "zero_one": 4,
        "std": 5,
        "rmse_norm_mad_norm_std": 6,
        "rmse_norm_mad_norm_galaxies": 7,
        "rmse_std": 8,
    }
    # speedynn/base.py:24-27
    model_types = {"regression": 0, "classification": 1}

    def __init__(self, Xtrain, Ytrain, k_vals=None, auto_feature=False, num_features=2,
                 k_fold=10, method="backward", loss_functions=["rmse"],
                 model_type="regression", use_gpu=True, form="precision", device=None):
        # speedynn/base.py:29-58
        self.Xtrain = np.array(Xtrain)
        self.Ytrain = np.asarray(Ytrain)
        self.dim = len(Xtrain[0])
        self.auto_feature = int(auto_feature)
        if k_vals is None:
            self.k_vals = np.array(range(1, 10))
        else:
            self.k_vals = np.array(k_vals)
        self.num_desired_features = num_features
        assert self.num_desired_features <= self.dim
        assert self.num_desired_features > 0
        if auto_feature:
            self.num_desired_features = self.dim
        self.dtype_int = np.int32
        self.dtype_float = np.float32
        self.Xtrain = np.ascontiguousarray(self.Xtrain.astype(self.dtype_float))
        self.Ytrain = np.ascontiguousarray(self.Ytrain.astype(self.dtype_float)).reshape(-1)
        self.k_vals = self.k_vals.astype(self.dtype_int)
        self.k_fold = k_fold
        self.method = method
        self.loss_functions = np.array([self.loss_dict[l] for l in loss_functions],
                                       dtype=self.dtype_int)
        self.model_type = self.model_types[model_type]
        self.use_gpu = use_gpu
        if not use_gpu:
            warnings.warn("speedynn_amd runs the selection on the GPU; use_gpu=False is accepted "
                          "for compatibility and gives the same results", stacklevel=2)
        if form not in _lib.FORMS:
            raise ValueError(f"form must be one of {sorted(_lib.FORMS)}")
        self.form = form
        self.device = device
        self.summary = None

    def _engine(self) -> _lib.Engine:
        if self.device is None:
            return _lib.default_engine()
        return _lib.Engine(self.device)

    def train(self):
        # speedynn/base.py:60-85
        if self.method not in ("backward", "forward"):
            raise ValueError("Unknown feature selection method: %r" % (self.method,))
        res = self._engine().select(
            self.Xtrain, self.Ytrain, self.k_vals, self.loss_functions,
            method=1 if self.method == "forward" else 0, model_type=self.model_type,
            k_fold=self.k_fold, num_features=self.num_desired_features,
            auto_feature=bool(self.auto_feature), form=_lib.FORMS[self.form])
        self.num_final_features = res["num_final"]
        self.selected_dimensions = res["selected_dimensions"]
        self.selected_ordering = res["selected_ordering"]
        flat = res["selected_errors"]
        nl = self.loss_functions.shape[0]
        errs = np.zeros((self.dim, nl), dtype=self.dtype_float)
        for l in range(nl):
            errs[:, l] = flat[l * self.dim:(l + 1) * self.dim]
        self.selected_errors = errs
        self.optimal_model_k = res["optimal_k"]
        self.summary = res
        return self

    def fit(self):
        """Alias of train() returning self."""
        return self.train()

    def get_selected_dimensions(self):
        return self.selected_dimensions

    def get_selected_ordering(self):
        return self.selected_ordering

    def get_feature_ranks(self):
        # speedynn/base.py:93-107
        ranks = np.zeros(self.dim)
        if self.method == "backward":
            removed = self.dim - self.num_final_features
            for i in range(removed):
                ranks[self.selected_ordering[self.dim - i - 1]] = -removed + i
            return ranks
        if self.method == "forward":
            for i in range(self.num_final_features):
                ranks[self.selected_ordering[i]] = self.num_final_features - i
            return ranks
        raise ValueError("Unknown feature selection method: %r" % (self.method,))

    def get_selected_errors(self):
        return self.selected_errors

    def get_optimal_model_k(self):
        return self.optimal_model_k

    def get_loss(self, labels, predictions, loss_func="rmse"):
        # speedynn/base.py:115-121
        code = self.loss_dict[loss_func]
        return self._engine().loss(np.asarray(labels, dtype=self.dtype_float),
                                   np.asarray(predictions, dtype=self.dtype_float), code)
