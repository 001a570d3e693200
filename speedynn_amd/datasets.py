"""Deterministic synthetic workloads for kNN wrapper feature selection.

Every random quantity comes from a counter-based splitmix64 stream, and the
Gaussian transform (Box-Muller) uses a logarithm and a cosine built from IEEE
add/multiply/divide/sqrt/frexp only, evaluated in float64 by numpy.  Those
operations are correctly rounded on every x86-64 host, so the same seed gives
bit-identical float32 arrays on any machine (no libm or SIMD-dispatch
differences).  Distribution quality is that of a textbook Box-Muller
transform; the polynomial log/cos are accurate to a few ulp of float64.

Workloads (n points, d features, k neighbours, leave-one-out):

* ``classification_sign``  -- y = sign(X_S w) in {-1,+1}, |S| informative
  features, w ~ N(0,1), sign(0) = +1, then exactly round(flip*n) labels flipped.
* ``classification_argmax`` -- y = argmax of C random projections of X_S,
  then round(flip*n) labels moved to a uniformly random other class.
* ``regression_sum`` -- y = sum of |S| informative columns + N(0, noise^2).

The five named configurations mirror the workload table of the selection
benchmark (n=2000/20000/65536/<=16000/32768).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_LN2 = 0.6931471805599453
_HALF_PI = 1.5707963267948966
_SQRT_HALF = 0.7071067811865476
_LOG2_10 = 3.321928094887362


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _stream_key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        k = _mix(np.array([np.uint64(seed) * _GAMMA + np.uint64(stream + 1) * _M1], dtype=np.uint64))
    return k[0]


def uint64s(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """``count`` raw 64-bit words of stream ``stream`` of ``seed``."""
    key = _stream_key(seed, stream)
    with np.errstate(over="ignore"):
        ctr = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        return _mix(key + ctr * _GAMMA)


def uniforms(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """Uniform float64 in the open interval (0, 1), 53-bit resolution."""
    w = uint64s(seed, stream, count, offset) >> np.uint64(11)
    return (w.astype(np.float64) + 0.5) * (2.0 ** -53)


def _log(x: np.ndarray) -> np.ndarray:
    """Natural log of x > 0 from frexp and the atanh series (basic ops only)."""
    m, e = np.frexp(x)
    small = m < _SQRT_HALF
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e).astype(np.float64)
    s = (m - 1.0) / (m + 1.0)          # |s| <= 0.1716
    s2 = s * s
    acc = np.full_like(s, 1.0 / 25.0)
    for j in range(11, -1, -1):
        acc = acc * s2 + 1.0 / (2 * j + 1)
    return 2.0 * s * acc + e * _LN2


def _sin_poly(a: np.ndarray) -> np.ndarray:
    a2 = a * a
    acc = np.full_like(a, 1.0)
    for j in range(12, 0, -1):
        acc = 1.0 - acc * a2 / ((2 * j) * (2 * j + 1))
    return a * acc


def _cos_poly(a: np.ndarray) -> np.ndarray:
    a2 = a * a
    acc = np.full_like(a, 1.0)
    for j in range(12, 0, -1):
        acc = 1.0 - acc * a2 / ((2 * j - 1) * (2 * j))
    return acc


def _cos_2pi(v: np.ndarray) -> np.ndarray:
    """cos(2*pi*v) for v in [0, 1): quadrant reduction, then Taylor series."""
    t = v * 4.0
    q = np.floor(t)
    r = (t - q) * _HALF_PI
    q = q.astype(np.int64) & 3
    c, s = _cos_poly(r), _sin_poly(r)
    return np.select([q == 0, q == 1, q == 2], [c, -s, -c], s)


def exp2(x: np.ndarray) -> np.ndarray:
    """2**x from ldexp and a Taylor series on the fractional part."""
    i = np.floor(x)
    f = (x - i) * _LN2
    acc = np.full_like(f, 1.0)
    for j in range(20, 0, -1):
        acc = 1.0 + acc * f / j
    return np.ldexp(acc, i.astype(np.int64))


def normals(seed: int, stream: int, count: int) -> np.ndarray:
    """Standard normal float64 samples (Box-Muller, first output only)."""
    u1 = uniforms(seed, 2 * stream, count)
    u2 = uniforms(seed, 2 * stream + 1, count)
    return np.sqrt(-2.0 * _log(u1)) * _cos_2pi(u2)


def randint(seed: int, stream: int, count: int, low: int, high: int) -> np.ndarray:
    """Integers uniform in [low, high)."""
    u = uniforms(seed, stream, count)
    return (low + np.floor(u * (high - low))).astype(np.int64)


def choice(seed: int, stream: int, population: int, k: int) -> np.ndarray:
    """k distinct integers from range(population) (partial Fisher-Yates)."""
    perm = np.arange(population, dtype=np.int64)
    u = uniforms(seed, stream, k)
    for i in range(k):
        j = i + int(np.floor(u[i] * (population - i)))
        perm[i], perm[j] = perm[j], perm[i]
    return perm[:k].copy()


def _ordered_dot(X: np.ndarray, cols: np.ndarray, w: np.ndarray) -> np.ndarray:
    """sum_f X[:, cols[f]] * w[f] in float64, summed in column order."""
    acc = np.zeros(X.shape[0], dtype=np.float64)
    for f, c in enumerate(cols):
        acc = acc + X[:, c].astype(np.float64) * w[f]
    return acc


# streams: 0 X, 1 informative set, 2 weights, 3 flip set, 4 flip target,
# 5 scales, 6 regression noise, 7 copy sources, 8 copy noise, 9 problem shapes
def gaussian_matrix(seed: int, n: int, d: int) -> np.ndarray:
    return normals(seed, 0, n * d).reshape(n, d).astype(np.float32)


def classification_sign(seed: int, n: int, d: int, n_informative: int, flip: float = 0.05,
                        X: np.ndarray | None = None, standardize: np.ndarray | None = None):
    if X is None:
        X = gaussian_matrix(seed, n, d)
    inf = choice(seed, 1, d, n_informative)
    w = normals(seed, 2, n_informative)
    Z = X if standardize is None else (X.astype(np.float64) / standardize[None, :])
    score = _ordered_dot(Z, inf, w)
    y = np.where(score >= 0.0, 1.0, -1.0)
    nflip = int(round(flip * n))
    if nflip:
        idx = choice(seed, 3, n, nflip)
        y[idx] = -y[idx]
    return X, y.astype(np.float32), inf


def classification_argmax(seed: int, n: int, d: int, n_informative: int, n_classes: int = 3,
                          flip: float = 0.05):
    X = gaussian_matrix(seed, n, d)
    inf = choice(seed, 1, d, n_informative)
    W = normals(seed, 2, n_classes * n_informative).reshape(n_classes, n_informative)
    scores = np.stack([_ordered_dot(X, inf, W[c]) for c in range(n_classes)], axis=1)
    y = np.argmax(scores, axis=1).astype(np.int64)
    nflip = int(round(flip * n))
    if nflip:
        idx = choice(seed, 3, n, nflip)
        shift = randint(seed, 4, nflip, 1, n_classes)
        y[idx] = (y[idx] + shift) % n_classes
    return X, y.astype(np.float32), inf


def regression_sum(seed: int, n: int, d: int, n_informative: int, noise: float = 0.5):
    X = gaussian_matrix(seed, n, d)
    inf = choice(seed, 1, d, n_informative)
    y = np.zeros(n, dtype=np.float64)
    for c in inf:
        y = y + X[:, c].astype(np.float64)
    y = y + noise * normals(seed, 6, n)
    return X, y.astype(np.float32), inf


def scaled_sign(seed: int, n: int, d: int, n_informative: int):
    """Features N(0,1) times scales 10^U(-2,2), labels as classification_sign
    on the standardised informative columns (the scales do not change them)."""
    scales = exp2(_LOG2_10 * (4.0 * uniforms(seed, 5, d) - 2.0))
    Z = normals(seed, 0, n * d).reshape(n, d)
    X = (Z * scales[None, :]).astype(np.float32)
    del Z
    X, y, inf = classification_sign(seed, n, d, n_informative, 0.05, X=X, standardize=scales)
    return X, y, inf, scales


def correlated_regression(seed: int, n: int, d_base: int, n_informative: int):
    """d_base N(0,1) columns followed by d_base copies of randomly chosen base
    columns plus N(0, 0.1^2) noise; y = sum of n_informative base columns plus
    N(0, 0.5^2)."""
    base = gaussian_matrix(seed, n, d_base)
    src = randint(seed, 7, d_base, 0, d_base)
    noise = normals(seed, 8, n * d_base).reshape(n, d_base)
    copies = (base[:, src].astype(np.float64) + 0.1 * noise).astype(np.float32)
    X = np.ascontiguousarray(np.concatenate([base, copies], axis=1))
    inf = choice(seed, 1, d_base, n_informative)
    y = np.zeros(n, dtype=np.float64)
    for c in inf:
        y = y + X[:, c].astype(np.float64)
    y = (y + 0.5 * normals(seed, 6, n)).astype(np.float32)
    return X, y, inf


def quantized(X: np.ndarray, step: float) -> np.ndarray:
    """Round features to a grid: produces many exactly tied distances."""
    return (np.round(X.astype(np.float64) / step) * step).astype(np.float32)


# --------------------------------------------------------------------------
# named workloads
# --------------------------------------------------------------------------

def config(name: str) -> dict:
    """Return the workload dict for one of c0..c4 (c3 returns a list of problems)."""
    if name == "c0":
        X, y, inf = classification_sign(0, 2000, 50, 5, 0.05)
        return dict(name=name, X=X, y=y, informative=inf, k_vals=[10], k_fold=2000,
                    num_features=10, method="forward", model_type="classification",
                    loss_functions=["zero_one"])
    if name == "c1":
        X, y, inf = classification_argmax(1, 20000, 200, 20, 3, 0.05)
        return dict(name=name, X=X, y=y, informative=inf, k_vals=[10], k_fold=20000,
                    num_features=20, method="forward", model_type="classification",
                    loss_functions=["zero_one"])
    if name == "c2":
        X, y, inf, scales = scaled_sign(2, 65536, 512, 32)
        return dict(name=name, X=X, y=y, informative=inf, scales=scales, k_vals=[20],
                    k_fold=65536, num_features=32, method="forward",
                    model_type="classification", loss_functions=["zero_one"])
    if name == "c4":
        X, y, inf = correlated_regression(4, 32768, 512, 16)
        return dict(name=name, X=X, y=y, informative=inf, k_vals=[15], k_fold=32768,
                    num_features=16, method="backward", model_type="regression",
                    loss_functions=["rmse"])
    raise ValueError(f"unknown workload {name!r}")


def reduced(name: str, n: int, d: int, num_features: int) -> dict:
    """A workload's recipe at another size: c2's log-scaled sign labels (d
    features, min(32, d) informative, forward to num_features) or c4's
    correlated copies (d/2 base columns and d/2 noisy copies, regression on
    min(16, d/2) base columns, backward to num_features); seeds as the full
    workload's, so only the sizes differ."""
    if name == "c2":
        X, y, inf, scales = scaled_sign(2, n, d, min(32, d))
        return dict(name=f"c2@{n}x{d}", X=X, y=y, informative=inf, scales=scales,
                    k_vals=[20], k_fold=n, num_features=num_features, method="forward",
                    model_type="classification", loss_functions=["zero_one"])
    if name == "c4":
        X, y, inf = correlated_regression(4, n, d // 2, min(16, d // 2))
        return dict(name=f"c4@{n}x{d}", X=X, y=y, informative=inf, k_vals=[15], k_fold=n,
                    num_features=num_features, method="backward", model_type="regression",
                    loss_functions=["rmse"])
    raise ValueError(f"no reduced recipe for {name!r}")


def batch_problem_shapes(count: int = 64, seed: int = 1000):
    """(n, d, k) of the batched workload: n~U{1000..16000}, d~U{32..256}, k in {5,10,20}."""
    ns = randint(seed, 9, count, 1000, 16001)
    ds = randint(seed, 10, count, 32, 257)
    ks = np.array([5, 10, 20])[randint(seed, 11, count, 0, 3)]
    return [(int(a), int(b), int(c)) for a, b, c in zip(ns, ds, ks)]


def batch_problem(p: int, count: int = 64):
    n, d, k = batch_problem_shapes(count)[p]
    X, y, inf = classification_sign(1000 + p, n, d, 8, 0.05)
    return dict(name=f"c3.{p}", X=X, y=y, informative=inf, k_vals=[k], k_fold=n,
                num_features=8, method="forward", model_type="classification",
                loss_functions=["zero_one"])


def checksum(a: np.ndarray) -> str:
    """Order-sensitive 64-bit FNV-1a over the raw bytes (hex)."""
    b = np.ascontiguousarray(a).reshape(-1).view(np.uint8)
    h = np.uint64(0xCBF29CE484222325)
    # fold 8-byte words for speed: exact and deterministic
    pad = (-b.size) % 8
    words = np.concatenate([b, np.zeros(pad, np.uint8)]).view(np.uint64)
    with np.errstate(over="ignore"):
        mixed = _mix(words + np.arange(words.size, dtype=np.uint64) * _GAMMA)
        acc = np.bitwise_xor.reduce(mixed) if mixed.size else np.uint64(0)
        h = _mix(np.array([h ^ acc ^ np.uint64(b.size)], dtype=np.uint64))[0]
    return f"{int(h):016x}"
