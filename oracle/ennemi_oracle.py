"""CPU restatement of ennemi's nearest-neighbour estimators (the oracle).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product in
``ennemi_amd``.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` may import it; the product never does.

It restates, function by function, the reference files

* ``ennemi/_entropy_estimators.py`` (estimators, ``_psi``), and
* ``ennemi/_driver.py`` (lag windowing, masks, NaN drop, rescaling, task
  enumeration, result assembly),

with a selectable neighbour backend for the two SciPy calls the reference
makes (``cKDTree.query(k=[k+1], p=inf)`` and
``cKDTree.query_ball_point(r, p=inf, return_length=True)``):

* ``"brute"``  -- the C brute force in ``oracle_knn.c`` (exact definition);
* ``"numpy"``  -- a pure NumPy brute force for tiny inputs;
* ``"kdtree"`` -- ``scipy.spatial.cKDTree`` called exactly as the reference
  calls it.  This is the CPU baseline ("port") timed by ``bench.py``.

Pins (see ``tests/test_oracle.py``): the three backends agree bit for bit on
radii and counts, including exact ties, duplicates and negative radii; ``psi``
is checked against closed forms and ``scipy.special.psi``; the estimator
formulas against analytic values; and the composed driver path against the
golden vectors in ``tests/golden`` that were written by running the reference
in the build container.
"""

from __future__ import annotations

import ctypes
import itertools
import math
import os
import warnings
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_knn.so")
_lib = None

DISCRETE_NORMALIZATION_WARNING = (
    "You have set normalize=True while at least one variable is discrete. "
    "The correlation coefficient formula assumes both variables to be continuous, "
    "and the results are not comparable across different discrete variables. "
    "Compare the raw MI against entropy of the discrete variable instead. "
    "If you really want to calculate correlation coefficients, you can suppress "
    "this warning by setting normalize=False and calling normalize_mi() on the results.")

PREPROCESS_CONSTANT_DATA_WARNING = (
    "A variable not marked as discrete takes only a single value, "
    "or values in a very small numerical range. "
    "If this is intentional, you can suppress this warning by passing preprocess=False. "
    "Note that this disables rescaling on all other variables as well.")


def build_oracle() -> str:
    """Compile oracle_knn.c with the committed Makefile; return the .so path."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build_oracle()
        lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_int64)
        lib.orc_knn_eps.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, dp]
        lib.orc_knn_eps.restype = None
        lib.orc_ball_count.argtypes = [dp, ctypes.c_int64, dp, ctypes.c_int64,
                                       ctypes.c_int32, dp, lp]
        lib.orc_ball_count.restype = None
        _lib = lib
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------------------
# Neighbour primitives (replace cKDTree.query / query_ball_point)
# ---------------------------------------------------------------------------

def _as_points(a) -> np.ndarray:
    """np.column_stack((a,)) as the reference does (_entropy_estimators.py:93-94)."""
    return np.ascontiguousarray(np.column_stack((np.asarray(a, dtype=np.float64),)))


def knn_eps(points, k: int, backend: str = "brute") -> np.ndarray:
    """(k+1)-th smallest max-norm distance from each point, self included.

    Restates ``cKDTree(points).query(points, k=[k+1], p=np.inf)[0].flatten()``
    (ennemi/_entropy_estimators.py:39,108,142,194).  Fewer than k+1 points
    gives +inf, as cKDTree does.
    """
    pts = _as_points(points)
    n, d = pts.shape
    if backend == "kdtree":
        from scipy.spatial import cKDTree
        return cKDTree(pts).query(pts, k=[k + 1], p=np.inf)[0].flatten()
    if backend == "numpy":
        if n <= k:
            return np.full(n, np.inf)
        dist = np.max(np.abs(pts[:, None, :] - pts[None, :, :]), axis=2)
        return np.sort(dist, axis=1)[:, k].copy()
    eps = np.empty(n)
    _load().orc_knn_eps(_dptr(pts), n, d, k, _dptr(eps))
    return eps


def ball_counts(queries, candidates, r, backend: str = "brute") -> np.ndarray:
    """Inclusive max-norm ball counts, ``#{j : D(q_i, c_j) <= r_i}``.

    Restates ``cKDTree(c).query_ball_point(q, r, p=np.inf, return_length=True)``
    (ennemi/_entropy_estimators.py:109-110,152-154,196).  Negative radii give 0.
    """
    q = _as_points(queries)
    c = _as_points(candidates)
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
    if backend == "kdtree":
        from scipy.spatial import cKDTree
        return np.asarray(cKDTree(c).query_ball_point(q, r, p=np.inf, return_length=True))
    if backend == "numpy":
        dist = np.max(np.abs(q[:, None, :] - c[None, :, :]), axis=2)
        return np.sum(dist <= r[:, None], axis=1).astype(np.int64)
    out = np.empty(q.shape[0], dtype=np.int64)
    _load().orc_ball_count(_dptr(q), q.shape[0], _dptr(c), c.shape[0], q.shape[1],
                           _dptr(r), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out


# ---------------------------------------------------------------------------
# Digamma (ennemi/_entropy_estimators.py:327-350)
# ---------------------------------------------------------------------------

def psi(x) -> np.ndarray:
    """The reference's truncated digamma series, for non-negative integers.

    Any zero anywhere makes the whole result the scalar +inf (line 338-339);
    psi(1) is the SciPy constant (342-343); otherwise the asymptotic series
    ``ln y - y^-6 (y^2 (y^2 (y/2 + 1/12) - 1/120) + 1/252)`` (347-348).
    """
    x = np.asarray(x)
    if np.any(x == 0):
        return np.asarray(np.inf)
    mask = (x != 1)
    result = np.full(x.shape, -0.5772156649015331)
    y = np.asarray(x[mask], dtype=np.float64)
    result[mask] = np.log(y) - np.power(y, -6) * (
        np.power(y, 2) * (np.power(y, 2) * (y / 2 + 1 / 12) - 1 / 120) + 1 / 252)
    return result


# ---------------------------------------------------------------------------
# Estimators (ennemi/_entropy_estimators.py)
# ---------------------------------------------------------------------------

def estimate_single_entropy(x, k: int = 3, backend: str = "brute", detail: bool = False):
    """Kozachenko-Leonenko entropy (ennemi/_entropy_estimators.py:21-42)."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x.reshape((x.size, 1))
    n, ndim = x.shape
    eps = knn_eps(x, k, backend)
    with np.errstate(divide="ignore"):
        value = psi(n) - psi(k) + ndim * (np.mean(np.log(eps)) + np.log(2))
    value = float(value)
    return (value, {"eps": eps}) if detail else value


def estimate_single_mi(x, y, k: int = 3, backend: str = "brute", detail: bool = False):
    """KSG estimator (ennemi/_entropy_estimators.py:69-113)."""
    n = len(x)
    x = np.column_stack((x,))
    y = np.column_stack((y,))
    xy = np.column_stack((x, y))
    eps = knn_eps(xy, k, backend)
    r = eps - 1e-12
    nx = ball_counts(x, x, r, backend)
    ny = ball_counts(y, y, r, backend)
    value = float(psi(n) + psi(k) - np.mean(psi(nx) + psi(ny)))
    return (value, {"eps": eps, "nx": nx, "ny": ny}) if detail else value


def estimate_conditional_mi(x, y, cond, k: int = 3, backend: str = "brute",
                            detail: bool = False):
    """Frenzel-Pompe conditional MI (ennemi/_entropy_estimators.py:116-156)."""
    cond = np.column_stack((cond,))
    xyz = np.column_stack((x, y, cond))
    eps = knn_eps(xyz, k, backend)
    xz = np.column_stack((x, cond))
    yz = np.column_stack((y, cond))
    r = eps - 1e-12
    nxz = ball_counts(xz, xz, r, backend)
    nyz = ball_counts(yz, yz, r, backend)
    nz = ball_counts(cond, cond, r, backend)
    with np.errstate(invalid="ignore"):
        value = float(psi(k) - np.mean(psi(nxz) + psi(nyz) - psi(nz)))
    return (value, {"eps": eps, "nxz": nxz, "nyz": nyz, "nz": nz}) if detail else value


def estimate_semidiscrete_mi(x, y, k: int = 3, backend: str = "brute", detail: bool = False):
    """Ross discrete-continuous MI (ennemi/_entropy_estimators.py:159-200)."""
    n = len(x)
    x = np.column_stack((x,))
    y_values, y_counts = np.unique(y, return_counts=True)
    if len(y_values) > n / 4:
        warnings.warn("The discrete variable has relatively many unique values." +
                      " Did you pass y and x in correct order?", UserWarning)
    n_full = np.empty(n)
    eps_full = np.empty(n)
    for val in y_values:
        subset = x[y == val]
        eps = knn_eps(subset, k, backend)
        eps_full[y == val] = eps
        n_full[y == val] = ball_counts(subset, x, eps - 1e-12, backend)
    weighted_y_counts_mean = np.sum(np.dot(psi(y_counts), y_counts / n))
    value = float(psi(n) + psi(k) - np.mean(psi(n_full)) - weighted_y_counts_mean)
    return (value, {"eps": eps_full, "n_full": n_full}) if detail else value


def _verify_not_continuous(values, n: int) -> None:
    if len(values) > n / 4:
        warnings.warn("A discrete variable has relatively many unique values." +
                      " Have you set marked the discrete variables in correct order?" +
                      " If both X and Y are discrete, the conditioning variable cannot be continuous" +
                      " (this limitation can be lifted in the future).", UserWarning)


def estimate_conditional_semidiscrete_mi(x, y, cond, k: int = 3, backend: str = "brute",
                                         detail: bool = False):
    """Conditional Ross MI (ennemi/_entropy_estimators.py:203-247)."""
    n = len(y)
    cond = np.column_stack((cond,))
    y_values = np.unique(y)
    _verify_not_continuous(y_values, n)
    xz = np.column_stack((x, cond))
    nxz = np.empty(n)
    nyz = np.empty(n)
    nz = np.empty(n)
    eps_full = np.empty(n)
    for val in y_values:
        subset = y == val
        eps = knn_eps(xz[subset], k, backend)
        eps_full[subset] = eps
        r = eps - 1e-12
        nxz[subset] = ball_counts(xz[subset], xz, r, backend)
        nyz[subset] = ball_counts(cond[subset], cond[subset], r, backend)
        nz[subset] = ball_counts(cond[subset], cond, r, backend)
    with np.errstate(invalid="ignore"):
        value = float(psi(k) - np.mean(psi(nxz) + psi(nyz) - psi(nz)))
    return (value, {"eps": eps_full, "nxz": nxz, "nyz": nyz, "nz": nz}) if detail else value


def _assert_not_object(x) -> None:
    if x.dtype.kind == "O":
        raise TypeError("Data type 'object' is not supported." +
                        " Please pass only numeric, boolean, or string data." +
                        " If your data is in a pandas DataFrame, convert string categories" +
                        " to integers (pandas stores strings as objects).")


def estimate_discrete_entropy(x) -> float:
    """Plug-in discrete entropy (ennemi/_entropy_estimators.py:44-56)."""
    n = x.shape[0]
    _assert_not_object(x)
    _, counts = np.unique(x, axis=0, return_counts=True)
    probs = counts / n
    return -np.sum(np.dot(probs, np.log(probs)))


def estimate_discrete_mi(x, y) -> float:
    """Plug-in discrete MI (ennemi/_entropy_estimators.py:257-289)."""
    n = len(x)
    data = np.column_stack((x, y))
    _assert_not_object(data)
    x_vals, x_counts = np.unique(data[:, 0], return_counts=True)
    x_dict = dict(zip(x_vals, x_counts))
    y_vals, y_counts = np.unique(data[:, 1], return_counts=True)
    y_dict = dict(zip(y_vals, y_counts))
    joint_vals, joint_counts = np.unique(data, axis=0, return_counts=True)
    _verify_not_continuous(x_vals, n)
    _verify_not_continuous(y_vals, n)

    def sum_term(a):
        return int(a[2]) * np.log(n * int(a[2]) / (x_dict[a[0]] * y_dict[a[1]]))
    return np.sum(np.apply_along_axis(sum_term, 1, np.column_stack((joint_vals, joint_counts)))) / n


def estimate_conditional_discrete_mi(x, y, cond) -> float:
    """Plug-in conditional discrete MI (ennemi/_entropy_estimators.py:291-320)."""
    n = len(x)
    _assert_not_object(cond)
    cond_vals, cond_inverses, cond_counts = np.unique(cond, axis=0, return_inverse=True,
                                                      return_counts=True)
    cond_inverses = np.ravel(cond_inverses)
    cond_probs = np.zeros(len(cond_vals))
    for i in range(len(cond_vals)):
        cond_probs[i] = cond_counts[i] * estimate_discrete_mi(x[cond_inverses == i],
                                                              y[cond_inverses == i])
    return np.sum(cond_probs).item() / n


# ---------------------------------------------------------------------------
# Driver (ennemi/_driver.py)
# ---------------------------------------------------------------------------

def _normalize(mi: float) -> float:
    """ennemi/_driver.py:77-81."""
    if mi <= 0.0:
        return mi
    return np.sqrt(1 - np.exp(-2 * mi))


def normalize_mi(mi):
    """ennemi/_driver.py:43-75 (NumPy branch)."""
    return np.vectorize(_normalize, otypes=[float])(mi)


def apply_masks(xs, ys, zs, mask, min_lag, max_lag, drop_nan):
    """ennemi/_driver.py:834-854."""
    if mask is not None:
        mask_subset = mask[max_lag: len(mask) + min_lag]
        xs = xs[mask_subset]
        ys = ys[mask_subset]
        if zs is not None:
            zs = zs[mask_subset]
    if drop_nan:
        notna = ~(np.isnan(xs) | np.isnan(ys))
        if zs is not None:
            notna &= ~np.max(np.isnan(zs), axis=1)
            zs = zs[notna]
        xs = xs[notna]
        ys = ys[notna]
    return xs, ys, zs


def validate_masked_data(xs, ys, zs, k, discrete_x, discrete_y) -> None:
    """ennemi/_driver.py:856-869."""
    if len(ys) <= k:
        raise ValueError("k must be smaller than number of observations (after lag and mask)")
    nan_msg = "input contains NaNs (after applying the mask), pass drop_nan=True to ignore"
    if (not discrete_x or xs.dtype.kind in "iufc") and np.isnan(xs).any():
        raise ValueError(nan_msg)
    if (not discrete_y or ys.dtype.kind in "iufc") and np.isnan(ys).any():
        raise ValueError(nan_msg)
    if not (discrete_x and discrete_y) and zs is not None and np.isnan(zs).any():
        raise ValueError(nan_msg)


def rescale_data(xs, ys, zs, discrete_x, discrete_y):
    """ennemi/_driver.py:871-902: standardize, then add 1e-10 noise from a fixed seed."""
    rng = np.random.default_rng(2_718281828)
    if not discrete_x:
        std = xs.std()
        if np.abs(std) < 1e-20:
            warnings.warn(PREPROCESS_CONSTANT_DATA_WARNING)
        else:
            xs = (xs - xs.mean()) / std
            xs += rng.normal(0.0, 1e-10, xs.shape)
    if not discrete_y:
        std = ys.std()
        if np.abs(std) < 1e-20:
            warnings.warn(PREPROCESS_CONSTANT_DATA_WARNING)
        else:
            ys = (ys - ys.mean()) / std
            ys += rng.normal(0.0, 1e-10, ys.shape)
    if zs is not None and not (discrete_x and discrete_y):
        std = zs.std(axis=0)
        if np.any(np.abs(std) < 1e-20):
            warnings.warn(PREPROCESS_CONSTANT_DATA_WARNING)
        else:
            zs = (zs - zs.mean(axis=0)) / std
            zs += rng.normal(0.0, 1e-10, zs.shape)
    return xs, ys, zs


def lagged_task(param_tuple, backend: str = "brute"):
    """One estimation task, ennemi/_driver.py:788-832.

    Returns (value, prepared) where prepared holds the preprocessed arrays
    handed to the estimator, for checking the product's host planner.
    """
    (x, y, lag, max_lag, min_lag, k, mask, cond, cond_lag,
     discrete_x, discrete_y, preprocess, drop_nan) = param_tuple
    min_lag = min(min_lag, 0)
    max_lag = max(max_lag, 0)
    xs = x[max_lag - lag: len(x) - lag + min_lag]
    ys = y[max_lag: len(y) + min_lag]
    if cond is not None:
        zs = np.column_stack([cond[max_lag - cond_lag[i]: len(cond) - cond_lag[i] + min_lag, i]
                              for i in range(len(cond_lag))])
    else:
        zs = None
    xs, ys, zs = apply_masks(xs, ys, zs, mask, min_lag, max_lag, drop_nan)
    validate_masked_data(xs, ys, zs, k, discrete_x, discrete_y)
    if preprocess:
        xs, ys, zs = rescale_data(xs, ys, zs, discrete_x, discrete_y)
    prepared = (xs, ys, zs)
    if zs is None:
        if discrete_x and discrete_y:
            return estimate_discrete_mi(xs, ys), prepared
        elif discrete_x:
            return estimate_semidiscrete_mi(ys, xs, k, backend), prepared
        if discrete_y:
            return estimate_semidiscrete_mi(xs, ys, k, backend), prepared
        return estimate_single_mi(xs, ys, k, backend), prepared
    if discrete_x and discrete_y:
        return estimate_conditional_discrete_mi(xs, ys, zs), prepared
    elif discrete_x:
        return estimate_conditional_semidiscrete_mi(ys, xs, zs, k, backend), prepared
    if discrete_y:
        return estimate_conditional_semidiscrete_mi(xs, ys, zs, k, backend), prepared
    return estimate_conditional_mi(xs, ys, zs, k, backend), prepared


def _map_tasks(params, backend: str, threads: int):
    """Fan-out of ennemi/_driver.py:736-785 (results kept in task order)."""
    if threads <= 1 or len(params) <= 1:
        return [lagged_task(p, backend)[0] for p in params]
    import concurrent.futures
    with concurrent.futures.ThreadPoolExecutor(threads, "oracle-work") as ex:
        return list(ex.map(lambda p: lagged_task(p, backend)[0], params))


def time_estimate(n: int, cond, k: int) -> float:
    """ennemi/_driver.py:726-734."""
    n_cond = 0 if cond is None else cond.shape[1]
    return n ** (1.0 + 0.05 * n_cond) * (0.9 + 0.1 * math.sqrt(k)) * 1e-5


def threads_for(n_tasks: int, n: int, cond, k: int) -> int:
    """The reference's serial/threaded switch (ennemi/_driver.py:745-752)."""
    if n_tasks * time_estimate(n, cond, k) < 0.2:
        return 1
    return os.cpu_count() or 1


def estimate_mi(y, x, lag=0, *, k: int = 3, cond=None, cond_lag=0, mask=None,
                discrete_y: bool = False, discrete_x: bool = False, preprocess: bool = True,
                drop_nan: bool = False, normalize: bool = False, backend: str = "brute",
                threads: Optional[int] = None):
    """ennemi/_driver.py:229-361 and 452-498 (NumPy inputs only)."""
    x = np.asarray(x)
    y = np.asarray(y)
    if cond is not None:
        cond = np.column_stack((np.asarray(cond),))
        ncond = cond.shape[1]
    else:
        ncond = 1
    if mask is not None:
        mask = np.asarray(mask)
    lag = np.atleast_1d(lag)
    cond_lag = np.broadcast_to(np.column_stack((cond_lag,)), (lag.shape[0], ncond))
    min_lag = min(np.min(lag), np.min(cond_lag))
    max_lag = max(np.max(lag), np.max(cond_lag))
    if max_lag - min_lag >= y.size or max_lag >= y.size or min_lag <= -y.size:
        raise ValueError("lag is too large, no observations left")
    nvar = 1 if x.ndim == 1 else x.shape[1]
    indices = list(itertools.product(range(len(lag)), range(nvar)))
    params = [((x if x.ndim == 1 else x[:, v]), y, lag[li], max_lag, min_lag, k, mask, cond,
               cond_lag[li], discrete_x, discrete_y, preprocess, drop_nan)
              for (li, v) in indices]
    if threads is None:
        threads = threads_for(len(params), len(y), cond, k)
    res = _map_tasks(params, backend, threads)
    result = np.empty((len(lag), nvar))
    for index, r in zip(indices, res):
        result[index] = r
    if normalize:
        result = normalize_mi(result)
    return result


def pairwise_mi(data, *, k: int = 3, cond=None, mask=None, discrete=False,
                preprocess: bool = True, drop_nan: bool = False, normalize: bool = False,
                backend: str = "brute", threads: Optional[int] = None):
    """ennemi/_driver.py:540-624 and 680-723 (NumPy inputs only)."""
    data = np.asarray(data)
    if cond is not None:
        cond = np.column_stack((np.asarray(cond),))
    if mask is not None:
        mask = np.asarray(mask)
    if data.ndim == 1 or data.shape[1] == 1:
        return np.full((1, 1), np.nan)
    discrete = np.broadcast_to(discrete, data.shape[1])
    nobs, nvar = data.shape
    cond_lag = np.full(cond.shape[1], 0) if cond is not None else np.asarray(0)
    indices, params = [], []
    for i in range(nvar):
        for j in range(i + 1, nvar):
            indices.append((i, j))
            params.append((data[:, i], data[:, j], 0, 0, 0, k, mask, cond, cond_lag,
                           discrete[i], discrete[j], preprocess, drop_nan))
    if threads is None:
        threads = threads_for(len(params), nobs, cond, k)
    res = _map_tasks(params, backend, threads)
    result = np.full((nvar, nvar), np.nan)
    for (i, j), r in zip(indices, res):
        result[i, j] = r
        result[j, i] = r
    if normalize:
        result = normalize_mi(result)
    return result


def _mask_and_validate_entropy(x, mask, drop_nan, discrete, k):
    """ennemi/_driver.py:180-200."""
    if mask is not None:
        x = x[np.asanyarray(mask)]
    if drop_nan and not discrete:
        if x.ndim > 1:
            x = x[~np.max(np.isnan(x), axis=1)]
        else:
            x = x[~np.isnan(x)]
    if k >= x.shape[0]:
        raise ValueError("k must be smaller than number of observations (after lag and mask)")
    if not discrete and np.any(np.isnan(x)):
        raise ValueError("input contains NaNs (after applying the mask), pass drop_nan=True to ignore")
    return x


def _call_entropy_func(xs, k, discrete, backend):
    """ennemi/_driver.py:222-226."""
    if discrete:
        return estimate_discrete_entropy(xs)
    return estimate_single_entropy(xs, k, backend)


def _estimate_entropy(x, k, multidim, mask, discrete, drop_nan, backend):
    """ennemi/_driver.py:165-178."""
    if multidim or x.ndim == 1:
        x = _mask_and_validate_entropy(x, mask, drop_nan, discrete, k)
        return np.asarray(_call_entropy_func(x, k, discrete, backend))
    nvar = x.shape[1]
    result = np.empty(nvar)
    for i in range(nvar):
        xs = _mask_and_validate_entropy(x[:, i], mask, drop_nan, discrete, k)
        result[i] = _call_entropy_func(xs, k, discrete, backend)
    return result


def estimate_entropy(x, *, k: int = 3, multidim: bool = False, discrete: bool = False,
                     mask=None, cond=None, drop_nan: bool = False, backend: str = "brute"):
    """ennemi/_driver.py:84-162 and 202-220 (NumPy inputs only)."""
    x = np.asarray(x)
    if cond is None:
        return _estimate_entropy(x, k, multidim, mask, discrete, drop_nan, backend)
    cond = np.asarray(cond)
    marginal = _estimate_entropy(cond, k, True, mask, discrete, drop_nan, backend)
    if multidim or x.ndim == 1:
        xs = _mask_and_validate_entropy(np.column_stack((x, cond)), mask, drop_nan, discrete, k)
        return np.asarray(_call_entropy_func(xs, k, discrete, backend) - marginal)
    nvar = x.shape[1]
    joint = np.empty(nvar)
    for i in range(nvar):
        xs = _mask_and_validate_entropy(np.column_stack((x[:, i], cond)), mask, drop_nan,
                                        discrete, k)
        joint[i] = _call_entropy_func(xs, k, discrete, backend)
    return joint - marginal
