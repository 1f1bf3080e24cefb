"""Estimator-level entry points, each one GPU instance.

These mirror the private functions the reference's unit tests import
(ennemi/_entropy_estimators.py:21-247): the arrays are used as given (no
rescaling), x may be multi-dimensional.  With ``detail=True`` they also return
the per-point radii and ball counts, in the caller's row order.
"""

from __future__ import annotations

from warnings import warn

import numpy as np

from . import _native as nat
from ._batch import Batch
from ._discrete import verify_not_continuous


def _cols(a):
    a = np.column_stack((np.asarray(a, dtype=np.float64),))
    return [np.ascontiguousarray(a[:, c]) for c in range(a.shape[1])]


def _run(batch: Batch, detail: bool, names, unperm=None):
    if not detail:
        return float(batch.run()[0])
    vals, eps, cnts = batch.run(export=True)
    e, c = eps[0], cnts[0]
    if unperm is not None:
        inv = np.empty_like(unperm)
        inv[unperm] = np.arange(len(unperm))
        e = e[inv]
        c = c[:, inv]
    out = {"eps": e}
    for s, name in enumerate(names):
        out[name] = c[s].astype(np.int64)
    return float(vals[0]), out


def estimate_single_entropy(x, k: int = 3, detail: bool = False):
    """Kozachenko-Leonenko entropy of x (n,) or (n, d) (:21-42)."""
    b = Batch()
    b.add(nat.KIND_KL, k, _cols(x))
    return _run(b, detail, [])


def estimate_single_mi(x, y, k: int = 3, detail: bool = False):
    """KSG mutual information (:69-113)."""
    b = Batch()
    b.add(nat.KIND_KSG, k, _cols(x), _cols(y))
    return _run(b, detail, ["nx", "ny"])


def estimate_conditional_mi(x, y, cond, k: int = 3, detail: bool = False):
    """Frenzel-Pompe conditional MI (:116-156)."""
    b = Batch()
    b.add(nat.KIND_CMI, k, _cols(x), _cols(y), _cols(cond))
    return _run(b, detail, ["nxz", "nyz", "nz"])


def _grouped(y, arrays):
    values, inverse, counts = np.unique(y, return_inverse=True, return_counts=True)
    order = np.argsort(np.ravel(inverse), kind="stable")
    return values, counts, order, [np.ascontiguousarray(a[order]) for a in arrays]


def estimate_semidiscrete_mi(x, y, k: int = 3, detail: bool = False):
    """Ross MI between continuous x (n,) or (n, d) and discrete y (:159-200)."""
    y = np.asarray(y)
    cols = _cols(x)
    values, counts, order, grouped = _grouped(y, cols)
    if len(values) > len(cols[0]) / 4:
        warn("The discrete variable has relatively many unique values." +
             " Did you pass y and x in correct order?", UserWarning)
    b = Batch()
    b.add(nat.KIND_ROSS, k, grouped, class_sizes=counts)
    return _run(b, detail, ["n_full"], order)


def estimate_conditional_semidiscrete_mi(x, y, cond, k: int = 3, detail: bool = False):
    """Conditional Ross MI (:203-247)."""
    y = np.asarray(y)
    xc, zc = _cols(x), _cols(cond)
    values, counts, order, grouped = _grouped(y, xc + zc)
    verify_not_continuous(values, len(y))
    b = Batch()
    b.add(nat.KIND_CROSS, k, grouped[:len(xc)], z=grouped[len(xc):], class_sizes=counts)
    return _run(b, detail, ["nxz", "nyz", "nz"], order)


def psi(x):
    """The reference's digamma series (:327-350), host copy for callers that
    need it outside a batch; the estimators evaluate it on the GPU."""
    x = np.asarray(x)
    if np.any(x == 0):
        return np.asarray(np.inf)
    mask = (x != 1)
    result = np.full(x.shape, -0.5772156649015331)
    y = np.asarray(x[mask], dtype=np.float64)
    result[mask] = np.log(y) - np.power(y, -6) * (
        np.power(y, 2) * (np.power(y, 2) * (y / 2 + 1 / 12) - 1 / 120) + 1 / 252)
    return result
