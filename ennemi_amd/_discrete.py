"""Plug-in estimators for discrete variables: host fallback.

The plug-in entropy and MI of discrete variables need no neighbour search,
only the number of points per distinct value (or row).  This module computes
them by factorising each variable into dense codes (the position of its value
among the sorted distinct values, so code order is np.unique order) and
counting code combinations with a histogram.  It serves the inputs that the
device counting path does not take (strings and other non-numeric data); the
estimates follow ennemi/_entropy_estimators.py:44-56 (entropy), :257-289
(MI) and :291-320 (conditional MI), and the error and warning texts are the
reference's.
"""

from __future__ import annotations

from warnings import warn

import numpy as np

OBJECT_MESSAGE = ("Data type 'object' is not supported." +
                  " Please pass only numeric, boolean, or string data." +
                  " If your data is in a pandas DataFrame, convert string categories" +
                  " to integers (pandas stores strings as objects).")

MANY_VALUES_MESSAGE = ("A discrete variable has relatively many unique values." +
                       " Have you set marked the discrete variables in correct order?" +
                       " If both X and Y are discrete, the conditioning variable cannot be continuous" +
                       " (this limitation can be lifted in the future).")


def reject_object(a) -> None:
    """TypeError for object arrays (_entropy_estimators.py:58-66)."""
    if a.dtype.kind == "O":
        raise TypeError(OBJECT_MESSAGE)


def verify_not_continuous(values, n: int) -> None:
    """Warn when more than a quarter of the points are distinct values
    (_entropy_estimators.py:249-254)."""
    if len(values) > n / 4:
        warn(MANY_VALUES_MESSAGE, UserWarning)


def factorize(a) -> tuple[np.ndarray, np.ndarray]:
    """(code of every row, points per code) of a 1-D or 2-D array, codes in
    the sorted order of the distinct values or rows."""
    _, inverse, counts = np.unique(a, axis=0, return_inverse=True, return_counts=True)
    return np.ravel(inverse).astype(np.int64), counts.astype(np.int64)


def mi_from_codes(cx, nx, cy, ny, n: int) -> float:
    """sum over occupied (x, y) cells of c log(n c / (n_x n_y)), divided by n.

    The cells are visited in (x code, y code) order, the order of the
    reference's np.unique over the stacked columns, and summed by np.sum."""
    nbin = len(nx) * len(ny)
    joint = np.bincount(cx * len(ny) + cy, minlength=nbin)
    cells = np.flatnonzero(joint)
    c = joint[cells]
    a = nx[cells // len(ny)]
    b = ny[cells % len(ny)]
    return float(np.sum(c * np.log(n * c / (a * b)))) / n


def estimate_discrete_entropy(x) -> float:
    """-sum p log p over the distinct rows of x."""
    reject_object(x)
    _, counts = factorize(x)
    p = counts / x.shape[0]
    return -float(np.dot(p, np.log(p)))


def estimate_discrete_mi(x, y) -> float:
    """Plug-in MI of two discrete variables (both converted to their common
    type first, as np.column_stack does in the reference)."""
    n = len(x)
    pair = np.column_stack((x, y))
    reject_object(pair)
    cx, nx = factorize(pair[:, 0])
    cy, ny = factorize(pair[:, 1])
    verify_not_continuous(nx, n)
    verify_not_continuous(ny, n)
    return mi_from_codes(cx, nx, cy, ny, n)


def estimate_conditional_discrete_mi(x, y, cond) -> float:
    """sum over condition values z of n_z MI(x; y | z), divided by n."""
    n = len(x)
    reject_object(cond)
    cz, nz = factorize(cond)
    # Group the points by condition value, keeping their order in a group.
    order = np.argsort(cz, kind="stable")
    ends = np.cumsum(nz)
    terms = np.zeros(len(nz))
    for g in range(len(nz)):
        rows = order[ends[g] - nz[g]:ends[g]]
        terms[g] = nz[g] * estimate_discrete_mi(x[rows], y[rows])
    return float(np.sum(terms)) / n
