"""Plug-in estimators for discrete variables (host, NumPy).

These have no neighbour search and are not part of the accelerated path:
they count unique values (np.unique), as the reference does in
ennemi/_entropy_estimators.py:44-66 and 249-320.  They complete the
``discrete_x=discrete_y=True`` and ``estimate_entropy(discrete=True)`` cases
of the public API.
"""

from __future__ import annotations

from warnings import warn

import numpy as np


def _assert_not_object(x) -> None:
    if x.dtype.kind == "O":
        raise TypeError("Data type 'object' is not supported." +
                        " Please pass only numeric, boolean, or string data." +
                        " If your data is in a pandas DataFrame, convert string categories" +
                        " to integers (pandas stores strings as objects).")


def verify_not_continuous(values, n: int) -> None:
    """_entropy_estimators.py:249-254."""
    if len(values) > n / 4:
        warn("A discrete variable has relatively many unique values." +
             " Have you set marked the discrete variables in correct order?" +
             " If both X and Y are discrete, the conditioning variable cannot be continuous" +
             " (this limitation can be lifted in the future).", UserWarning)


def estimate_discrete_entropy(x) -> float:
    """-sum p log p over the unique rows of x (_entropy_estimators.py:44-56)."""
    n = x.shape[0]
    _assert_not_object(x)
    _, counts = np.unique(x, axis=0, return_counts=True)
    probs = counts / n
    return -np.sum(np.dot(probs, np.log(probs)))


def estimate_discrete_mi(x, y) -> float:
    """sum p(x,y) log(p(x,y) / (p(x) p(y))) (_entropy_estimators.py:257-289)."""
    n = len(x)
    data = np.column_stack((x, y))
    _assert_not_object(data)
    x_vals, x_counts = np.unique(data[:, 0], return_counts=True)
    x_dict = dict(zip(x_vals, x_counts))
    y_vals, y_counts = np.unique(data[:, 1], return_counts=True)
    y_dict = dict(zip(y_vals, y_counts))
    joint_vals, joint_counts = np.unique(data, axis=0, return_counts=True)
    verify_not_continuous(x_vals, n)
    verify_not_continuous(y_vals, n)

    def sum_term(a):
        joint_weight = int(a[2])
        return joint_weight * np.log(n * joint_weight / (x_dict[a[0]] * y_dict[a[1]]))

    return np.sum(np.apply_along_axis(sum_term, 1, np.column_stack((joint_vals, joint_counts)))) / n


def estimate_conditional_discrete_mi(x, y, cond) -> float:
    """sum_z p(z) MI(x; y | z) (_entropy_estimators.py:291-320)."""
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
