"""Plug-in estimators for discrete variables.

The plug-in entropy and MI of discrete variables need no neighbour search,
only the points per distinct value or row: np.unique(rows, axis=0,
return_inverse=True, return_counts=True), which is where the reference
(ennemi/_entropy_estimators.py:44-56 entropy, :257-289 MI, :291-320
conditional MI) spends its time.  Numeric data is factorised on the GPU
(enm_unique_rows: a sort of each column, run heads, mixed-radix row keys);
the host then evaluates the reference's formulas over the distinct rows, in
the reference's order (lexicographic rows, np.sum), so each estimate is the
reference's bit for bit given the same counts.  Non-numeric values
(strings), which have no float64 representation, are factorised on the
host with np.unique.  The error and warning texts are the reference's.
"""

from __future__ import annotations

from typing import List, Sequence, Tuple
from warnings import warn

import numpy as np

from . import _native as nat

OBJECT_MESSAGE = ("Data type 'object' is not supported." +
                  " Please pass only numeric, boolean, or string data." +
                  " If your data is in a pandas DataFrame, convert string categories" +
                  " to integers (pandas stores strings as objects).")

MANY_VALUES_MESSAGE = ("A discrete variable has relatively many unique values." +
                       " Have you set marked the discrete variables in correct order?" +
                       " If both X and Y are discrete, the conditioning variable cannot be continuous" +
                       " (this limitation can be lifted in the future).")

_EXACT_INT = float(2 ** 53)


def reject_object(a) -> None:
    """TypeError for object arrays (_entropy_estimators.py:58-66)."""
    if a.dtype.kind == "O":
        raise TypeError(OBJECT_MESSAGE)


def verify_not_continuous(values, n: int) -> None:
    """Warn when more than a quarter of the points are distinct values
    (_entropy_estimators.py:249-254)."""
    if len(values) > n / 4:
        warn(MANY_VALUES_MESSAGE, UserWarning)


def _numeric(a: np.ndarray) -> bool:
    """Values that float64 holds exactly, with their order and equality."""
    if a.dtype.kind == "b":
        return True
    if a.dtype.kind in "iu":
        return a.size == 0 or float(np.max(np.abs(a.astype(np.float64)))) < _EXACT_INT
    return a.dtype.kind == "f" and a.dtype.itemsize <= 8


def _columns(a: np.ndarray) -> List[np.ndarray]:
    a = np.asarray(a)
    if a.ndim == 1:
        return [a]
    return [a[:, c] for c in range(a.shape[1])]


def _device_unique(row_sets: Sequence[Sequence[np.ndarray]]) -> List[Tuple[np.ndarray, np.ndarray]]:
    """(code of every row, points per code) of each row set (a list of
    equally long numeric columns), codes in lexicographic row order, from
    enm_unique_rows."""
    import torch
    from ._batch import _stream
    ctx = nat.context()
    dev = torch.device("cuda", ctx.device)
    st = _stream()
    offsets = {}
    parts = []
    total = 0
    for cols in row_sets:
        for c in cols:
            if id(c) not in offsets:
                offsets[id(c)] = total
                parts.append(np.asarray(c, dtype=np.float64))
                total += len(c)
    values = np.concatenate(parts) if parts else np.empty(0)
    n_sets = len(row_sets)
    sets = (nat.EnmRowset * n_sets)()
    out_off = [0]
    for s, cols in enumerate(row_sets):
        sets[s].n = len(cols[0])
        sets[s].ncol = len(cols)
        for j, c in enumerate(cols):
            sets[s].col[j] = offsets[id(c)]
        out_off.append(out_off[-1] + len(cols[0]))
    h_distinct = np.zeros(n_sets, dtype=np.int64)
    with torch.cuda.stream(st):
        d_values = torch.from_numpy(values).to(dev)
        d_codes = torch.empty(max(out_off[-1], 1), dtype=torch.int32, device=dev)
        d_counts = torch.empty(max(out_off[-1], 1), dtype=torch.int64, device=dev)
    with ctx.lock:
        rc = ctx.unique_rows(sets, n_sets, d_values.data_ptr(), d_codes.data_ptr(), d_counts.data_ptr(),
                             h_distinct, st.cuda_stream)
    if rc == nat.E_UNSUPPORTED:
        # Too many value combinations for one float64 key: number the rows
        # a column at a time (the codes of a prefix keep its lexicographic
        # order, and are far fewer than the combinations).
        out = []
        for cols in row_sets:
            codes, counts = _device_unique([cols[:1]])[0]
            for c in cols[1:]:
                codes, counts = _device_unique([[codes.astype(np.float64), c]])[0]
            out.append((codes, counts))
        return out
    st.synchronize()
    codes = d_codes.cpu().numpy()
    counts = d_counts.cpu().numpy()
    return [(codes[out_off[s]:out_off[s + 1]].astype(np.int64),
             counts[out_off[s]:out_off[s] + int(h_distinct[s])].copy()) for s in range(n_sets)]


def _host_unique(cols: Sequence[np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    rows = cols[0] if len(cols) == 1 else np.column_stack(cols)
    _, inverse, counts = np.unique(rows, axis=0, return_inverse=True, return_counts=True)
    return np.ravel(inverse).astype(np.int64), counts.astype(np.int64)


def _unique(row_sets: Sequence[Sequence[np.ndarray]]) -> List[Tuple[np.ndarray, np.ndarray]]:
    if all(_numeric(c) for cols in row_sets for c in cols):
        return _device_unique(row_sets)
    return [_host_unique(cols) for cols in row_sets]


def _mi_terms(n: int, cj: np.ndarray, nj: np.ndarray, cx: np.ndarray, nx: np.ndarray, cy: np.ndarray,
              ny: np.ndarray) -> np.ndarray:
    """c log(n c / (n_x n_y)) per distinct (x, y) row, in row order."""
    xj = np.empty(len(nj), dtype=np.int64)
    yj = np.empty(len(nj), dtype=np.int64)
    xj[cj] = cx
    yj[cj] = cy
    return nj * np.log(n * nj / (nx[xj] * ny[yj]))


def estimate_discrete_entropy(x) -> float:
    """-sum p log p over the distinct rows of x."""
    reject_object(x)
    _, counts = _unique([_columns(x)])[0]
    p = counts / x.shape[0]
    return -float(np.dot(p, np.log(p)))


def estimate_discrete_mi(x, y) -> float:
    """Plug-in MI of two discrete variables (both converted to their common
    type first, as np.column_stack does in the reference)."""
    n = len(x)
    pair = np.column_stack((x, y))
    reject_object(pair)
    xs, ys = pair[:, 0], pair[:, 1]
    (cx, nx), (cy, ny), (cj, nj) = _unique([[xs], [ys], [xs, ys]])
    verify_not_continuous(nx, n)
    verify_not_continuous(ny, n)
    return float(np.sum(_mi_terms(n, cj, nj, cx, nx, cy, ny))) / n


def estimate_conditional_discrete_mi(x, y, cond) -> float:
    """sum over condition rows z of n_z MI(x; y | z), divided by n.  The
    (z, x), (z, y) and (z, x, y) rows are numbered together, so each
    condition row's (x, y) cells are consecutive, in the order the
    reference visits them."""
    n = len(x)
    reject_object(cond)
    zc = _columns(cond)
    pair = np.column_stack((x, y))
    reject_object(pair)
    xs, ys = pair[:, 0], pair[:, 1]
    (cz, nz), (czx, nzx), (czy, nzy), (czxy, nzxy) = _unique([zc, zc + [xs], zc + [ys], zc + [xs, ys]])
    # Condition row of every (z, x), (z, y) and (z, x, y) row.
    g_zx = np.empty(len(nzx), dtype=np.int64)
    g_zx[czx] = cz
    g_zy = np.empty(len(nzy), dtype=np.int64)
    g_zy[czy] = cz
    g_zxy = np.empty(len(nzxy), dtype=np.int64)
    g_zxy[czxy] = cz
    # Distinct x (y) values per condition row: the reference's warning on
    # each subset (_entropy_estimators.py:281-282 via :315).
    for distinct in (np.bincount(g_zx, minlength=len(nz)), np.bincount(g_zy, minlength=len(nz))):
        if np.any(distinct > nz / 4):
            warn(MANY_VALUES_MESSAGE, UserWarning)
    zx_of = np.empty(len(nzxy), dtype=np.int64)
    zx_of[czxy] = czx
    zy_of = np.empty(len(nzxy), dtype=np.int64)
    zy_of[czxy] = czy
    n_sub = nz[g_zxy]
    terms = nzxy * np.log(n_sub * nzxy / (nzx[zx_of] * nzy[zy_of]))
    ends = np.cumsum(np.bincount(g_zxy, minlength=len(nz)))
    weighted = np.zeros(len(nz))
    for g in range(len(nz)):
        a = ends[g - 1] if g else 0
        weighted[g] = nz[g] * (np.sum(terms[a:ends[g]]) / nz[g])
    return float(np.sum(weighted)) / n
