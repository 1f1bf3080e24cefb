"""Public estimation functions with the call surface of ennemi 1.5.0.

Same signatures, argument meaning, validation messages, warnings and result
layout as the reference (ennemi/_driver.py).  What changes is where the work
runs: every (lag, variable) or (i, j) task of a call is validated and
rescaled on the host with the reference's own NumPy operations
(_driver.py:788-902, so the float64 inputs of each estimator are bit for bit
the reference's), then all tasks run as one batch of HIP instances on the GPU
(replacing the thread-pool fan-out of _driver.py:736-785).

Deviations, by design:
* ``max_threads`` is accepted and ignored (there is no host thread pool).
* ``callback`` is called after the batch, in task order (the reference calls
  it as tasks complete; its tests only check the set of calls).
* An invalid task raises at planning time, as the reference's sequential
  path does; the reference's threaded path would leave NaN instead.
"""

from __future__ import annotations

import itertools
import sys
from typing import Callable, List, Optional, Sequence, Tuple, Union
from warnings import warn

import numpy as np

from . import _native as nat
from ._batch import Batch
from . import _discrete

DISCRETE_NORMALIZATION_WARNING = \
    "You have set normalize=True while at least one variable is discrete. " +\
    "The correlation coefficient formula assumes both variables to be continuous, " +\
    "and the results are not comparable across different discrete variables. " +\
    "Compare the raw MI against entropy of the discrete variable instead. " +\
    "If you really want to calculate correlation coefficients, you can suppress " +\
    "this warning by setting normalize=False and calling normalize_mi() on the results."

PREPROCESS_CONSTANT_DATA_WARNING = \
    "A variable not marked as discrete takes only a single value, " +\
    "or values in a very small numerical range. " +\
    "If this is intentional, you can suppress this warning by passing preprocess=False. " +\
    "Note that this disables rescaling on all other variables as well."

K_TOO_LARGE_MSG = "k must be smaller than number of observations (after lag and mask)"
NANS_LEFT_MSG = "input contains NaNs (after applying the mask), pass drop_nan=True to ignore"

# The fixed seed of the reference's rescaling noise (_driver.py:873-874).
NOISE_SEED = 2_718281828


# ---------------------------------------------------------------------------
# Normalisation (_driver.py:43-81)
# ---------------------------------------------------------------------------

def normalize_mi(mi):
    """Normalize mutual information values to the unit interval.

    The result matches the correlation coefficient of two Gaussian variables
    with the same MI, sqrt(1 - exp(-2 MI)); negative values are kept as-is.
    pandas ``DataFrame``/``Series`` inputs keep their columns and index.
    """
    if "pandas" in sys.modules:
        import pandas
        if isinstance(mi, (pandas.DataFrame, pandas.Series)):
            return mi.map(_normalize)
    return np.vectorize(_normalize, otypes=[float])(mi)


def _normalize(mi: float) -> float:
    if mi <= 0.0:
        return mi
    return np.sqrt(1 - np.exp(-2 * mi))


# ---------------------------------------------------------------------------
# Parameter validation (_driver.py:501-537)
# ---------------------------------------------------------------------------

def _check_parameters(x, y, k, cond, mask) -> None:
    _validate_k_type(k)
    if not 1 <= len(x.shape) <= 2:
        raise ValueError("x must be one- or two-dimensional")
    if y is not None:
        if len(y.shape) > 1:
            raise ValueError("y must be one-dimensional")
        if x.shape[0] != y.shape[0]:
            raise ValueError("x and y must have same length")
    if mask is not None:
        _validate_mask(mask, x.shape[0])
    if cond is not None:
        _validate_cond(cond, x.shape[0])


def _validate_k_type(k) -> None:
    if not isinstance(k, int):
        raise TypeError("k must be int")
    if k <= 0:
        raise ValueError("k must be greater than zero")


def _validate_mask(mask, input_len: int) -> None:
    if len(mask.shape) > 1:
        raise ValueError("mask must be one-dimensional")
    if len(mask) != input_len:
        raise ValueError("mask length does not match input length")
    if mask.dtype != bool:
        raise TypeError("mask must contain only booleans")


def _validate_cond(cond, input_len: int) -> None:
    if not 1 <= cond.ndim <= 2:
        raise ValueError("cond must be one- or two-dimensional")
    if input_len != len(cond):
        raise ValueError("x and cond must have same length")


# ---------------------------------------------------------------------------
# Task preparation: lag windows, masks, NaNs, rescaling (_driver.py:788-902)
# ---------------------------------------------------------------------------

class _Noise:
    """The reference's rescaling noise.  Every task draws from a fresh
    default_rng(2_718281828) in the order x, y, z, so all draws of all tasks
    are prefixes of one stream; it is generated once per call and sliced."""

    def __init__(self) -> None:
        self.buf = np.empty(0)

    def take(self, offset: int, shape) -> np.ndarray:
        size = int(np.prod(shape))
        if offset + size > len(self.buf):
            total = max(offset + size, 2 * len(self.buf))
            self.buf = np.random.default_rng(NOISE_SEED).normal(0.0, 1e-10, total)
        return self.buf[offset:offset + size].reshape(shape)


def _rescale(xs, ys, zs, discrete_x: bool, discrete_y: bool, noise: _Noise):
    """_rescale_data (_driver.py:871-902) with the noise stream shared."""
    drawn = 0
    if not discrete_x:
        std = xs.std()
        if np.abs(std) < 1e-20:
            warn(PREPROCESS_CONSTANT_DATA_WARNING)
        else:
            xs = (xs - xs.mean()) / std
            xs += noise.take(drawn, xs.shape)
            drawn += xs.size
    if not discrete_y:
        std = ys.std()
        if np.abs(std) < 1e-20:
            warn(PREPROCESS_CONSTANT_DATA_WARNING)
        else:
            ys = (ys - ys.mean()) / std
            ys += noise.take(drawn, ys.shape)
            drawn += ys.size
    if zs is not None and not (discrete_x and discrete_y):
        std = zs.std(axis=0)
        if np.any(np.abs(std) < 1e-20):
            warn(PREPROCESS_CONSTANT_DATA_WARNING)
        else:
            zs = (zs - zs.mean(axis=0)) / std
            zs += noise.take(drawn, zs.shape)
            drawn += zs.size
    return xs, ys, zs


def _apply_masks(xs, ys, zs, mask, min_lag, max_lag, drop_nan):
    """_driver.py:834-854."""
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


def _validate_masked_data(xs, ys, zs, k, discrete_x, discrete_y) -> None:
    """_driver.py:856-869."""
    if len(ys) <= k:
        raise ValueError(K_TOO_LARGE_MSG)
    if (not discrete_x or xs.dtype.kind in "iufc") and np.isnan(xs).any():
        raise ValueError(NANS_LEFT_MSG)
    if (not discrete_y or ys.dtype.kind in "iufc") and np.isnan(ys).any():
        raise ValueError(NANS_LEFT_MSG)
    if not (discrete_x and discrete_y) and zs is not None and np.isnan(zs).any():
        raise ValueError(NANS_LEFT_MSG)


def _group_by_class(labels, arrays):
    """Points grouped by class in np.unique order (as _estimate_semidiscrete_mi
    iterates over y values, _entropy_estimators.py:177-196)."""
    values, inverse, counts = np.unique(labels, return_inverse=True, return_counts=True)
    order = np.argsort(np.ravel(inverse), kind="stable")
    return values, counts, [np.ascontiguousarray(a[order]) for a in arrays]


class _Tasks:
    """Collects tasks of one call: GPU instances in a Batch, discrete-only
    tasks evaluated on the host, and where each result goes."""

    def __init__(self) -> None:
        self.batch = Batch()
        self.slots: List[Tuple[str, int]] = []   # ("gpu", inst) or ("host", value index)
        self.host_values: List[float] = []
        # Work left for after the batch's launch (warnings that read device
        # results), so that it does not hold the launch back.
        self.deferred: List[Callable[[], None]] = []

    def finish(self) -> None:
        for f in self.deferred:
            f()
        self.deferred = []

    def host(self, value: float) -> None:
        self.slots.append(("host", len(self.host_values)))
        self.host_values.append(float(value))

    def gpu(self, idx: int) -> None:
        self.slots.append(("gpu", idx))

    def run(self) -> List[float]:
        gpu_vals = self.batch.run() if len(self.batch) else np.empty(0)
        self.finish()
        return [float(gpu_vals[i]) if kind == "gpu" else self.host_values[i] for kind, i in self.slots]


def _add_mi_task(tasks: _Tasks, xs, ys, zs, k: int, discrete_x: bool, discrete_y: bool) -> None:
    """Dispatch of _lagged_mi (_driver.py:815-832) onto instance kinds."""
    b = tasks.batch
    if zs is None:
        if discrete_x and discrete_y:
            tasks.host(_discrete.estimate_discrete_mi(xs, ys))
            return
        if discrete_x or discrete_y:
            cont, disc = (ys, xs) if discrete_x else (xs, ys)
            cont = np.column_stack((cont,))
            values, counts, grouped = _group_by_class(disc, [cont[:, c] for c in range(cont.shape[1])])
            if len(values) > len(disc) / 4:
                warn("The discrete variable has relatively many unique values." +
                     " Did you pass y and x in correct order?", UserWarning)
            tasks.gpu(b.add(nat.KIND_ROSS, k, grouped, class_sizes=counts))
            return
        tasks.gpu(b.add(nat.KIND_KSG, k, _cols(xs), _cols(ys)))
        return
    if discrete_x and discrete_y:
        tasks.host(_discrete.estimate_conditional_discrete_mi(xs, ys, zs))
        return
    if discrete_x or discrete_y:
        cont, disc = (ys, xs) if discrete_x else (xs, ys)
        cont = np.column_stack((cont,))
        zz = np.column_stack((zs,))
        values, counts, grouped = _group_by_class(
            disc, [cont[:, c] for c in range(cont.shape[1])] + [zz[:, c] for c in range(zz.shape[1])])
        _discrete.verify_not_continuous(values, len(disc))
        dx = cont.shape[1]
        tasks.gpu(b.add(nat.KIND_CROSS, k, grouped[:dx], z=grouped[dx:], class_sizes=counts))
        return
    tasks.gpu(b.add(nat.KIND_CMI, k, _cols(xs), _cols(ys), _cols(zs)))


def _cols(a) -> List[np.ndarray]:
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        return [a]
    return [np.ascontiguousarray(a[:, c]) for c in range(a.shape[1])]


def _prepare_lagged(param_tuple, noise: _Noise):
    """_lagged_mi up to the estimator call (_driver.py:788-812)."""
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
    xs, ys, zs = _apply_masks(xs, ys, zs, mask, min_lag, max_lag, drop_nan)
    _validate_masked_data(xs, ys, zs, k, discrete_x, discrete_y)
    if preprocess:
        xs, ys, zs = _rescale(xs, ys, zs, discrete_x, discrete_y, noise)
    return xs, ys, zs


# ---------------------------------------------------------------------------
# estimate_mi / estimate_corr (_driver.py:229-498)
# ---------------------------------------------------------------------------

def estimate_mi(y, x, lag=0, *, k: int = 3, cond=None, cond_lag=0, mask=None,
                discrete_y: bool = False, discrete_x: bool = False, preprocess: bool = True,
                drop_nan: bool = False, normalize: bool = False, max_threads: Optional[int] = None,
                callback: Optional[Callable[[int, int], None]] = None):
    """Estimate the mutual information between y and each x variable.

    - Unconditional MI: the default (KSG, Kraskov et al. 2004).
    - Conditional MI: pass a ``cond`` array (Frenzel & Pompe 2007).
    - Discrete-continuous MI: set ``discrete_y`` or ``discrete_x`` (Ross 2014).

    Returns MI in nats (or correlation scale with ``normalize``) as a 2-D array
    indexed [lag, x column], or a pandas DataFrame for pandas ``x``.  The time
    lag is interpreted as ``y(t) ~ x(t - lag) | z(t - cond_lag)``.  Parameters
    are those of ``ennemi.estimate_mi``; ``max_threads`` is ignored.
    """
    x_arr = np.asarray(x)
    y_arr = np.asarray(y)
    if cond is not None:
        cond_arr = np.column_stack((np.asarray(cond),))
        ncond = cond_arr.shape[1]
    else:
        cond_arr = None
        ncond = 1
    mask_arr = None
    if mask is not None:
        mask_arr = np.asarray(mask)
    lag_arr = np.atleast_1d(lag)
    cond_lag_arr = np.broadcast_to(np.column_stack((cond_lag,)), (lag_arr.shape[0], ncond))

    result = _estimate_mi(y_arr, x_arr, lag_arr, k, cond_arr, cond_lag_arr, mask_arr,
                          discrete_x, discrete_y, preprocess, drop_nan, callback)

    if normalize:
        if discrete_x or discrete_y:
            warn(DISCRETE_NORMALIZATION_WARNING)
        result = normalize_mi(result)

    if "pandas" in sys.modules:
        import pandas
        if isinstance(x, pandas.DataFrame):
            return pandas.DataFrame(result, index=lag_arr, columns=x.columns)
        elif isinstance(x, pandas.Series):
            return pandas.DataFrame(result, index=lag_arr, columns=[x.name])
    return result


def estimate_corr(y, x, lag=0, *, k: int = 3, cond=None, cond_lag=0, mask=None,
                  preprocess: bool = True, drop_nan: bool = False,
                  max_threads: Optional[int] = None,
                  callback: Optional[Callable[[int, int], None]] = None):
    """``estimate_mi`` with ``normalize=True`` (_driver.py:363-449)."""
    return estimate_mi(y, x, lag, k=k, cond=cond, cond_lag=cond_lag, mask=mask,
                       preprocess=preprocess, drop_nan=drop_nan, normalize=True,
                       max_threads=max_threads, callback=callback)


def _estimate_mi(y, x, lag, k, cond, cond_lag, mask, discrete_x, discrete_y, preprocess,
                 drop_nan, callback):
    tasks, indices, nvar = _plan_mi(y, x, lag, k, cond, cond_lag, mask, discrete_x, discrete_y,
                                    preprocess, drop_nan)
    values = tasks.run()
    result = np.empty((len(lag), nvar))
    for index, res in zip(indices, values):
        result[index] = res
    if callback is not None:
        for (li, v) in indices:
            callback(v, lag[li])
    return result


def plan_mi(y, x, lag=0, *, k: int = 3, cond=None, cond_lag=0, mask=None,
            discrete_y: bool = False, discrete_x: bool = False, preprocess: bool = True,
            drop_nan: bool = False, prep: str = "auto"):
    """Validate and rescale every task of an ``estimate_mi`` call and return
    its batch of GPU instances without running it (for timing the device
    work alone).  Returns (batch, task indices [(lag index, x column)]),
    the indices in the batch's instance order.
    prep="host" prepares the tasks with NumPy on the host even where the
    device could (for comparing the two)."""
    x_arr = np.asarray(x)
    y_arr = np.asarray(y)
    cond_arr, ncond = (np.column_stack((np.asarray(cond),)), None) if cond is not None else (None, 1)
    if cond_arr is not None:
        ncond = cond_arr.shape[1]
    lag_arr = np.atleast_1d(lag)
    cond_lag_arr = np.broadcast_to(np.column_stack((cond_lag,)), (lag_arr.shape[0], ncond))
    mask_arr = np.asarray(mask) if mask is not None else None
    tasks, indices, _ = _plan_mi(y_arr, x_arr, lag_arr, k, cond_arr, cond_lag_arr, mask_arr,
                                 discrete_x, discrete_y, preprocess, drop_nan, prep)
    tasks.finish()
    if tasks.host_values:
        raise ValueError("plan_mi covers the GPU estimators only (not discrete-discrete tasks)")
    in_batch = [None] * len(indices)
    for index, (_, inst) in zip(indices, tasks.slots):
        in_batch[inst] = index
    return tasks.batch, in_batch


def _plan_mi(y, x, lag, k, cond, cond_lag, mask, discrete_x, discrete_y, preprocess, drop_nan,
             prep: str = "auto"):
    _check_parameters(x, y, k, cond, mask)
    min_lag = min(np.min(lag), np.min(cond_lag))
    max_lag = max(np.max(lag), np.max(cond_lag))
    if max_lag - min_lag >= y.size or max_lag >= y.size or min_lag <= -y.size:
        raise ValueError("lag is too large, no observations left")
    nvar = 1 if x.ndim == 1 else x.shape[1]

    indices = list(itertools.product(range(len(lag)), range(nvar)))
    if prep != "host" and _device_prep_eligible(x, y, cond, discrete_x, discrete_y):
        return _plan_mi_device(y, x, lag, k, cond, cond_lag, mask, preprocess, drop_nan,
                               min_lag, max_lag, indices), indices, nvar
    tasks = _Tasks()
    noise = _Noise()
    for (li, v) in indices:
        xv = x if x.ndim == 1 else x[:, v]
        params = (xv, y, lag[li], max_lag, min_lag, k, mask, cond, cond_lag[li],
                  discrete_x, discrete_y, preprocess, drop_nan)
        xs, ys, zs = _prepare_lagged(params, noise)
        _add_mi_task(tasks, xs, ys, zs, k, discrete_x, discrete_y)
    return tasks, indices, nvar


def _device_prep_eligible(x, y, cond, discrete_x: bool, discrete_y: bool) -> bool:
    """Tasks whose windows, masks, NaN drop and rescaling run on the device
    (enm_prep_count / enm_prep_write): continuous float64 (or integer or
    boolean, which NumPy computes with in float64) x, y and condition."""
    if discrete_x or discrete_y:
        return False
    for a in (x, y) if cond is None else (x, y, cond):
        if a.dtype.kind not in "fiub" or (a.dtype.kind == "f" and a.dtype != np.float64):
            return False
    if cond is not None and cond.shape[1] > nat.PREP_MAXZ:
        return False
    # Without a usable GPU the host prepares (and validates) the tasks; the
    # estimates themselves then fail with NativeUnavailable.
    try:
        nat.context()
    except nat.NativeUnavailable:
        return False
    return True


_noise_cache = {}


def _device_noise(n: int, dev):
    """The first n draws of default_rng(NOISE_SEED).normal(0, 1e-10, .) in
    HBM: every task draws a prefix of this one stream, so the longest one
    drawn so far serves every shorter call (kept per device)."""
    import torch
    have = _noise_cache.get(str(dev))
    if have is None or have.numel() < n:
        have = torch.from_numpy(np.random.default_rng(NOISE_SEED).normal(0.0, 1e-10, n)).to(dev)
        _noise_cache[str(dev)] = have
    return have


class DevicePrep:
    """The device preparation of one estimate_mi call's continuous tasks
    (windows, mask and NaN drop, rescaling with the reference's noise, and
    the sort hints of the prepared columns): the raw series live in HBM, and
    count() then write() rebuild the batch arena from them
    (enm_prep_count, enm_prep_write, enm_prep_hints).  Tasks are kept in
    device order (by x column, then lag); place[t] is task t's slot."""

    def __init__(self, y, x, lag, cond, cond_lag, mask, preprocess, drop_nan, min_lag, max_lag, indices,
                 hints: bool = True):
        import torch
        from ._batch import _stream
        self.ctx = nat.context()
        self.dev = torch.device("cuda", self.ctx.device)
        self.st = _stream()
        n_rows = len(y)
        nvar = 1 if x.ndim == 1 else x.shape[1]
        self.dz = dz = 0 if cond is None else cond.shape[1]
        series = [y] + ([x] if x.ndim == 1 else [x[:, v] for v in range(nvar)]) + \
            ([] if cond is None else [cond[:, c] for c in range(dz)])
        # The series go up from pinned host memory, asynchronously on the
        # preparation stream (ahead of the kernels that read them), while
        # the host builds the task table.
        self.raw_host = torch.empty(n_rows * len(series), dtype=torch.float64, pin_memory=True)
        raw = self.raw_host.numpy()
        for i, col in enumerate(series):
            raw[i * n_rows:(i + 1) * n_rows] = col
        with torch.cuda.stream(self.st):
            self.d_raw = self.raw_host.to(self.dev, non_blocking=True)
            self.d_mask = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.uint8)).to(self.dev) \
                if mask is not None else None
        min_lag = min(min_lag, 0)
        max_lag = max(max_lag, 0)
        w = n_rows - max_lag + min_lag
        self.n_tasks = n_tasks = len(indices)
        # Device order: by x column, then lag, so a kernel's consecutive
        # blocks (one XCD's share) read one x series.
        idx = np.asarray(indices, dtype=np.int64).reshape(n_tasks, 2)
        order = np.lexsort((idx[:, 0], idx[:, 1]))
        self.place = np.empty(n_tasks, dtype=np.int64)   # task -> its device slot
        self.place[order] = np.arange(n_tasks)
        self.tasks = (nat.EnmPrepTask * n_tasks)()
        li, v = idx[order, 0], idx[order, 1]
        tab = np.ctypeslib.as_array(self.tasks)   # the same memory, field by field
        tab["x_col"] = (1 + v) * n_rows
        tab["y_col"] = 0
        tab["x0"] = max_lag - np.asarray(lag, dtype=np.int64)[li]
        tab["y0"] = max_lag
        tab["w"] = w
        tab["dz"] = dz
        if dz:
            cl = np.asarray(cond_lag, dtype=np.int64).reshape(len(lag), -1)
            tab["z_col"][:, :dz] = (1 + nvar + np.arange(dz)) * n_rows
            tab["z0"][:, :dz] = max_lag - cl[li, :dz]
        self.preprocess = preprocess
        self.hints = hints
        self.flags = (nat.P_DROP_NAN if drop_nan else 0) | (nat.P_RESCALE if preprocess else 0) | \
            (nat.P_HINTS if hints else 0)
        self.h_n = np.zeros(n_tasks, dtype=np.int64)
        self.h_nan = np.zeros(n_tasks, dtype=np.uint32)
        self.arena = self.hint = self.d_const = self.d_noise = None
        self.noise_len = 0

    def _ptrs(self):
        return self.d_raw.data_ptr(), self.d_mask.data_ptr() if self.d_mask is not None else 0

    def count(self, k: int) -> None:
        """Kept rows of every task; validation in task order with the
        reference's messages."""
        raw, mask = self._ptrs()
        with self.ctx.lock:
            self.ctx.prep_count(self.tasks, self.n_tasks, raw, mask, self.flags & ~nat.P_HINTS,
                                self.h_n, self.h_nan, self.st.cuda_stream)
        small = self.h_n[self.place] <= k
        nan = self.h_nan[self.place] != 0
        failing = np.flatnonzero(small | nan)
        if len(failing):
            # The first failing task in task order, its first failed check.
            raise ValueError(K_TOO_LARGE_MSG if small[failing[0]] else NANS_LEFT_MSG)

    def write(self) -> None:
        """The arena (and hints) from the counts of the last count(); the
        buffers are kept while the counts stay the same."""
        import torch
        per = self.h_n * (2 + self.dz)
        self.arena_off = np.concatenate([[0], np.cumsum(per)[:-1]]).astype(np.int64)
        total = int(per.sum())
        noise_len = int(per.max()) if self.preprocess else 0
        with torch.cuda.stream(self.st):
            if self.arena is None or self.arena.numel() != total:
                self.arena = torch.empty(total, dtype=torch.float64, device=self.dev)
                self.hint = torch.empty(total, dtype=torch.int32, device=self.dev) if self.hints else None
                self.d_const = torch.zeros(self.n_tasks, dtype=torch.int32, device=self.dev)
            if self.preprocess and noise_len > self.noise_len:
                self.d_noise = _device_noise(noise_len, self.dev)
                self.noise_len = noise_len
        raw, mask = self._ptrs()
        noise = self.d_noise.data_ptr() if self.preprocess else 0
        nlen = self.noise_len if self.preprocess else 0
        with self.ctx.lock:
            self.ctx.prep_write(self.tasks, self.n_tasks, raw, mask, noise, nlen, self.flags, self.arena_off,
                                self.arena.data_ptr(), self.d_const.data_ptr(), self.st.cuda_stream)
            if self.hints:
                self.ctx.prep_hints(self.tasks, self.n_tasks, raw, mask, noise, nlen, self.flags, self.arena_off,
                                    self.arena.data_ptr(), self.hint.data_ptr(), self.st.cuda_stream)

    def constant_flags(self) -> np.ndarray:
        """Per task (task order): bit 0/1/2 when x/y/the condition was
        constant (left unscaled).  Synchronises the preparation stream."""
        self.st.synchronize()
        if not self.preprocess:
            return np.zeros(self.n_tasks, dtype=np.int32)
        return self.d_const.cpu().numpy()[self.place]


def _plan_mi_device(y, x, lag, k, cond, cond_lag, mask, preprocess, drop_nan, min_lag, max_lag,
                    indices):
    """_plan_mi with each task prepared on the device: the raw series go up
    once, the kept rows of every task are counted there (the host validates
    with those counts, in task order, with the reference's messages), and
    the compacted, rescaled columns are written straight into the batch's
    arena, with the sort hints of every column."""
    from ._batch import ArenaBatch
    prep = DevicePrep(y, x, lag, cond, cond_lag, mask, preprocess, drop_nan, min_lag, max_lag, indices)
    prep.count(k)
    prep.write()
    dz = prep.dz

    def constant_warnings():
        # Read after the batch is launched (the flags are ready by then):
        # reading them first would wait for the preparation before the
        # batch could be planned and launched.
        const = prep.constant_flags()
        for t in range(prep.n_tasks):
            for bit in (1, 2, 4):
                if const[t] & bit and (bit != 4 or dz > 0):
                    warn(PREPROCESS_CONSTANT_DATA_WARNING)

    tasks_out = _Tasks()
    tasks_out.deferred.append(constant_warnings)
    tasks_out.batch = ArenaBatch(prep.arena, prep.hint)
    tasks_out.batch.prep = prep
    kind = nat.KIND_KSG if dz == 0 else nat.KIND_CMI
    offs = (prep.arena_off[:, None] + prep.h_n[:, None] * np.arange(2 + dz)).tolist()
    for n, o in zip(prep.h_n.tolist(), offs):
        tasks_out.batch.add(kind, k, n, o, 1, 1, dz)
    tasks_out.slots = [("gpu", p) for p in prep.place.tolist()]
    return tasks_out


# ---------------------------------------------------------------------------
# pairwise_mi / pairwise_corr (_driver.py:540-723)
# ---------------------------------------------------------------------------

def pairwise_mi(data, *, k: int = 3, cond=None, mask=None, discrete=False,
                preprocess: bool = True, drop_nan: bool = False, normalize: bool = False,
                max_threads: Optional[int] = None,
                callback: Optional[Callable[[int, int], None]] = None):
    """Estimate the pairwise MI between each pair of columns of ``data``.

    Returns a symmetric matrix with NaN on the diagonal (a DataFrame for
    pandas input).  Parameters are those of ``ennemi.pairwise_mi``;
    ``max_threads`` is ignored.
    """
    data_arr = np.asarray(data)
    cond_arr = None
    mask_arr = None
    if cond is not None:
        cond_arr = np.column_stack((np.asarray(cond),))
    if mask is not None:
        mask_arr = np.asarray(mask)
    if data_arr.ndim == 1 or data_arr.shape[1] == 1:
        return np.full((1, 1), np.nan)
    discrete_arr = np.broadcast_to(discrete, data_arr.shape[1])

    result = _pairwise_mi(data_arr, k, cond_arr, preprocess, drop_nan, mask_arr, discrete_arr, callback)

    if normalize:
        if discrete_arr.any():
            warn(DISCRETE_NORMALIZATION_WARNING)
        result = normalize_mi(result)

    if "pandas" in sys.modules:
        import pandas
        if isinstance(data, pandas.DataFrame):
            return pandas.DataFrame(result, index=data.columns, columns=data.columns)
    return result


def pairwise_corr(data, *, k: int = 3, cond=None, mask=None, preprocess: bool = True,
                  drop_nan: bool = False, max_threads: Optional[int] = None,
                  callback: Optional[Callable[[int, int], None]] = None):
    """``pairwise_mi`` with ``normalize=True`` (_driver.py:627-677)."""
    return pairwise_mi(data, k=k, cond=cond, mask=mask, preprocess=preprocess,
                       drop_nan=drop_nan, normalize=True, max_threads=max_threads,
                       callback=callback)


def _pairwise_mi(data, k, cond, preprocess, drop_nan, mask, discrete, callback):
    tasks, indices, nvar = _plan_pairwise(data, k, cond, preprocess, drop_nan, mask, discrete)
    values = tasks.run()
    result = np.full((nvar, nvar), np.nan)
    for (i, j), res in zip(indices, values):
        result[i, j] = res
        result[j, i] = res
    if callback is not None:
        for (i, j) in indices:
            callback(i, j)
    return result


def plan_pairwise_mi(data, *, k: int = 3, cond=None, mask=None, discrete=False,
                     preprocess: bool = True, drop_nan: bool = False):
    """Validate and rescale every pair of a ``pairwise_mi`` call and return
    its batch of GPU instances without running it.  Returns (batch, pairs)."""
    data_arr = np.asarray(data)
    cond_arr = np.column_stack((np.asarray(cond),)) if cond is not None else None
    mask_arr = np.asarray(mask) if mask is not None else None
    discrete_arr = np.broadcast_to(discrete, data_arr.shape[1])
    tasks, indices, _ = _plan_pairwise(data_arr, k, cond_arr, preprocess, drop_nan, mask_arr,
                                       discrete_arr)
    if tasks.host_values:
        raise ValueError("plan_pairwise_mi covers the GPU estimators only")
    return tasks.batch, indices


def _plan_pairwise(data, k, cond, preprocess, drop_nan, mask, discrete):
    _check_parameters(data, None, k, cond, mask)
    if (not (np.all(discrete) or np.all(~discrete))) and (cond is not None):
        raise ValueError("Conditioning is not supported with mixed discrete and continuous data. " +
                         "This is a limitation that can be lifted in the future (see " +
                         "https://github.com/polsys/ennemi/issues/87).")
    nobs, nvar = data.shape
    cond_lag = np.full(cond.shape[1], 0) if cond is not None else np.asarray(0)

    indices = [(i, j) for i in range(nvar) for j in range(i + 1, nvar)]
    tasks = _Tasks()
    noise = _Noise()
    # Without a mask, a NaN drop or a condition, a column's rescaled x-role
    # (or y-role, for a given noise offset) is the same in every pair, so it
    # is prepared once and shared; sharing the object also shares its sort.
    shareable = mask is None and cond is None and not drop_nan and preprocess
    cache = {}

    def role(col: int, is_x: bool, disc: bool, offset: int):
        key = (col, is_x, offset)
        if key not in cache:
            v = data[:, col]
            warned = False
            if not disc:
                std = v.std()
                if np.abs(std) < 1e-20:
                    warned = True
                else:
                    v = (v - v.mean()) / std
                    v += noise.take(offset, v.shape)
            cache[key] = (v, warned)
        v, warned = cache[key]
        if warned:
            warn(PREPROCESS_CONSTANT_DATA_WARNING)
        return v

    for (i, j) in indices:
        dx, dy = bool(discrete[i]), bool(discrete[j])
        if shareable:
            xs0, ys0 = data[:, i], data[:, j]
            _validate_masked_data(xs0, ys0, None, k, dx, dy)
            xs = role(i, True, dx, 0)
            x_drew = (not dx) and not (np.abs(xs0.std()) < 1e-20)
            ys = role(j, False, dy, len(xs0) if x_drew else 0)
        else:
            params = (data[:, i], data[:, j], 0, 0, 0, k, mask, cond, cond_lag, dx, dy,
                      preprocess, drop_nan)
            xs, ys, zs = _prepare_lagged(params, noise)
            _add_mi_task(tasks, xs, ys, zs, k, dx, dy)
            continue
        _add_mi_task(tasks, xs, ys, None, k, dx, dy)
    return tasks, indices, nvar


# ---------------------------------------------------------------------------
# estimate_entropy (_driver.py:84-226)
# ---------------------------------------------------------------------------

def estimate_entropy(x, *, k: int = 3, multidim: bool = False, discrete: bool = False,
                     mask=None, cond=None, drop_nan: bool = False):
    """Estimate the entropy of one or more continuous random variables
    (Kozachenko-Leonenko via the k-th neighbour radius, in nats).

    Each column is a separate variable unless ``multidim``; ``cond`` gives
    the conditional entropy by the chain rule H(X|Y) = H(X,Y) - H(Y).
    Parameters are those of ``ennemi.estimate_entropy``.
    """
    x_arr = np.asarray(x)
    if mask is not None:
        mask = np.asarray(mask)
        _validate_mask(mask, x_arr.shape[0])
    if not 1 <= x_arr.ndim <= 2:
        raise ValueError("x must be one- or two-dimensional")
    _validate_k_type(k)

    if cond is None:
        result = _estimate_entropy(x_arr, k, multidim, mask, discrete, drop_nan)
    else:
        cond_arr = np.asarray(cond)
        _validate_cond(cond_arr, x_arr.shape[0])
        result = _estimate_conditional_entropy(x_arr, cond_arr, k, multidim, mask, discrete, drop_nan)

    if not multidim and "pandas" in sys.modules:
        import pandas
        if isinstance(x, pandas.DataFrame):
            return pandas.DataFrame(np.atleast_2d(result), columns=x.columns)
        elif isinstance(x, pandas.Series):
            return pandas.DataFrame(np.atleast_2d(result), columns=[x.name])
    return result


def _mask_and_validate_entropy(x, mask, drop_nan, discrete, k):
    """_driver.py:180-200."""
    if mask is not None:
        x = x[np.asanyarray(mask)]
    if drop_nan and not discrete:
        if x.ndim > 1:
            x = x[~np.max(np.isnan(x), axis=1)]
        else:
            x = x[~np.isnan(x)]
    if k >= x.shape[0]:
        raise ValueError(K_TOO_LARGE_MSG)
    if not discrete and np.any(np.isnan(x)):
        raise ValueError(NANS_LEFT_MSG)
    return x


def _entropy_task(tasks: _Tasks, xs, k: int, discrete: bool) -> None:
    """_call_entropy_func (_driver.py:222-226)."""
    if discrete:
        tasks.host(_discrete.estimate_discrete_entropy(xs))
    else:
        tasks.gpu(tasks.batch.add(nat.KIND_KL, k, _cols(xs)))


def _estimate_entropy(x, k, multidim, mask, discrete, drop_nan):
    """_driver.py:165-178 (columns batched into one call)."""
    tasks = _Tasks()
    if multidim or x.ndim == 1:
        xs = _mask_and_validate_entropy(x, mask, drop_nan, discrete, k)
        _entropy_task(tasks, xs, k, discrete)
        return np.asarray(tasks.run()[0])
    nvar = x.shape[1]
    for i in range(nvar):
        xs = _mask_and_validate_entropy(x[:, i], mask, drop_nan, discrete, k)
        _entropy_task(tasks, xs, k, discrete)
    return np.asarray(tasks.run(), dtype=np.float64)


def _estimate_conditional_entropy(x, cond, k, multidim, mask, discrete, drop_nan):
    """_driver.py:202-220: H(X|Y) = H(X,Y) - H(Y), all terms in one batch."""
    tasks = _Tasks()
    cm = _mask_and_validate_entropy(cond, mask, drop_nan, discrete, k)
    _entropy_task(tasks, cm, k, discrete)
    if multidim or x.ndim == 1:
        xs = _mask_and_validate_entropy(np.column_stack((x, cond)), mask, drop_nan, discrete, k)
        _entropy_task(tasks, xs, k, discrete)
        vals = tasks.run()
        return np.asarray(vals[1] - vals[0])
    nvar = x.shape[1]
    for i in range(nvar):
        xs = _mask_and_validate_entropy(np.column_stack((x[:, i], cond)), mask, drop_nan, discrete, k)
        _entropy_task(tasks, xs, k, discrete)
    vals = tasks.run()
    joint = np.asarray(vals[1:], dtype=np.float64)
    return joint - vals[0]
