"""Sort hints of device-prepared batches (enm_prep_hints, enm_run_batch_hinted).

A hint is a column's ascending order derived from one sort per raw series;
the batch checks every hint in full and sorts the columns whose hint is
absent or wrong.  Pinned here:
  * every hint that prep_hints gives is exactly np.lexsort's order of the
    prepared column (value, then position), and columns are left without a
    hint only where ties in the raw data make long runs;
  * a hinted batch gives the unhinted batch's bits (values, radii, counts),
    for lag scans with NaNs, masks, conditions, ties and preprocess=False;
  * corrupted hints (swapped, reversed, out of range, duplicated, marked
    absent) change nothing.
"""

import numpy as np
import pytest

import tests.workloads as wl

pytestmark = pytest.mark.gpu


def _plan(*args, **kw):
    from ennemi_amd._driver import plan_mi
    batch, idx = plan_mi(*args, **kw)
    assert batch.device_hint is not None
    return batch, idx


def _columns(batch):
    arena = batch.device_arena.cpu().numpy()
    hint = batch.device_hint.cpu().numpy()
    for (kind, k, n, dx, dy, dz, _, _, offs) in batch._inst:
        for o in offs:
            yield arena[o:o + n], hint[o:o + n]


def _hinted_vs_plain(batch):
    a = batch.run(export=True, hints=True)
    b = batch.run(export=True, hints=False)
    np.testing.assert_array_equal(a[0], b[0])
    for u, v in zip(a[1], b[1]):
        np.testing.assert_array_equal(u, v)
    for u, v in zip(a[2], b[2]):
        np.testing.assert_array_equal(u, v)
    return a


def _check_hints(batch):
    with_hint = 0
    total = 0
    for col, h in _columns(batch):
        total += 1
        if h[0] < 0:
            continue
        with_hint += 1
        ref = np.lexsort((np.arange(len(col)), col))
        np.testing.assert_array_equal(h, ref)
    return with_hint, total


def test_lag_scan_hints_are_exact_orders():
    y, x = wl.lag_scan(30_000, seed=5)
    batch, _ = _plan(y, x, lag=np.arange(-6, 7), k=3, drop_nan=True)
    got, total = _check_hints(batch)
    assert total == 2 * 13 * 8
    assert got == total   # continuous data: every column has its hint
    _hinted_vs_plain(batch)


def test_masks_and_conditions():
    rng = np.random.default_rng(11)
    n = 12_000
    z = rng.standard_normal((n, 2))
    x = rng.standard_normal((n, 3)) + z[:, :1]
    y = x[:, 0] * 0.5 + z[:, 1] + rng.standard_normal(n)
    x[rng.choice(n, 500, replace=False), 1] = np.nan
    mask = rng.uniform(size=n) > 0.2
    for cond, cond_lag in ((z[:, 0], 0), (z, [[0, 1], [2, 0], [1, 1]])):
        batch, _ = _plan(y, x, lag=[0, 1, 3], k=4, cond=cond, cond_lag=cond_lag, mask=mask, drop_nan=True)
        got, total = _check_hints(batch)
        assert got == total
        _hinted_vs_plain(batch)


def test_ties_leave_columns_unhinted_and_results_unchanged():
    rng = np.random.default_rng(12)
    n = 20_000
    y = rng.standard_normal(n)
    x = np.column_stack([np.round(y + rng.standard_normal(n), 1),    # heavy ties
                         y + rng.standard_normal(n)])                  # none
    batch, idx = _plan(y, x, lag=[0, 2], k=3)
    n_without = 0
    for col, h in _columns(batch):
        if h[0] < 0:
            n_without += 1
        else:
            np.testing.assert_array_equal(h, np.lexsort((np.arange(len(col)), col)))
    assert n_without >= 2   # the rounded column at both lags
    _hinted_vs_plain(batch)


def test_preprocess_false_with_duplicates():
    rng = np.random.default_rng(13)
    n = 15_000
    y = rng.standard_normal(n)
    x = y + rng.standard_normal(n)
    x[1::97] = x[0:-1:97][:len(x[1::97])]   # pairs of equal values: runs of two
    batch, _ = _plan(y, x, lag=[0, 1], k=3, preprocess=False)
    got, total = _check_hints(batch)
    assert got == total
    _hinted_vs_plain(batch)


@pytest.mark.parametrize("how", ["swap", "reverse", "out_of_range", "negative", "duplicate", "absent"])
def test_corrupted_hints_change_nothing(how):
    import torch
    y, x = wl.lag_scan(8_000, seed=6)
    batch, _ = _plan(y, x[:, :3], lag=[0, 1], k=3, drop_nan=True)
    plain = batch.run(export=True, hints=False)
    h = batch.device_hint.cpu().numpy()
    kind, k, n, dx, dy, dz, _, _, offs = batch._inst[1]
    o = offs[0]
    if how == "swap":
        h[o + 10], h[o + 11] = h[o + 11], h[o + 10]
    elif how == "reverse":
        h[o:o + n] = h[o:o + n][::-1]
    elif how == "out_of_range":
        h[o + n - 1] = n + 5
    elif how == "negative":
        h[o + 7] = -3
    elif how == "duplicate":
        h[o + 100] = h[o + 101]
    else:
        h[o] = -1
    batch.device_hint = torch.from_numpy(h).to(batch.device_arena.device)
    hinted = batch.run(export=True, hints=True)
    np.testing.assert_array_equal(hinted[0], plain[0])
    for u, v in zip(hinted[1], plain[1]):
        np.testing.assert_array_equal(u, v)
    for u, v in zip(hinted[2], plain[2]):
        np.testing.assert_array_equal(u, v)


def test_public_api_matches_oracle_with_hints():
    from oracle import ennemi_oracle as orc
    import ennemi_amd
    y, x = wl.lag_scan(20_000, seed=7)
    lags = np.arange(-3, 4)
    out = ennemi_amd.estimate_mi(y, x, lag=lags, k=3, drop_nan=True)
    for li, j in [(0, 0), (3, 5), (6, 7), (2, 2)]:
        p = (x[:, j], y, int(lags[li]), 3, -3, 3, None, None, np.asarray([0]), False, False, True, True)
        w, _ = orc.lagged_task(p, backend="kdtree")
        assert abs(out[li, j] - w) <= 1e-12, (li, j, out[li, j], w)
