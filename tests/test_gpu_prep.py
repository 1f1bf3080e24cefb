"""Device-side preparation of estimate_mi tasks (enm_prep_count/_write).

The device writes each task's instance columns (lag windows, mask, NaN
drop, rescaling with the reference's noise) straight into the batch arena.
They must equal, bit for bit, the columns the host planner makes with
NumPy's own operations (ennemi/_driver.py:795-806, 834-854, 871-902), which
tests/test_host_cpu.py pins against the oracle.  Covered: every continuous
golden driver case, the masked lag scan (C3), conditions of one and several
columns (pairwise vs row-by-row sums), integer and boolean inputs, constant
variables (no rescaling, no noise drawn, the warning), and the validation
errors with the reference's messages.
"""

import warnings

import numpy as np
import pytest

from tests import workloads as wl
from tests.test_oracle import MI_CASES, mi_kwargs

pytestmark = pytest.mark.gpu


def _columns(batch):
    """Per instance: its columns as float64 arrays (host or device batch)."""
    from ennemi_amd._batch import ArenaBatch
    out = []
    if isinstance(batch, ArenaBatch):
        arena = batch.device_arena.cpu().numpy()
        for (kind, k, n, dx, dy, dz, _, _, offs) in batch._inst:
            out.append((kind, k, [arena[o:o + n] for o in offs]))
    else:
        for (kind, k, n, dx, dy, dz, _, _, cols) in batch._inst:
            out.append((kind, k, [batch._cols[c] for c in cols]))
    return out


def _compare(args, kw):
    from ennemi_amd._batch import ArenaBatch
    from ennemi_amd._driver import plan_mi
    with warnings.catch_warnings(record=True) as w_dev:
        warnings.simplefilter("always")
        dev, idx_d = plan_mi(*args, **kw)
    with warnings.catch_warnings(record=True) as w_host:
        warnings.simplefilter("always")
        host, idx_h = plan_mi(*args, prep="host", **kw)
    assert isinstance(dev, ArenaBatch)
    assert sorted(idx_d) == sorted(idx_h)
    cd, ch = _columns(dev), _columns(host)
    assert len(cd) == len(ch)
    # The device batch orders its instances by x column; match by task.
    cd = [cd[idx_d.index(ix)] for ix in idx_h]
    for t, ((kd, kkd, a), (kh, kkh, b)) in enumerate(zip(cd, ch)):
        assert (kd, kkd, len(a)) == (kh, kkh, len(b)), t
        for c, (u, v) in enumerate(zip(a, b)):
            np.testing.assert_array_equal(u, v, err_msg=f"task {t} column {c}")
    assert sorted(str(x.message) for x in w_dev) == sorted(str(x.message) for x in w_host)
    return dev, host


CONTINUOUS = [c for c in MI_CASES if "discrete" not in c]


@pytest.mark.parametrize("name", CONTINUOUS)
def test_device_columns_equal_host_golden(golden_driver, name):
    g = golden_driver
    kw = mi_kwargs(g, name)
    kw.pop("normalize")
    _compare((g[f"{name}/y"], g[f"{name}/x"], g[f"{name}/lag"]), kw)


def test_device_columns_equal_host_lag_scan():
    y, x = wl.lag_scan(100_096)
    dev, host = _compare((y, x, np.arange(-48, 49)), dict(k=3, drop_nan=True))
    assert len(dev) == 776


@pytest.mark.parametrize("dz", [1, 2, 3])
def test_device_columns_conditions_masks_and_lags(dz):
    rng = np.random.default_rng(40 + dz)
    n = 30_011
    y = rng.normal(size=n) * 3 + 1
    x = rng.normal(size=(n, 3)) * np.array([1.0, 1e-3, 1e4])
    cond = rng.normal(size=(n, dz)) + 5
    x[rng.choice(n, 900, replace=False), 1] = np.nan
    cond[rng.choice(n, 300, replace=False), dz - 1] = np.nan
    mask = rng.random(n) < 0.8
    cond_lag = np.array([list(range(dz))]) if dz > 1 else 2
    _compare((y, x, [-5, 0, 3, 17]), dict(k=4, cond=cond, cond_lag=cond_lag, mask=mask, drop_nan=True))
    _compare((y, x[:, 0], [1, -2]), dict(k=3, cond=np.nan_to_num(cond), preprocess=False))


def test_device_columns_integer_boolean_and_short_inputs():
    rng = np.random.default_rng(9)
    n = 9000
    yi = rng.integers(-50, 50, size=n)
    xb = rng.random(n) < 0.3
    _compare((yi, xb, [0, 2]), dict(k=3))
    # Windows shorter than a block, and of a few values.
    _compare((rng.normal(size=40), rng.normal(size=40), [0, 1, -3]), dict(k=2))
    _compare((rng.normal(size=12), rng.normal(size=12), 0), dict(k=1))


def test_device_constant_variables_skip_rescaling():
    """x constant: no rescaling or noise for x, and y's noise starts at the
    first value (the reference draws nothing for x); a constant condition
    column leaves the whole condition unscaled."""
    rng = np.random.default_rng(3)
    n = 5000
    y = rng.normal(size=n)
    x = np.full(n, 2.5)
    _compare((y, x, [0, 1]), dict(k=3))
    cond = np.column_stack([rng.normal(size=n), np.full(n, -1.0)])
    _compare((y, rng.normal(size=n), 0), dict(k=3, cond=cond))
    _compare((np.full(n, 7.0), rng.normal(size=n), 0), dict(k=3, cond=cond[:, :1]))


def test_device_validation_messages():
    import ennemi_amd
    from ennemi_amd._driver import K_TOO_LARGE_MSG, NANS_LEFT_MSG
    z = np.zeros(20)
    with pytest.raises(ValueError, match="k must be smaller"):
        ennemi_amd.estimate_mi(z, z, lag=16, k=4)
    m = np.ones(20, bool)
    m[2:] = False
    with pytest.raises(ValueError) as e:
        ennemi_amd.estimate_mi(np.arange(20.0), np.arange(20.0), mask=m, k=3)
    assert str(e.value) == K_TOO_LARGE_MSG
    xn = np.arange(20.0)
    xn[3] = np.nan
    with pytest.raises(ValueError) as e:
        ennemi_amd.estimate_mi(np.arange(20.0), xn)
    assert str(e.value) == NANS_LEFT_MSG
    with pytest.raises(ValueError) as e:
        ennemi_amd.estimate_mi(np.arange(20.0), np.arange(20.0), cond=xn)
    assert str(e.value) == NANS_LEFT_MSG
    # A NaN outside the mask is not an error; with drop_nan it is dropped.
    m = np.ones(20, bool)
    m[3] = False
    assert np.isfinite(ennemi_amd.estimate_mi(np.arange(20.0) ** 2, xn, mask=m)).all()
    assert np.isfinite(ennemi_amd.estimate_mi(np.arange(20.0) ** 2, xn, drop_nan=True)).all()
