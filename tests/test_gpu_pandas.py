"""pandas inputs and callbacks through the public API (GPU).

pandas DataFrame / Series inputs come back as DataFrames with the
reference's index and columns (ennemi/_driver.py:69-75, 154-161, 354-360,
619-623), holding exactly the values of the same call on NumPy arrays.
Callbacks fire once per task with the reference's arguments; the reference's
own tests check membership only (tests/unit/test_driver.py:1218-1233,
1523-1541), and ennemi_amd calls them after the batch, in task order.
tests/golden/example_data.csv is the data file of the reference's
tests/unit/test_driver.py:611-661.
"""

import os
from itertools import product

import numpy as np
import pandas as pd
import pytest

from oracle import ennemi_oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12


def _example():
    return pd.read_csv(os.path.join(HERE, "golden", "example_data.csv"))


def test_estimate_mi_data_frame_round_trip():
    """The reference's test_pandas_data_frame (test_driver.py:638-661), plus
    the values of the NumPy call and of the oracle."""
    import ennemi_amd as em
    data = _example()
    actual = em.estimate_mi(data["y"], data[["x1", "x2", "x3"]], lag=[0, 1, 3], k=5)
    assert isinstance(actual, pd.DataFrame)
    assert list(actual.index) == [0, 1, 3]
    assert list(actual.columns) == ["x1", "x2", "x3"]
    y = data["y"].to_numpy()
    x = data[["x1", "x2", "x3"]].to_numpy()
    plain = em.estimate_mi(y, x, lag=[0, 1, 3], k=5)
    np.testing.assert_array_equal(actual.to_numpy(), plain)
    ref = orc.estimate_mi(y, x, lag=[0, 1, 3], k=5, backend="kdtree")
    assert np.max(np.abs(plain - ref)) <= TOL
    # The reference's statistical assertions.
    assert abs(actual.loc[0, "x1"]) < 0.04 and actual.loc[1, "x1"] > 0.4 and abs(actual.loc[3, "x1"]) < 0.04
    for i in [0, 1, 3]:
        assert abs(actual.loc[i, "x2"]) < 0.04
    assert abs(actual.loc[0, "x3"]) < 0.04 and abs(actual.loc[1, "x3"]) < 0.04 and actual.loc[3, "x3"] > 0.15
    # A Series x gives a one-column frame named after the series.
    one = em.estimate_mi(data["y"], data["x1"], lag=[0, 1], k=5)
    assert isinstance(one, pd.DataFrame) and list(one.columns) == ["x1"]
    np.testing.assert_array_equal(one.to_numpy()[:, 0],
                                  em.estimate_mi(y, x[:, 0], lag=[0, 1], k=5)[:, 0])


def test_series_with_custom_index_and_mask():
    """test_driver.py:663-679: Series on a non-zero index, a boolean mask
    Series, NaN rows masked out."""
    import ennemi_amd as em
    rng = np.random.default_rng(20)
    data = rng.multivariate_normal([0, 0], [[1, 0.8], [0.8, 1]], size=1000)
    x = data[:, 0]
    y = data[:, 1].copy()
    y[150:450] = np.nan
    mask = np.full(1000, True)
    mask[150:450] = False
    df = pd.DataFrame({"x": x, "y": y, "mask": mask}, index=range(2000, 3000))
    masked = em.estimate_mi(df["y"], df["x"], mask=df["mask"])
    assert isinstance(masked, pd.DataFrame) and list(masked.columns) == ["x"]
    assert abs(masked.loc[0, "x"] - (-0.5 * np.log(1 - 0.8 ** 2))) < 0.03
    plain = em.estimate_mi(y, x, mask=mask)
    np.testing.assert_array_equal(masked.to_numpy(), plain)
    ref = orc.estimate_mi(y, x, mask=mask, backend="kdtree")
    assert abs(plain[0, 0] - ref[0, 0]) <= TOL


def test_pairwise_and_entropy_data_frames_and_normalize():
    import ennemi_amd as em
    data = _example()
    frame = em.pairwise_mi(data, k=4)
    assert isinstance(frame, pd.DataFrame)
    assert list(frame.index) == list(data.columns) and list(frame.columns) == list(data.columns)
    plain = em.pairwise_mi(data.to_numpy(), k=4)
    np.testing.assert_array_equal(frame.to_numpy(), plain)
    ref = orc.pairwise_mi(data.to_numpy(), k=4, backend="kdtree")
    off = ~np.eye(4, dtype=bool)
    assert np.max(np.abs(plain[off] - ref[off])) <= TOL
    # normalize_mi keeps the frame.
    norm = em.normalize_mi(frame)
    assert isinstance(norm, pd.DataFrame) and list(norm.columns) == list(data.columns)
    np.testing.assert_array_equal(norm.to_numpy(), em.normalize_mi(plain))
    # Entropy of every column: a one-row frame; multidim: a plain value.
    ent = em.estimate_entropy(data, k=3)
    assert isinstance(ent, pd.DataFrame) and ent.shape == (1, 4)
    assert list(ent.columns) == list(data.columns)
    np.testing.assert_array_equal(ent.to_numpy()[0], em.estimate_entropy(data.to_numpy(), k=3))
    for c, name in enumerate(data.columns):
        assert abs(ent.loc[0, name] - orc.estimate_single_entropy(data[name].to_numpy(), 3,
                                                                 backend="kdtree")) <= TOL
    ser = em.estimate_entropy(data["x2"], k=3)
    assert isinstance(ser, pd.DataFrame) and list(ser.columns) == ["x2"]
    md = em.estimate_entropy(data[["x1", "x2"]], k=3, multidim=True)
    assert not isinstance(md, pd.DataFrame)


def test_pairwise_corr_on_a_time_indexed_frame():
    """The shape of the reference's pandas workflow test (hourly data with
    a DatetimeIndex, an afternoon mask, a condition), on synthetic data."""
    import ennemi_amd as em
    rng = np.random.default_rng(77)
    idx = pd.date_range("2010-01-01", periods=24 * 365 * 10, freq="h")
    doy = idx.dayofyear.to_numpy().astype(float)
    temp = 10 * np.sin(2 * np.pi * (doy - 100) / 365) + rng.normal(size=len(idx))
    wind = rng.uniform(0, 360, size=len(idx)) + 5 * temp
    df = pd.DataFrame({"Temperature": temp, "WindDir": wind, "DayOfYear": doy}, index=idx)
    mask = df.index.hour == 13
    uncond = em.pairwise_corr(df, mask=mask)
    cond = em.pairwise_corr(df, mask=mask, cond=df["DayOfYear"])
    for res in (uncond, cond):
        assert isinstance(res, pd.DataFrame) and res.shape == (3, 3)
        assert res.loc["Temperature", "DayOfYear"] == res.loc["DayOfYear", "Temperature"]
    assert uncond.loc["Temperature", "DayOfYear"] > 0.8
    plain = em.pairwise_corr(df.to_numpy(), mask=mask)
    np.testing.assert_array_equal(uncond.to_numpy(), plain)
    auto = em.estimate_corr(df["Temperature"], df["Temperature"], lag=[0, -24, -240],
                            cond=df["DayOfYear"], mask=mask)
    assert isinstance(auto, pd.DataFrame) and list(auto.index) == [0, -24, -240]
    assert list(auto.columns) == ["Temperature"]
    assert abs(auto.loc[0, "Temperature"] - 1) < 0.02
    plain_auto = em.estimate_corr(temp, temp, lag=[0, -24, -240], cond=doy, mask=mask)
    np.testing.assert_array_equal(auto.to_numpy(), plain_auto)


@pytest.mark.parametrize("n", [100, 3000])
def test_estimate_mi_callback_membership(n):
    """test_driver.py:1218-1233."""
    import ennemi_amd as em
    seen = []
    rng = np.random.default_rng(18)
    data = rng.multivariate_normal([0] * 11, np.eye(11), n)
    em.estimate_mi(data[:, 0], data[:, 1:], lag=[7, -3, 2], callback=lambda v, lag: seen.append((v, lag)))
    assert len(seen) == 30
    for (i, lag) in product(range(10), [7, -3, 2]):
        assert (i, lag) in seen
    seen.clear()
    em.estimate_mi([1, 2, 3, 4, 5, 6, 7], [2, 3, 4, 5, 6, 7, 8], callback=lambda v, lag: seen.append((v, lag)))
    assert seen == [(0, 0)]


@pytest.mark.parametrize("n", [100, 3000])
def test_pairwise_callback_membership(n):
    """test_driver.py:1523-1541."""
    import ennemi_amd as em
    seen = []
    rng = np.random.default_rng(105)
    data = rng.multivariate_normal([0] * 10, np.eye(10), n)
    em.pairwise_mi(data, callback=lambda i, j: seen.append((i, j)))
    assert len(seen) == 45
    for (i, j) in product(range(10), range(10)):
        assert ((i, j) in seen) == (i < j)
