"""Discrete plug-in estimators with device counting (enm_unique_rows).

Numeric discrete data is factorised on the GPU; the estimates must equal
the reference's outputs recorded in tests/golden/discrete_cases.npz (the
formulas are evaluated in the reference's order, so bit for bit), with the
reference's warnings.  Strings are counted on the host.  Large and
many-valued inputs are checked against the oracle, including row sets whose
value combinations do not fit one float64 key.
"""

import os
import warnings

import numpy as np
import pytest

from oracle import ennemi_oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(HERE, "golden", "discrete_cases.npz"), allow_pickle=False)


@pytest.fixture
def device_only(monkeypatch):
    """Fail if numeric data reaches the host counting path."""
    from ennemi_amd import _discrete
    host = _discrete._host_unique

    def no_host(cols):
        if all(_discrete._numeric(c) for c in cols):
            raise AssertionError("numeric data counted on the host")
        return host(cols)
    monkeypatch.setattr(_discrete, "_host_unique", no_host)


def test_estimator_level_matches_reference(g, device_only):
    from ennemi_amd import _discrete as d
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert d.estimate_conditional_discrete_mi(g["est_cond_disc/x"], g["est_cond_disc/y"],
                                                  g["est_cond_disc/cond"]) == float(g["est_cond_disc/out"])
        for name in ("est_disc_mi_mixed", "est_disc_mi_bool", "est_disc_mi_str", "est_disc_mi_many"):
            assert d.estimate_discrete_mi(g[f"{name}/x"], g[f"{name}/y"]) == float(g[f"{name}/out"]), name
        assert d.estimate_conditional_discrete_mi(g["est_cond_disc_many/x"], g["est_cond_disc_many/y"],
                                                  g["est_cond_disc_many/cond"]) == \
            float(g["est_cond_disc_many/out"])
        assert abs(d.estimate_discrete_entropy(g["est_disc_entropy_2d/x"])
                   - float(g["est_disc_entropy_2d/out"])) <= 1e-14


def test_public_api_matches_reference(g, device_only):
    import ennemi_amd
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for name in ("disc_cond1", "disc_cond2"):
            out = ennemi_amd.estimate_mi(g[f"{name}/y"], g[f"{name}/x"], g[f"{name}/lag"],
                                         cond=g[f"{name}/cond"], discrete_x=True, discrete_y=True)
            np.testing.assert_array_equal(out, g[f"{name}/out"])
        out = ennemi_amd.estimate_entropy(g["entropy_discrete_multidim/x"], discrete=True, multidim=True)
        assert abs(float(out) - float(g["entropy_discrete_multidim/out"])) <= 1e-14
        out = ennemi_amd.estimate_entropy(g["entropy_discrete_cond/x"], cond=g["entropy_discrete_cond/cond"],
                                          discrete=True)
        assert abs(float(out) - float(g["entropy_discrete_cond/out"])) <= 1e-14


def test_many_values_warn(g):
    from ennemi_amd import _discrete as d
    with pytest.warns(UserWarning, match="relatively many unique values"):
        d.estimate_discrete_mi(g["est_disc_mi_many/x"], g["est_disc_mi_many/y"])
    with pytest.warns(UserWarning, match="relatively many unique values"):
        d.estimate_conditional_discrete_mi(g["est_cond_disc_many/x"], g["est_cond_disc_many/y"],
                                           g["est_cond_disc_many/cond"])
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        d.estimate_discrete_mi(g["est_disc_mi_mixed/x"], g["est_disc_mi_mixed/y"])


@pytest.mark.parametrize("n", [1, 7, 2049, 1_000_003])
def test_large_and_tiny_inputs_against_oracle(n, device_only):
    from ennemi_amd import _discrete as d
    rng = np.random.default_rng(n)
    z = rng.integers(0, 6, size=(n, 2))
    x = (z[:, 0] * 3 + rng.integers(0, 7, n)) % 11 - 5.0
    y = (x.astype(np.int64) + z[:, 1] + rng.integers(0, 3, n)) % 9
    x[::5] = -0.0  # -0.0 and 0.0 are one value
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert d.estimate_discrete_mi(x, y) == orc.estimate_discrete_mi(x, y)
        assert d.estimate_conditional_discrete_mi(x, y, z) == orc.estimate_conditional_discrete_mi(x, y, z)
        h = d.estimate_discrete_entropy(np.column_stack([x, z]))
        assert abs(h - orc.estimate_discrete_entropy(np.column_stack([x, z]))) <= 1e-12


def test_row_keys_beyond_float64_are_numbered_a_column_at_a_time(device_only):
    """Four columns of 30,000 distinct values each: 8.1e17 combinations, more
    than a float64 key holds exactly."""
    from ennemi_amd import _discrete as d
    rng = np.random.default_rng(8)
    n = 60_000
    rows = np.column_stack([rng.permutation(n) % 30_000 for _ in range(4)]).astype(np.int64)
    rows[100:200] = rows[:100]   # some repeated rows
    h = d.estimate_discrete_entropy(rows)
    assert abs(h - orc.estimate_discrete_entropy(rows)) <= 1e-12
