"""Pins for the oracle (CPU only).

The oracle in ``oracle/`` is the checker for the HIP product, so it is checked
first against things that do not depend on it:

* per-point radii and counts recorded from ``scipy.spatial.cKDTree`` (called as
  ennemi calls it) on inputs with exact ties, duplicates and negative radii;
* the reference's own outputs for its public and estimator-level functions
  (``tests/golden/driver_cases.npz``, written by running the reference);
* closed forms: psi values, Gaussian MI / entropy, and invariants.

A planted mistake in the oracle (strict instead of inclusive count, dropped
1e-12 shift, k-th instead of (k+1)-th neighbour, true digamma instead of the
truncated series) must make at least one of these checks fail; the last test
plants each one and checks that it does.
"""

import math
import warnings

import numpy as np
import pytest
from scipy.special import psi as true_psi

from oracle import ennemi_oracle as orc


KNN_SETS = ["normal2", "rounded2", "rounded_noise2", "normal1", "rounded1", "normal4",
            "heavy3", "dup5"]


@pytest.mark.parametrize("name", KNN_SETS)
@pytest.mark.parametrize("k", [1, 3, 7])
@pytest.mark.parametrize("backend", ["brute", "numpy", "kdtree"])
def test_radii_and_counts_match_ckdtree(golden_knn, name, k, backend):
    pts = golden_knn[f"{name}/pts"]
    eps = orc.knn_eps(pts, k, backend)
    np.testing.assert_array_equal(eps, golden_knn[f"{name}/k{k}/eps"])
    cnt = orc.ball_counts(pts, pts, eps - 1e-12, backend)
    np.testing.assert_array_equal(cnt, golden_knn[f"{name}/k{k}/count"])
    cnt0 = orc.ball_counts(pts, pts, eps, backend)
    np.testing.assert_array_equal(cnt0, golden_knn[f"{name}/k{k}/count_at_eps"])


def test_golden_sets_exercise_edge_cases(golden_knn):
    # The fixtures must actually contain ties, zero radii and negative radii.
    eps = golden_knn["rounded2/k1/eps"]
    assert np.sum(eps == 0) > 10
    assert np.sum(golden_knn["rounded2/k1/count"] == 0) > 10
    assert np.sum(golden_knn["dup5/k1/eps"] == 0) == 300


def test_fewer_points_than_k_gives_inf():
    pts = np.random.default_rng(0).normal(size=(3, 2))
    for backend in ["brute", "numpy", "kdtree"]:
        assert np.all(np.isinf(orc.knn_eps(pts, 3, backend)))
        assert list(orc.ball_counts(pts, pts, np.full(3, np.inf), backend)) == [3, 3, 3]


def test_psi_matches_reference_and_closed_form(golden_driver):
    ref = golden_driver["psi/out"]
    for i in range(50):
        assert float(orc.psi(i)) == ref[i]
    assert orc.psi(0) == np.inf
    assert float(orc.psi(1)) == -0.5772156649015331
    for i in range(2, 200):
        assert abs(float(orc.psi(i)) - true_psi(i)) < 1e-4
    # The series is not the true digamma: at 2 it differs by about 1e-5.
    assert abs(float(orc.psi(2)) - true_psi(2)) > 1e-7
    # Any zero in an array gives the scalar inf.
    assert orc.psi(np.array([3, 0, 5])).shape == ()


def _assert_same(actual, expected):
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    assert actual.shape == expected.shape
    both_nan = np.isnan(actual) & np.isnan(expected)
    same_inf = np.isinf(actual) & (actual == expected)
    fin = ~(both_nan | same_inf)
    assert np.all(np.isfinite(actual[fin])) and np.all(np.isfinite(expected[fin]))
    np.testing.assert_allclose(actual[fin], expected[fin], rtol=0, atol=1e-12)


def _opt(g, name, key):
    full = f"{name}/{key}"
    return g[full] if full in g.files else None


MI_CASES = ["mi_lags", "mi_k1_norm", "mi_cond1", "mi_cond2", "mi_mask", "mi_nan",
            "mi_nan_uncond", "mi_discrete_y", "mi_discrete_x", "mi_discrete_y_cond",
            "mi_discrete_x_cond", "mi_discrete_both", "mi_dup_nopre", "mi_dup_nopre_cond",
            "mi_ties_pre", "mi_nopre"]


def mi_kwargs(g, name):
    kw = dict(k=int(g[f"{name}/k"]))
    for key in ["cond", "cond_lag", "mask"]:
        v = _opt(g, name, key)
        if v is not None:
            kw[key] = v
    kw["normalize"] = name.endswith("_norm")
    kw["discrete_y"] = "discrete_y" in name or "discrete_both" in name
    kw["discrete_x"] = "discrete_x" in name or "discrete_both" in name
    kw["drop_nan"] = "_nan" in name
    kw["preprocess"] = "nopre" not in name
    return kw


@pytest.mark.parametrize("name", MI_CASES)
def test_oracle_estimate_mi_matches_reference(golden_driver, name):
    g = golden_driver
    kw = mi_kwargs(g, name)
    with np.errstate(all="ignore"):
        out = orc.estimate_mi(g[f"{name}/y"], g[f"{name}/x"], g[f"{name}/lag"], threads=1, **kw)
    _assert_same(out, g[f"{name}/out"])


def test_oracle_corr_matches_reference(golden_driver):
    g = golden_driver
    out = orc.estimate_mi(g["corr/y"], g["corr/x"], g["corr/lag"], k=4, normalize=True, threads=1)
    _assert_same(out, g["corr/out"])


PAIRWISE_CASES = {"pairwise": {}, "pairwise_norm": {"normalize": True},
                  "pairwise_cond_mask": {}, "pairwise_mixed": {},
                  "pairwise_nopre": {"preprocess": False}, "pairwise_nan": {"drop_nan": True}}


@pytest.mark.parametrize("name", sorted(PAIRWISE_CASES))
def test_oracle_pairwise_matches_reference(golden_driver, name):
    g = golden_driver
    kw = dict(PAIRWISE_CASES[name])
    for key in ["cond", "mask", "discrete"]:
        v = _opt(g, name, key)
        if v is not None:
            kw[key] = v
    out = orc.pairwise_mi(g[f"{name}/data"], k=int(g[f"{name}/k"]), threads=1, **kw)
    _assert_same(out, g[f"{name}/out"])


ENTROPY_CASES = {"entropy_cols": {}, "entropy_multidim": {"multidim": True},
                 "entropy_mask": {}, "entropy_cond": {},
                 "entropy_cond_multidim": {"multidim": True},
                 "entropy_nan_multidim": {"multidim": True, "drop_nan": True},
                 "entropy_nan_cols": {"drop_nan": True}, "entropy_dup": {},
                 "entropy_discrete": {"discrete": True}}


@pytest.mark.parametrize("name", sorted(ENTROPY_CASES))
def test_oracle_entropy_matches_reference(golden_driver, name):
    g = golden_driver
    kw = dict(ENTROPY_CASES[name])
    for key in ["cond", "mask"]:
        v = _opt(g, name, key)
        if v is not None:
            kw[key] = v
    with np.errstate(all="ignore"):
        out = orc.estimate_entropy(g[f"{name}/x"], k=int(g[f"{name}/k"]), **kw)
    _assert_same(out, g[f"{name}/out"])


def test_oracle_estimators_match_reference(golden_driver):
    g = golden_driver
    _assert_same(orc.estimate_single_mi(g["est_single_mi/x"], g["est_single_mi/y"], 3),
                 g["est_single_mi/out"])
    _assert_same(orc.estimate_conditional_mi(g["est_cmi/x"], g["est_cmi/y"], g["est_cmi/cond"], 4),
                 g["est_cmi/out"])
    _assert_same(orc.estimate_semidiscrete_mi(g["est_ross_multid/x"], g["est_ross_multid/y"], 3),
                 g["est_ross_multid/out"])
    _assert_same(orc.estimate_semidiscrete_mi(g["est_ross_tiny_class/x"],
                                              g["est_ross_tiny_class/y"], 3),
                 g["est_ross_tiny_class/out"])
    _assert_same(orc.estimate_conditional_semidiscrete_mi(g["est_cross/x"], g["est_cross/y"],
                                                          g["est_cross/cond"], 2),
                 g["est_cross/out"])
    _assert_same(orc.estimate_single_entropy(g["est_entropy_4d/x"], 3), g["est_entropy_4d/out"])


def test_oracle_c1_headline_config(golden_driver):
    import hashlib
    rng = np.random.default_rng(101)
    d1 = rng.multivariate_normal([0, 0], [[1, 0.6], [0.6, 1]], size=10_000)
    assert hashlib.sha256(d1.tobytes()).hexdigest() == str(golden_driver["c1/sha256"])
    out = orc.estimate_mi(d1[:, 1], d1[:, 0], k=3, threads=1)
    _assert_same(out, golden_driver["c1/out_raw"])
    # Sanity: the analytic MI of a bivariate Gaussian with rho = 0.6.
    assert abs(out[0, 0] - (-0.5 * math.log(1 - 0.36))) < 0.02


def test_gaussian_kats():
    rng = np.random.default_rng(0)
    cov = np.array([[1, 0.5], [0.5, 1]])
    data = rng.multivariate_normal([0, 0], cov, size=2000)
    mi = orc.estimate_single_mi(data[:, 0], data[:, 1], 3)
    assert abs(mi - (-0.5 * math.log(1 - 0.25))) < 0.02
    # x <-> y symmetry (test_entropy_estimators.py:149-159 pins it to 1e-5)
    assert abs(mi - orc.estimate_single_mi(data[:, 1], data[:, 0], 3)) < 1e-12
    h = orc.estimate_single_entropy(data, 3)
    assert abs(h - 0.5 * math.log(np.linalg.det(2 * math.pi * math.e * cov))) < 0.03


def test_noise_stream_is_a_prefix():
    # The rescaling draws noise for x, y, z in order from a fresh fixed-seed
    # generator; the draws are a prefix of one stream and equal 1e-10 times the
    # standard normals (the product's batched planner relies on both facts).
    a = np.random.default_rng(2_718281828)
    first = a.normal(0.0, 1e-10, 1000)
    second = a.normal(0.0, 1e-10, (500, 3))
    b = np.random.default_rng(2_718281828).standard_normal(2500)
    assert np.array_equal(np.concatenate([first, second.ravel()]), 0.0 + 1e-10 * b)


def test_planted_mistakes_are_caught(golden_knn, golden_driver, monkeypatch):
    pts = golden_knn["rounded_noise2/pts"]
    want_eps = golden_knn["rounded_noise2/k3/eps"]
    want_cnt = golden_knn["rounded_noise2/k3/count"]

    # (1) k-th instead of (k+1)-th neighbour.
    assert not np.array_equal(orc.knn_eps(pts, 2, "numpy"), want_eps)
    # (2) strict instead of inclusive count on tied data.
    tied = golden_knn["rounded1/pts"]
    e = golden_knn["rounded1/k3/eps"]
    q = np.column_stack((tied,))
    strict = np.sum(np.abs(q - q.T) < e[:, None], axis=1)
    assert not np.array_equal(strict, golden_knn["rounded1/k3/count_at_eps"])
    # (3) the 1e-12 shift dropped.
    assert not np.array_equal(orc.ball_counts(pts, pts, want_eps, "numpy"), want_cnt)
    # (4) the true digamma instead of the truncated series.
    g = golden_driver
    monkeypatch.setattr(orc, "psi", lambda x: (np.asarray(np.inf) if np.any(np.asarray(x) == 0)
                                               else true_psi(np.asarray(x, dtype=float))))
    out = orc.estimate_single_mi(g["est_single_mi/x"], g["est_single_mi/y"], 3)
    assert abs(out - float(g["est_single_mi/out"])) > 1e-9


@pytest.fixture(scope="module")
def golden_discrete():
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "discrete_cases.npz"),
                   allow_pickle=False)


def test_oracle_discrete_estimators_match_reference(golden_discrete):
    """The oracle's plug-in estimators against the reference's outputs
    (tests/golden/discrete_cases.npz): conditional MI with one- and
    two-column conditions, mixed types, booleans, strings, entropy of rows."""
    g = golden_discrete
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert orc.estimate_conditional_discrete_mi(g["est_cond_disc/x"], g["est_cond_disc/y"],
                                                    g["est_cond_disc/cond"]) == float(g["est_cond_disc/out"])
        assert orc.estimate_conditional_discrete_mi(g["est_cond_disc_many/x"], g["est_cond_disc_many/y"],
                                                    g["est_cond_disc_many/cond"]) == \
            float(g["est_cond_disc_many/out"])
        for name in ("est_disc_mi_mixed", "est_disc_mi_bool", "est_disc_mi_str", "est_disc_mi_many"):
            assert orc.estimate_discrete_mi(g[f"{name}/x"], g[f"{name}/y"]) == float(g[f"{name}/out"]), name
        assert abs(orc.estimate_discrete_entropy(g["est_disc_entropy_2d/x"])
                   - float(g["est_disc_entropy_2d/out"])) <= 1e-14
        for name in ("disc_cond1", "disc_cond2"):
            out = orc.estimate_mi(g[f"{name}/y"], g[f"{name}/x"], g[f"{name}/lag"], cond=g[f"{name}/cond"],
                                  discrete_x=True, discrete_y=True, threads=1)
            _assert_same(out, g[f"{name}/out"])
        _assert_same(orc.estimate_entropy(g["entropy_discrete_multidim/x"], discrete=True, multidim=True),
                     g["entropy_discrete_multidim/out"])
        _assert_same(orc.estimate_entropy(g["entropy_discrete_cond/x"], cond=g["entropy_discrete_cond/cond"],
                                          discrete=True), g["entropy_discrete_cond/out"])


def test_planted_mistake_in_conditional_discrete_mi_is_caught(golden_discrete, monkeypatch):
    """Weighting each condition row by 1 instead of n_z must fail the pin."""
    g = golden_discrete
    x, y, z = g["est_cond_disc/x"], g["est_cond_disc/y"], g["est_cond_disc/cond"]
    _, inv, cnt = np.unique(z, axis=0, return_inverse=True, return_counts=True)
    inv = np.ravel(inv)
    wrong = sum(orc.estimate_discrete_mi(x[inv == i], y[inv == i]) for i in range(len(cnt))) / len(cnt)
    assert abs(wrong - float(g["est_cond_disc/out"])) > 1e-6
