"""The public API on the GPU against the reference's own outputs.

Each case of tests/golden/driver_cases.npz was produced by running ennemi
1.5.0; ennemi_amd must reproduce it within 1e-12 (and the same +-inf / NaN).
"""

import warnings

import numpy as np
import pytest

from tests.test_oracle import (ENTROPY_CASES, MI_CASES, PAIRWISE_CASES, _assert_same, _opt,
                               mi_kwargs)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", MI_CASES)
def test_estimate_mi_matches_reference(golden_driver, name):
    import ennemi_amd
    g = golden_driver
    kw = mi_kwargs(g, name)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = ennemi_amd.estimate_mi(g[f"{name}/y"], g[f"{name}/x"], g[f"{name}/lag"], **kw)
    _assert_same(out, g[f"{name}/out"])


def test_estimate_corr_matches_reference(golden_driver):
    import ennemi_amd
    g = golden_driver
    out = ennemi_amd.estimate_corr(g["corr/y"], g["corr/x"], g["corr/lag"], k=4)
    _assert_same(out, g["corr/out"])


@pytest.mark.parametrize("name", sorted(PAIRWISE_CASES))
def test_pairwise_matches_reference(golden_driver, name):
    import ennemi_amd
    g = golden_driver
    kw = dict(PAIRWISE_CASES[name])
    for key in ["cond", "mask", "discrete"]:
        v = _opt(g, name, key)
        if v is not None:
            kw[key] = v
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = ennemi_amd.pairwise_mi(g[f"{name}/data"], k=int(g[f"{name}/k"]), **kw)
    _assert_same(out, g[f"{name}/out"])


@pytest.mark.parametrize("name", sorted(ENTROPY_CASES))
def test_entropy_matches_reference(golden_driver, name):
    import ennemi_amd
    g = golden_driver
    kw = dict(ENTROPY_CASES[name])
    for key in ["cond", "mask"]:
        v = _opt(g, name, key)
        if v is not None:
            kw[key] = v
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = ennemi_amd.estimate_entropy(g[f"{name}/x"], k=int(g[f"{name}/k"]), **kw)
    _assert_same(out, g[f"{name}/out"])


def test_estimators_match_reference(golden_driver):
    from ennemi_amd import _estimators as est
    g = golden_driver
    _assert_same(est.estimate_single_mi(g["est_single_mi/x"], g["est_single_mi/y"], 3),
                 g["est_single_mi/out"])
    _assert_same(est.estimate_conditional_mi(g["est_cmi/x"], g["est_cmi/y"], g["est_cmi/cond"], 4),
                 g["est_cmi/out"])
    _assert_same(est.estimate_semidiscrete_mi(g["est_ross_multid/x"], g["est_ross_multid/y"], 3),
                 g["est_ross_multid/out"])
    _assert_same(est.estimate_semidiscrete_mi(g["est_ross_tiny_class/x"],
                                              g["est_ross_tiny_class/y"], 3),
                 g["est_ross_tiny_class/out"])
    _assert_same(est.estimate_conditional_semidiscrete_mi(g["est_cross/x"], g["est_cross/y"],
                                                          g["est_cross/cond"], 2),
                 g["est_cross/out"])
    _assert_same(est.estimate_single_entropy(g["est_entropy_4d/x"], 3), g["est_entropy_4d/out"])


def test_c1_headline_config(golden_driver):
    import hashlib
    import ennemi_amd
    rng = np.random.default_rng(101)
    d1 = rng.multivariate_normal([0, 0], [[1, 0.6], [0.6, 1]], size=10_000)
    assert hashlib.sha256(d1.tobytes()).hexdigest() == str(golden_driver["c1/sha256"])
    _assert_same(ennemi_amd.estimate_mi(d1[:, 1], d1[:, 0], k=3, normalize=True),
                 golden_driver["c1/out"])
    _assert_same(ennemi_amd.estimate_mi(d1[:, 1], d1[:, 0], k=3), golden_driver["c1/out_raw"])
