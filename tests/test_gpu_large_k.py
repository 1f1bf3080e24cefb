"""Any k: the heap-list kNN (k+1 above the 32-slot register lists).

k+1 <= 32 uses register lists; 33..128 a per-lane max-heap in LDS; larger k
a heap in a global buffer.  Radii and counts must stay bit-exact against the
oracle and values within 1e-12, in every mode, with several list classes in
one batch.  The reference's own large-k integration tests
(tests/integration/test_unbiasedness.py, test_mixture_distribution.py:45)
run through the public API and are checked both for their statistical
assertions and against the oracle's cKDTree path.
"""

import zlib
from math import e, log, pi

import numpy as np
import pytest

from oracle import ennemi_oracle as orc
from tests.test_gpu_parity import MODE_NAMES, _batch_one, _close, _cols, _data, _modes, _nat

pytestmark = pytest.mark.gpu

K_VALUES = [31, 32, 40, 100, 127, 128, 150, 300]


@pytest.mark.parametrize("mode", MODE_NAMES)
@pytest.mark.parametrize("k", K_VALUES)
def test_large_k_ksg(k, mode):
    nat = _nat()
    a = _data(zlib.crc32(b"lk_ksg") + k, 1400, 2, rounding=2 if k == 40 else None)
    X, Y = a[:, :1], a[:, 1:]
    v, eps, c = _batch_one(nat.KIND_KSG, k, _cols(X), _cols(Y), flags=_modes()[mode])
    w, ref = orc.estimate_single_mi(X, Y, k, detail=True)
    np.testing.assert_array_equal(eps, ref["eps"])
    np.testing.assert_array_equal(c[0], ref["nx"])
    np.testing.assert_array_equal(c[1], ref["ny"])
    assert _close(v, w), (v, w)


@pytest.mark.parametrize("mode", ["f32", "fp64", "brute"])
@pytest.mark.parametrize("k", [32, 100, 127, 200])
def test_large_k_cmi_and_entropy(k, mode):
    nat = _nat()
    a = _data(zlib.crc32(b"lk_cmi") + k, 1000, 4)
    X, Y, Z = a[:, :1], a[:, 1:2], a[:, 2:]
    v, eps, c = _batch_one(nat.KIND_CMI, k, _cols(X), _cols(Y), _cols(Z), flags=_modes()[mode])
    w, ref = orc.estimate_conditional_mi(X, Y, Z, k, detail=True)
    np.testing.assert_array_equal(eps, ref["eps"])
    for s, key in enumerate(("nxz", "nyz", "nz")):
        np.testing.assert_array_equal(c[s], ref[key], err_msg=key)
    assert _close(v, w), (v, w)
    v, eps, _ = _batch_one(nat.KIND_KL, k, _cols(a[:, :3]), flags=_modes()[mode])
    w, ref = orc.estimate_single_entropy(a[:, :3], k, detail=True)
    np.testing.assert_array_equal(eps, ref["eps"])
    assert _close(v, w), (v, w)


def test_large_k_mixed_batch_and_k_at_least_n():
    """Register, LDS-heap and global-heap instances in one batch; an
    instance with k >= n (every radius +inf) and a Ross class with at most
    k points."""
    from ennemi_amd._batch import Batch
    from ennemi_amd import _estimators as est
    nat = _nat()
    b = Batch()
    cases = []
    for t, (k, n) in enumerate([(3, 1500), (40, 1300), (100, 1100), (250, 900), (60, 50), (8, 9)]):
        a = _data(700 + t, n, 2)
        b.add(nat.KIND_KSG, k, _cols(a[:, :1]), _cols(a[:, 1:]))
        cases.append((k, a))
    rng = np.random.default_rng(5)
    y = rng.choice(3, size=900, p=[0.6, 0.35, 0.05])
    x = rng.normal(size=900) + y
    values, counts, order, grouped = est._grouped(y, _cols(x))
    b.add(nat.KIND_ROSS, 60, grouped, class_sizes=counts)
    vals, eps, cnts = b.run(export=True)
    for i, (k, a) in enumerate(cases):
        w, ref = orc.estimate_single_mi(a[:, :1], a[:, 1:], k, detail=True)
        np.testing.assert_array_equal(eps[i], ref["eps"], err_msg=f"k={k}")
        np.testing.assert_array_equal(cnts[i][0], ref["nx"])
        np.testing.assert_array_equal(cnts[i][1], ref["ny"])
        assert _close(vals[i], w), (k, vals[i], w)
    w, ref = orc.estimate_semidiscrete_mi(x, y, 60, detail=True)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    np.testing.assert_array_equal(eps[-1][inv], ref["eps"])
    assert _close(vals[-1], w), (vals[-1], w)


def test_reference_unbiasedness_calls():
    """The reference's tests/integration/test_unbiasedness.py, with the
    values also checked against the oracle (cKDTree) within 1e-12."""
    import ennemi_amd as em

    def both(*args, **kw):
        got = em.estimate_mi(*args, **kw).item()
        ref = orc.estimate_mi(*args, backend="kdtree", **kw).item()
        assert _close(got, ref), (kw.get("k"), got, ref)
        return got

    rng = np.random.default_rng(0)
    data = rng.multivariate_normal([0, 0], [[1, 0.8], [0.8, 1]], size=20_000)
    mi_3 = both(data[:, 0], data[:, 1], k=3)
    mi_100 = both(data[:, 0], data[:, 1], k=100)
    expected = -0.5 * log(1 - 0.8 ** 2)
    assert abs(mi_3 - expected) < 0.005
    assert abs(mi_100 - expected) > abs(mi_3 - expected) + 0.005

    rng = np.random.default_rng(0)
    data = rng.multivariate_normal([0, 0], [[1, 0], [0, 1]], size=20_000)
    mi_3 = both(data[:, 0], data[:, 1], k=3)
    mi_100 = both(data[:, 0], data[:, 1], k=100)
    assert abs(mi_100) < 0.004
    assert mi_3 - mi_100 > 0.002

    rng = np.random.default_rng(0)
    x = rng.normal(0.0, 1.0, size=20_000)
    y = rng.normal(0.0, 1.0, size=20_000)
    mi_3 = both(x, x + y, cond=x, k=3)
    mi_100 = both(x, x + y, cond=x, k=100)
    assert abs(mi_100) < 0.005
    assert abs(mi_3 - mi_100) > 0.05

    rng = np.random.default_rng(0)
    x = rng.normal(size=20_000)
    h_1 = em.estimate_entropy(x, k=1).item()
    h_100 = em.estimate_entropy(x, k=100).item()
    assert _close(h_100, orc.estimate_single_entropy(x, 100, backend="kdtree"))
    expected = 0.5 * log(2 * pi * e)
    assert h_1 > expected + 0.005 and h_100 < expected - 0.005
    assert abs(h_1 - expected) < 0.04 and abs(h_100 - expected) < 0.01

    rng = np.random.default_rng(0)
    cov = np.asarray([[1, 0.6, 0.3], [0.6, 2, 0.1], [0.3, 0.1, 1]])
    data = rng.multivariate_normal([0, 0, 0], cov, size=20_000)
    h_5 = em.estimate_entropy(data[:, :2], cond=data[:, 2], multidim=True, k=5).item()
    h_50 = em.estimate_entropy(data[:, :2], cond=data[:, 2], multidim=True, k=50).item()
    ref_50 = (orc.estimate_single_entropy(data, 50, backend="kdtree")
              - orc.estimate_single_entropy(data[:, 2], 50, backend="kdtree"))
    assert abs(h_50 - ref_50) <= 2e-12, (h_50, ref_50)
    expected = 0.5 * (log(np.linalg.det(2 * pi * e * cov)) - log(2 * pi * e))
    assert abs(h_5 - expected) < 0.005
    assert abs(h_50 - expected) < 0.03
