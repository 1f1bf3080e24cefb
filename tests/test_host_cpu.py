"""Host-side checks that need no GPU.

* The C ABI: the library loads and exports every function include/ennemi_hip.h
  declares; the ctypes structures match the C layout.
* Validation: the reference's exact error messages and warnings
  (tests/unit/test_driver.py:26-37 of ennemi), raised before any GPU work.
* The host planner: the float64 columns handed to the GPU for every task are
  bit-identical to the oracle's restatement of _lagged_mi -> _apply_masks ->
  _rescale_data, including the shared noise stream and pairwise role sharing.
"""

import ctypes
import os
import re
import warnings

import numpy as np
import pytest

from oracle import ennemi_oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_function():
    import ennemi_amd._native as nat
    header = open(os.path.join(ROOT, "include", "ennemi_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(enm_\w+)\s*\(", header, re.M))
    assert declared == set(nat.EXPORTED_SYMBOLS)
    lib = nat.load_library()
    for name in declared:
        assert hasattr(lib, name), name


def test_struct_layouts_match_header():
    import ennemi_amd._native as nat
    # enm_instance: 2*int32 + int64 + 4*int32 + int64 + 16*int64
    assert ctypes.sizeof(nat.EnmInstance) == 8 + 8 + 16 + 8 + 16 * 8
    # enm_stats: 6*uint64 + 2*uint32 + 6*float + float (+ tail padding to 8)
    assert ctypes.sizeof(nat.EnmStats) == 48 + 8 + 24 + 4 + 4
    header = open(os.path.join(ROOT, "include", "ennemi_hip.h")).read()
    for name, val in (("F_FP64", 0x1), ("F_BRUTE", 0x2), ("F_GRAPH", 0x4),
                      ("F_PER_INSTANCE", 0x8), ("F_TIMING", 0x10)):
        assert getattr(nat, name) == val
        assert re.search(rf"#define ENM_{name}\s+0x{val:x}u", header), name
    for name, val in (("KSG", 0), ("CMI", 1), ("ROSS", 2), ("CROSS", 3), ("KL", 4)):
        assert getattr(nat, f"KIND_{name}") == val
        assert re.search(rf"#define ENM_KIND_{name}\s+{val}\b", header), name


def test_init_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import ennemi_amd._native as nat
    lib = nat.load_library()
    ptr = ctypes.c_void_p()
    rc = lib.enm_init(0, ctypes.byref(ptr))
    assert rc != 0 and not ptr.value
    assert lib.enm_last_error(None)
    with pytest.raises(nat.NativeUnavailable):
        nat.Context(0)


# ---------------------------------------------------------------------------
# Validation messages (ennemi tests/unit/test_driver.py:26-37)
# ---------------------------------------------------------------------------

X_Y_DIFFERENT_LENGTH_MSG = "x and y must have same length"
X_COND_DIFFERENT_LENGTH_MSG = "x and cond must have same length"
X_WRONG_DIMENSION_MSG = "x must be one- or two-dimensional"
Y_WRONG_DIMENSION_MSG = "y must be one-dimensional"
COND_WRONG_DIMENSION_MSG = "cond must be one- or two-dimensional"
MASK_WRONG_DIMENSION_MSG = "mask must be one-dimensional"
K_TOO_LARGE_MSG = "k must be smaller than number of observations (after lag and mask)"
K_NEGATIVE_MSG = "k must be greater than zero"
TOO_LARGE_LAG_MSG = "lag is too large, no observations left"
INVALID_MASK_LENGTH_MSG = "mask length does not match input length"
INVALID_MASK_TYPE_MSG = "mask must contain only booleans"
NANS_LEFT_MSG = "input contains NaNs (after applying the mask), pass drop_nan=True to ignore"


def _raises(exc, msg, fn, *a, **kw):
    with pytest.raises(exc) as ei:
        fn(*a, **kw)
    if msg is not None:
        assert str(ei.value) == msg


def test_validation_messages():
    from ennemi_amd import estimate_entropy, estimate_mi, pairwise_mi
    z = np.zeros(20)
    _raises(ValueError, K_TOO_LARGE_MSG, estimate_entropy, np.zeros(3), k=3)
    _raises(ValueError, K_NEGATIVE_MSG, estimate_entropy, z, k=0)
    _raises(TypeError, None, estimate_entropy, z, k=2.71828)
    _raises(ValueError, X_WRONG_DIMENSION_MSG, estimate_entropy, np.zeros((20, 2, 1)))
    _raises(ValueError, INVALID_MASK_LENGTH_MSG, estimate_entropy, z, mask=np.ones(5, bool))
    _raises(TypeError, INVALID_MASK_TYPE_MSG, estimate_entropy, z, mask=np.ones(20))
    _raises(ValueError, NANS_LEFT_MSG, estimate_entropy, np.full(20, np.nan))
    _raises(ValueError, X_COND_DIFFERENT_LENGTH_MSG, estimate_entropy, z, cond=np.zeros(5))

    _raises(ValueError, X_Y_DIFFERENT_LENGTH_MSG, estimate_mi, z, np.zeros(10))
    _raises(ValueError, X_WRONG_DIMENSION_MSG, estimate_mi, z, np.zeros((20, 2, 2)))
    _raises(ValueError, Y_WRONG_DIMENSION_MSG, estimate_mi, np.zeros((20, 2)), z)
    _raises(ValueError, COND_WRONG_DIMENSION_MSG, estimate_mi, z, z, cond=np.zeros((20, 2, 2)))
    _raises(ValueError, X_COND_DIFFERENT_LENGTH_MSG, estimate_mi, z, z, cond=np.zeros(19))
    _raises(ValueError, MASK_WRONG_DIMENSION_MSG, estimate_mi, z, z, mask=np.ones((20, 2), bool))
    _raises(ValueError, INVALID_MASK_LENGTH_MSG, estimate_mi, z, z, mask=np.ones(19, bool))
    _raises(TypeError, INVALID_MASK_TYPE_MSG, estimate_mi, z, z, mask=np.ones(20))
    _raises(ValueError, K_NEGATIVE_MSG, estimate_mi, z, z, k=-1)
    _raises(TypeError, None, estimate_mi, z, z, k=3.0)
    _raises(ValueError, TOO_LARGE_LAG_MSG, estimate_mi, z, z, lag=20)
    _raises(ValueError, TOO_LARGE_LAG_MSG, estimate_mi, z, z, lag=[-10, 10])
    _raises(ValueError, K_TOO_LARGE_MSG, estimate_mi, z, z, lag=16, k=4)
    m = np.ones(20, bool)
    m[2:] = False
    _raises(ValueError, K_TOO_LARGE_MSG, estimate_mi, z, z, mask=m, k=3)
    xn = np.arange(20.0)
    xn[3] = np.nan
    _raises(ValueError, NANS_LEFT_MSG, estimate_mi, np.arange(20.0), xn)
    _raises(ValueError, NANS_LEFT_MSG, estimate_mi, np.arange(20.0), np.arange(20.0), cond=xn)
    _raises(ValueError, None, pairwise_mi, np.zeros((20, 3)), cond=z,
            discrete=[True, False, False])
    # One variable only: the trivial result, without any estimation.
    assert np.isnan(pairwise_mi(np.zeros(10))).all()


def test_normalize_mi_matches_reference_formula():
    from ennemi_amd import normalize_mi
    vals = np.array([-0.5, 0.0, 0.1, 1.0, 3.0])
    np.testing.assert_array_equal(normalize_mi(vals), orc.normalize_mi(vals))
    assert normalize_mi(0.5) == orc.normalize_mi(0.5)


# ---------------------------------------------------------------------------
# The host planner feeds the GPU the reference's exact float64 inputs
# ---------------------------------------------------------------------------

def _batch_columns(batch, i):
    kind, k, n, dx, dy, dz, n_segs, seg_index, cols = batch._inst[i]
    return [batch._cols[c] for c in cols], (dx, dy, dz)


def _oracle_params(g, name):
    from tests.test_oracle import mi_kwargs
    kw = mi_kwargs(g, name)
    y = g[f"{name}/y"]
    x = g[f"{name}/x"]
    lag = np.atleast_1d(g[f"{name}/lag"])
    cond = kw.get("cond")
    ncond = 1
    if cond is not None:
        cond = np.column_stack((np.asarray(cond),))
        ncond = cond.shape[1]
    cond_lag = np.broadcast_to(np.column_stack((kw.get("cond_lag", 0),)), (lag.shape[0], ncond))
    min_lag = min(np.min(lag), np.min(cond_lag))
    max_lag = max(np.max(lag), np.max(cond_lag))
    nvar = 1 if x.ndim == 1 else x.shape[1]
    out = []
    for li in range(len(lag)):
        for v in range(nvar):
            out.append(((x if x.ndim == 1 else x[:, v]), y, lag[li], max_lag, min_lag, kw["k"],
                        kw.get("mask"), cond, cond_lag[li], kw["discrete_x"], kw["discrete_y"],
                        kw["preprocess"], kw["drop_nan"]))
    return kw, out


@pytest.mark.parametrize("name", ["mi_lags", "mi_cond1", "mi_cond2", "mi_mask", "mi_nan",
                                  "mi_nan_uncond", "mi_nopre", "mi_ties_pre", "mi_k1_norm"])
def test_planner_inputs_bitwise_equal_oracle(golden_driver, name):
    from ennemi_amd._driver import plan_mi
    g = golden_driver
    kw, params = _oracle_params(g, name)
    kw2 = {k: v for k, v in kw.items() if k != "normalize"}
    batch, indices = plan_mi(g[f"{name}/y"], g[f"{name}/x"], g[f"{name}/lag"], **kw2)
    assert len(batch) == len(params)
    for i, p in enumerate(params):
        _, (xs, ys, zs) = orc.lagged_task(p)
        cols, (dx, dy, dz) = _batch_columns(batch, i)
        want = [xs, ys] + ([] if zs is None else [zs[:, c] for c in range(zs.shape[1])])
        assert len(cols) == len(want)
        for a, b in zip(cols, want):
            assert np.array_equal(a, b), f"{name} task {i}"
            assert a.tobytes() == np.ascontiguousarray(b).tobytes()


def test_pairwise_shared_roles_bitwise_equal_oracle(golden_driver):
    from ennemi_amd import _driver as drv
    g = golden_driver
    data = g["pairwise/data"]
    captured = []
    orig = drv._add_mi_task

    def capture(tasks, xs, ys, zs, k, dx, dy):
        captured.append((xs, ys))
        tasks.host(0.0)
    drv._add_mi_task = capture
    try:
        drv.pairwise_mi(data, k=3)
    finally:
        drv._add_mi_task = orig
    nvar = data.shape[1]
    pairs = [(i, j) for i in range(nvar) for j in range(i + 1, nvar)]
    for (i, j), (xs, ys) in zip(pairs, captured):
        p = (data[:, i], data[:, j], 0, 0, 0, 3, None, None, np.asarray(0), False, False, True, False)
        _, (ox, oy, _) = orc.lagged_task(p)
        assert xs.tobytes() == ox.tobytes() and ys.tobytes() == oy.tobytes(), (i, j)
    # The x role of a column is one shared object across its pairs (shared sort).
    assert captured[0][0] is captured[1][0]


def test_noise_slices_equal_fresh_generators():
    from ennemi_amd._driver import NOISE_SEED, _Noise
    nz = _Noise()
    a = nz.take(0, (700,))
    b = nz.take(700, (300, 2))
    r = np.random.default_rng(NOISE_SEED)
    assert np.array_equal(a, r.normal(0.0, 1e-10, 700))
    assert np.array_equal(b, r.normal(0.0, 1e-10, (300, 2)))
    c = nz.take(5000, (10,))  # growth regenerates the same stream
    assert np.array_equal(c, np.random.default_rng(NOISE_SEED).normal(0.0, 1e-10, 5010)[5000:])


def test_constant_variable_warns_and_skips_noise():
    from ennemi_amd._driver import plan_mi, PREPROCESS_CONSTANT_DATA_WARNING
    y = np.random.default_rng(0).normal(size=50)
    x = np.ones(50)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        batch, _ = plan_mi(y, x, k=3)
    assert any(str(m.message) == PREPROCESS_CONSTANT_DATA_WARNING for m in w)
    cols, _ = _batch_columns(batch, 0)
    p = (x, y, 0, 0, 0, 3, None, None, np.zeros(1, int), False, False, True, False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        _, (ox, oy, _) = orc.lagged_task(p)
    assert cols[0].tobytes() == ox.tobytes() and cols[1].tobytes() == oy.tobytes()


def test_ross_grouping_is_class_ordered():
    from ennemi_amd._driver import _group_by_class
    y = np.array([3, 1, 3, 2, 1, 1])
    x = np.arange(6.0)
    values, counts, grouped = _group_by_class(y, [x])
    assert list(values) == [1, 2, 3] and list(counts) == [3, 1, 2]
    assert list(grouped[0]) == [1.0, 4.0, 5.0, 3.0, 0.0, 2.0]
