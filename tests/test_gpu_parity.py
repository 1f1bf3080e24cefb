"""Parity of the HIP path with the oracle (GPU).

Every test calls the product through the C ABI (ennemi_amd -> ctypes ->
libennemi_hip.so) and checks against the oracle on the same seeded inputs:
k-th-neighbour radii and ball counts bit-exact, estimates within 1e-12
absolute (the FP64 reference's own summation order differs from the
device's fixed-order tree; both are float64).
"""

import zlib

import numpy as np
import pytest

from oracle import ennemi_oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _nat():
    import ennemi_amd._native as nat
    return nat


def _modes():
    nat = _nat()
    return {
        "f32": 0,
        "fp64": nat.F_FP64,
        "brute": nat.F_BRUTE,
        "brute_fp64": nat.F_BRUTE | nat.F_FP64,
        "per_instance_fp64": nat.F_PER_INSTANCE | nat.F_FP64,
    }


MODE_NAMES = ["f32", "fp64", "brute", "brute_fp64", "per_instance_fp64"]


def _close(a, b):
    a = float(a)
    b = float(b)
    if np.isnan(a) or np.isnan(b):
        return np.isnan(a) and np.isnan(b)
    if np.isinf(a) or np.isinf(b):
        return a == b
    return abs(a - b) <= TOL


def _batch_one(kind, k, x, y=(), z=(), classes=None, flags=0):
    from ennemi_amd._batch import Batch
    b = Batch()
    b.add(kind, k, x, y, z, class_sizes=classes)
    vals, eps, cnts = b.run(flags=flags, export=True)
    return vals[0], eps[0], cnts[0]


def _cols(a):
    a = np.column_stack((a,))
    return [np.ascontiguousarray(a[:, c]) for c in range(a.shape[1])]


# ---------------------------------------------------------------------------
# Radii against cKDTree golden vectors (ties, duplicates, heavy tails)
# ---------------------------------------------------------------------------

KNN_SETS = ["normal2", "rounded2", "rounded_noise2", "normal1", "rounded1", "normal4",
            "heavy3", "dup5"]


@pytest.mark.parametrize("mode", MODE_NAMES)
@pytest.mark.parametrize("name", KNN_SETS)
def test_radii_match_golden(golden_knn, name, mode):
    nat = _nat()
    pts = golden_knn[f"{name}/pts"]
    for k in (1, 3, 7):
        v, eps, _ = _batch_one(nat.KIND_KL, k, _cols(pts), flags=_modes()[mode])
        np.testing.assert_array_equal(eps, golden_knn[f"{name}/k{k}/eps"], err_msg=f"{name} k={k}")
        ref = orc.estimate_single_entropy(pts, k)
        assert _close(v, ref), (v, ref)


# ---------------------------------------------------------------------------
# Estimators against the oracle: radii, counts, values
# ---------------------------------------------------------------------------

def _data(seed, n, d, rounding=None, scale=1.0):
    rng = np.random.default_rng(seed)
    a = rng.multivariate_normal(np.zeros(d), np.eye(d) * 0.6 + 0.4, size=n) * scale
    if rounding is not None:
        a = np.round(a, rounding)
    return a


CASES = [
    # (name, kind, n, dx, dy, dz, k, rounding)
    ("ksg", "KSG", 1500, 1, 1, 0, 3, None),
    ("ksg_k1", "KSG", 900, 1, 1, 0, 1, None),
    ("ksg_ties", "KSG", 1200, 1, 1, 0, 3, 1),
    ("ksg_2x1", "KSG", 1000, 2, 1, 0, 4, None),
    ("ksg_k20", "KSG", 1100, 1, 1, 0, 20, None),
    ("cmi_z1", "CMI", 1200, 1, 1, 1, 3, None),
    ("cmi_z2", "CMI", 1000, 1, 1, 2, 5, None),
    ("cmi_z3_ties", "CMI", 1000, 1, 1, 3, 3, 1),
    ("cmi_z7", "CMI", 700, 1, 1, 7, 3, None),
    ("kl_1", "KL", 1000, 1, 0, 0, 3, None),
    ("kl_4", "KL", 1000, 4, 0, 0, 3, None),
    ("kl_ties", "KL", 800, 2, 0, 0, 2, 1),
]


def _run_case(case, mode):
    nat = _nat()
    name, kind, n, dx, dy, dz, k, rnd = case
    a = _data(zlib.crc32(name.encode()), n, dx + dy + dz, rnd)
    X, Y, Z = a[:, :dx], a[:, dx:dx + dy], a[:, dx + dy:]
    flags = _modes()[mode]
    if kind == "KSG":
        v, eps, c = _batch_one(nat.KIND_KSG, k, _cols(X), _cols(Y), flags=flags)
        w, ref = orc.estimate_single_mi(X, Y, k, detail=True)
        np.testing.assert_array_equal(eps, ref["eps"])
        np.testing.assert_array_equal(c[0], ref["nx"])
        np.testing.assert_array_equal(c[1], ref["ny"])
    elif kind == "CMI":
        v, eps, c = _batch_one(nat.KIND_CMI, k, _cols(X), _cols(Y), _cols(Z), flags=flags)
        w, ref = orc.estimate_conditional_mi(X, Y, Z, k, detail=True)
        np.testing.assert_array_equal(eps, ref["eps"])
        for s, key in enumerate(("nxz", "nyz", "nz")):
            np.testing.assert_array_equal(c[s], ref[key], err_msg=key)
    else:
        v, eps, c = _batch_one(nat.KIND_KL, k, _cols(X), flags=flags)
        w, ref = orc.estimate_single_entropy(X, k, detail=True)
        np.testing.assert_array_equal(eps, ref["eps"])
    assert _close(v, w), (v, w)


@pytest.mark.parametrize("mode", MODE_NAMES)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_estimator_parity(case, mode):
    _run_case(case, mode)


@pytest.mark.parametrize("mode", MODE_NAMES)
@pytest.mark.parametrize("dx", [1, 3])
def test_ross_parity(mode, dx):
    from ennemi_amd import _estimators as est
    rng = np.random.default_rng(11 + dx)
    n = 1500
    y = rng.choice(5, size=n, p=[0.4, 0.3, 0.2, 0.098, 0.002])
    y[:2] = 7  # a class with fewer than k+1 points: radius inf, count n
    x = rng.normal(size=(n, dx)) + y[:, None] * 0.3
    if dx == 1:
        x = x[:, 0]
    nat = _nat()
    values, counts, order, grouped = est._grouped(y, _cols(x))
    v, eps, c = _batch_one(nat.KIND_ROSS, 3, grouped, classes=counts, flags=_modes()[mode])
    w, ref = orc.estimate_semidiscrete_mi(x, y, 3, detail=True)
    inv = np.empty_like(order)
    inv[order] = np.arange(n)
    np.testing.assert_array_equal(eps[inv], ref["eps"])
    np.testing.assert_array_equal(c[0][inv], ref["n_full"].astype(np.int64))
    assert np.isinf(ref["eps"][:2]).all()
    assert _close(v, w), (v, w)


@pytest.mark.parametrize("mode", MODE_NAMES)
@pytest.mark.parametrize("dz", [1, 2])
def test_conditional_ross_parity(mode, dz):
    nat = _nat()
    from ennemi_amd import _estimators as est
    rng = np.random.default_rng(31 + dz)
    n = 1200
    y = rng.choice(4, size=n)
    x = rng.normal(size=n) + y * 0.5
    z = rng.normal(size=(n, dz)) + x[:, None] * 0.3
    values, counts, order, grouped = est._grouped(y, _cols(x) + _cols(z))
    v, eps, c = _batch_one(nat.KIND_CROSS, 3, grouped[:1], z=grouped[1:], classes=counts,
                           flags=_modes()[mode])
    w, ref = orc.estimate_conditional_semidiscrete_mi(x, y, z, 3, detail=True)
    inv = np.empty_like(order)
    inv[order] = np.arange(n)
    np.testing.assert_array_equal(eps[inv], ref["eps"])
    for s, key in enumerate(("nxz", "nyz", "nz")):
        np.testing.assert_array_equal(c[s][inv], ref[key].astype(np.int64), err_msg=key)
    assert _close(v, w), (v, w)


# ---------------------------------------------------------------------------
# Edge cases
# ---------------------------------------------------------------------------

def test_duplicates_give_minus_inf_and_nan():
    from ennemi_amd import _estimators as est
    rng = np.random.default_rng(3)
    x = np.round(rng.normal(size=400), 1)
    y = np.round(rng.normal(size=400), 1)
    z = np.round(rng.normal(size=400), 0)
    v = est.estimate_single_mi(x, y, 3)
    assert v == orc.estimate_single_mi(x, y, 3) == -np.inf
    with np.errstate(invalid="ignore"):
        w = orc.estimate_conditional_mi(x, y, z, 3)
    assert np.isnan(est.estimate_conditional_mi(x, y, z, 3)) == np.isnan(w)
    assert est.estimate_single_entropy(x, 3) == -np.inf


def test_minimum_points_and_magnitudes():
    from ennemi_amd import _estimators as est
    rng = np.random.default_rng(5)
    for scale in (1e-30, 1.0, 1e12, 1e35, 1e200):
        x = rng.normal(size=(5, 2)) * scale
        v, d = est.estimate_single_entropy(x, 4, detail=True)
        w, r = orc.estimate_single_entropy(x, 4, detail=True)
        np.testing.assert_array_equal(d["eps"], r["eps"])
        a = rng.normal(size=(600, 3)) * scale
        v, d = est.estimate_conditional_mi(a[:, 0], a[:, 1], a[:, 2], 3, detail=True)
        w, r = orc.estimate_conditional_mi(a[:, 0], a[:, 1], a[:, 2], 3, detail=True)
        np.testing.assert_array_equal(d["eps"], r["eps"])
        for key in ("nxz", "nyz", "nz"):
            np.testing.assert_array_equal(d[key], r[key])
        assert _close(v, w), (scale, v, w)


def test_modes_agree_bitwise():
    nat = _nat()
    from ennemi_amd._batch import Batch
    rng = np.random.default_rng(9)
    b = Batch()
    for t in range(12):
        n = 500 + 37 * t
        a = rng.normal(size=(n, 4))
        if t % 3 == 0:
            b.add(nat.KIND_KSG, 3, [a[:, 0]], [a[:, 1]])
        elif t % 3 == 1:
            b.add(nat.KIND_CMI, 3, [a[:, 0]], [a[:, 1]], [a[:, 2], a[:, 3]])
        else:
            b.add(nat.KIND_KL, 2, [a[:, 0], a[:, 1], a[:, 2]])
    ref = b.run(flags=nat.F_FP64 | nat.F_BRUTE | nat.F_PER_INSTANCE, export=True)
    for flags in (0, nat.F_FP64, nat.F_BRUTE, nat.F_PER_INSTANCE):
        got = b.run(flags=flags, export=True)
        np.testing.assert_array_equal(got[0], ref[0])
        for e1, e2 in zip(got[1], ref[1]):
            np.testing.assert_array_equal(e1, e2)
        for c1, c2 in zip(got[2], ref[2]):
            np.testing.assert_array_equal(c1, c2)


def test_graph_replays_match_eager():
    nat = _nat()
    from ennemi_amd._batch import Batch
    import torch
    rng = np.random.default_rng(12)
    b = Batch()
    for t in range(6):
        a = rng.normal(size=(800 + 100 * t, 3))
        b.add(nat.KIND_CMI, 3, [a[:, 0]], [a[:, 1]], [a[:, 2]])
    eager = b.run()
    db = b.device(flags=nat.F_GRAPH)
    for rep in range(4):
        db.launch()
        got = db.results()
        np.testing.assert_array_equal(got, eager)
        # Eager work and synchronisation between replays.
        torch.ones(1000, device="cuda").sum().item()
    assert db.stats()["graph_replayed"] == 1


def test_large_ksg_against_kdtree():
    # N = 100,000 at the headline shape: radii and counts vs cKDTree.
    nat = _nat()
    rng = np.random.default_rng(101)
    d = rng.multivariate_normal([0, 0], [[1, 0.6], [0.6, 1]], size=100_000)
    v, eps, c = _batch_one(nat.KIND_KSG, 3, [d[:, 0]], [d[:, 1]])
    w, ref = orc.estimate_single_mi(d[:, 0], d[:, 1], 3, backend="kdtree", detail=True)
    np.testing.assert_array_equal(eps, ref["eps"])
    np.testing.assert_array_equal(c[0], ref["nx"])
    np.testing.assert_array_equal(c[1], ref["ny"])
    assert _close(v, w), (v, w)


# ---------------------------------------------------------------------------
# Sample sort of segments longer than one block-sort tile (2048 keys)
# ---------------------------------------------------------------------------

def _sample_positions(n):
    """Positions the sample sort draws its splitter samples from in a
    segment of n keys (the device's stride-and-hash rule)."""
    nb = min(1024, -(-n // 768))
    ns = 16 * nb if nb > 1 else 0
    if ns == 0:
        return np.zeros(0, dtype=np.int64)
    step = n // ns
    i = np.arange(ns, dtype=np.uint64)
    h = (i * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x85EBCA77)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(13)
    return ((i * np.uint64(n)) // np.uint64(ns) + h % np.uint64(step)).astype(np.int64)


def _ksg_vs_kdtree(x, y, k, flags=0):
    nat = _nat()
    v, eps, c = _batch_one(nat.KIND_KSG, k, [x], [y], flags=flags)
    w, ref = orc.estimate_single_mi(x, y, k, backend="kdtree", detail=True)
    np.testing.assert_array_equal(eps, ref["eps"])
    np.testing.assert_array_equal(c[0], ref["nx"])
    np.testing.assert_array_equal(c[1], ref["ny"])
    assert _close(v, w), (v, w)


@pytest.mark.parametrize("n", [2049, 4097, 9000, 30011])
@pytest.mark.parametrize("rounding", [None, 1, 0])
def test_sample_sorted_sizes_and_ties(n, rounding):
    # Rounding to 0 decimals leaves a handful of distinct keys: the buckets
    # are cut inside runs of equal values by the point index.
    a = _data(7000 + n, n, 2, rounding)
    _ksg_vs_kdtree(np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1]), 3)


def test_sample_sort_oversized_bucket():
    # Every splitter sample holds a key above all others, so one bucket gets
    # ~97 % of the segment (far above one LDS run): the block merge-sort
    # fallback must give the same order.
    n = 20000
    rng = np.random.default_rng(99)
    x = rng.normal(size=n)
    pos = _sample_positions(n)
    assert len(pos) == 16 * 27 and len(np.unique(pos)) == len(pos) and pos.max() < n
    x[pos] = 1e3 + np.arange(len(pos))
    y = rng.normal(size=n)
    _ksg_vs_kdtree(x, y, 3)
    # Both columns adversarial, FP64 mode.
    y[pos] = -1e3 - np.arange(len(pos))
    _ksg_vs_kdtree(x, y, 3, flags=_nat().F_FP64)


def test_sample_sorted_class_segments():
    # Ross MI with classes from 3 to 9000 points: class segments of every
    # size sorted together (one bucket, several buckets, many runs).
    from ennemi_amd import _estimators as est
    nat = _nat()
    sizes = [3, 700, 2049, 9000, 5000, 40]
    rng = np.random.default_rng(5)
    y = np.repeat(np.arange(len(sizes)), sizes)
    rng.shuffle(y)
    x = rng.normal(size=len(y)) + 0.3 * y
    values, counts, order, grouped = est._grouped(y, [x])
    v, eps, c = _batch_one(nat.KIND_ROSS, 3, grouped, classes=counts)
    w, ref = orc.estimate_semidiscrete_mi(x, y, 3, detail=True)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(y))
    np.testing.assert_array_equal(eps[inv], ref["eps"])
    np.testing.assert_array_equal(c[0][inv], ref["n_full"].astype(np.int64))
    assert _close(v, w), (v, w)


def test_sample_sorted_modes_graphs_and_exports_agree():
    # Instances long enough for the sample sort (segments above 2048 keys):
    # every mode bitwise equal to FP64 brute force per instance, graph
    # replays equal to eager runs, and the values the same with and without
    # exported radii (without, no radius is kept in kNN order).
    nat = _nat()
    from ennemi_amd._batch import Batch
    import torch
    rng = np.random.default_rng(21)
    b = Batch()
    for t in range(6):
        n = 2500 + 1700 * t
        a = rng.normal(size=(n, 4))
        if t % 3 == 0:
            b.add(nat.KIND_KSG, 3, [a[:, 0]], [a[:, 1]])
        elif t % 3 == 1:
            b.add(nat.KIND_CMI, 3, [a[:, 0]], [a[:, 1]], [a[:, 2]])
        else:
            b.add(nat.KIND_KL, 3, [a[:, 0], a[:, 1]])
    ref = b.run(flags=nat.F_FP64 | nat.F_BRUTE | nat.F_PER_INSTANCE, export=True)
    for flags in (0, nat.F_FP64, nat.F_PER_INSTANCE | nat.F_FP64):
        got = b.run(flags=flags, export=True)
        np.testing.assert_array_equal(got[0], ref[0])
        for e1, e2 in zip(got[1], ref[1]):
            np.testing.assert_array_equal(e1, e2)
        for c1, c2 in zip(got[2], ref[2]):
            np.testing.assert_array_equal(c1, c2)
        np.testing.assert_array_equal(b.run(flags=flags), ref[0])
    db = b.device(flags=nat.F_GRAPH)
    for rep in range(3):
        db.launch()
        np.testing.assert_array_equal(db.results(), ref[0])
        torch.ones(1000, device="cuda").sum().item()
    assert db.stats()["graph_replayed"] == 1


# ---------------------------------------------------------------------------
# Tuning invariance: every shape the autotuner may pick gives the same bits
# ---------------------------------------------------------------------------

AUTOTUNE_STRIPS = [0, 256, 512, 1024, 2048, 4096, 8192]
AUTOTUNE_WAVES = [8, 16, 32]


def _mixed_tuning_batch():
    from ennemi_amd._batch import Batch
    from ennemi_amd import _estimators as est
    nat = _nat()
    rng = np.random.default_rng(2024)
    b = Batch()
    a = _data(31, 9000, 4, rounding=None)
    t = _data(32, 6000, 3, rounding=1)
    b.add(nat.KIND_KSG, 3, _cols(a[:, :1]), _cols(a[:, 1:2]))
    b.add(nat.KIND_KSG, 4, _cols(t[:, :2]), _cols(t[:, 2:]))
    b.add(nat.KIND_CMI, 3, _cols(a[:, :1]), _cols(a[:, 1:2]), _cols(a[:, 2:]))
    b.add(nat.KIND_CMI, 5, _cols(t[:, :1]), _cols(t[:, 1:2]), _cols(t[:, 2:]))
    b.add(nat.KIND_KL, 3, _cols(a[:, :3]))
    b.add(nat.KIND_KL, 40, _cols(t[:, :2]))
    y = rng.choice(4, size=7000, p=[0.5, 0.3, 0.15, 0.05])
    x = rng.normal(size=(7000, 2)) + y[:, None]
    values, counts, order, grouped = est._grouped(y, _cols(x))
    b.add(nat.KIND_ROSS, 3, grouped, class_sizes=counts)
    z = [np.ascontiguousarray(np.sin(g)) for g in grouped]
    b.add(nat.KIND_CROSS, 3, grouped[:1], z=z, class_sizes=counts)
    return b


AUTOTUNE_SLOTS = [0, 8, 16, 32]
AUTOTUNE_TICKETS = [0, 1, 2, 4, 64]


def _shapes():
    """Every strip x waves pair the autotuner sweeps, and list slots and
    items per ticket on top of the default."""
    for strip in AUTOTUNE_STRIPS:
        for waves in AUTOTUNE_WAVES:
            yield (strip, waves, 0, 0)
    for slots in AUTOTUNE_SLOTS:
        for ticket in AUTOTUNE_TICKETS:
            yield (512, 16, slots, ticket)


@pytest.mark.parametrize("flags_name", ["f32", "fp64"])
def test_every_autotuner_shape_gives_identical_results(flags_name):
    nat = _nat()
    flags = _modes()[flags_name]
    ctx = nat.context()
    old = ctx.get_tuning()
    b = _mixed_tuning_batch()
    try:
        ctx.set_tuning(1024, 32)
        ref_vals, ref_eps, ref_cnt = b.run(flags=flags, export=True)
        for shape in _shapes():
            ctx.set_tuning_ex(*shape)
            vals, eps, cnt = b.run(flags=flags, export=True)
            np.testing.assert_array_equal(vals, ref_vals, err_msg=f"shape {shape}")
            for i in range(len(vals)):
                np.testing.assert_array_equal(eps[i], ref_eps[i], err_msg=f"eps {i} shape {shape}")
                np.testing.assert_array_equal(cnt[i], ref_cnt[i], err_msg=f"counts {i} shape {shape}")
    finally:
        ctx.set_tuning(*old)


def test_autotune_keeps_the_winner_per_shape():
    """enm_autotune keeps its winner for the batch's shapes (joint dimension
    class, k+1 class, log2 n) and the results do not change."""
    nat = _nat()
    ctx = nat.context()
    old = ctx.get_tuning()
    b = _mixed_tuning_batch()
    try:
        ctx.set_tuning(1024, 32)
        ref = b.run(export=True)
        b.device(0).autotune()
        win_a = ctx.tuning_for(2, 3, 9000)             # the first KSG instance's shape
        assert ctx.tuning_for(11, 11, 3) == (1024, 32, 0, 0)   # other shapes: the default
        # A second batch of another shape keeps its own winner; the first
        # batch's shapes keep theirs.
        from ennemi_amd._batch import Batch
        other = Batch()
        other.add(nat.KIND_KL, 3, _cols(_data(77, 40_000, 3)))
        other.device(0).autotune()
        assert ctx.tuning_for(2, 3, 9000) == win_a
        assert ctx.tuning_for(11, 11, 3) == (1024, 32, 0, 0)
        ctx.set_tuning(256, 8)
        assert ctx.tuning_for(2, 3, 9000) == (256, 8, 0, 0)   # set_tuning forgets the kept shapes
        got = b.run(export=True)
        np.testing.assert_array_equal(got[0], ref[0])
        for i in range(len(got[0])):
            np.testing.assert_array_equal(got[1][i], ref[1][i])
            np.testing.assert_array_equal(got[2][i], ref[2][i])
    finally:
        ctx.set_tuning(*old)


def test_default_strips_widen_for_three_or_more_columns():
    """A fresh context gives instances of three or more joint columns strips
    of 4096 points and two-column ones 1024 until a tuning is set; results
    are those of one strip size for all."""
    nat = _nat()
    fresh = nat.Context(0)
    try:
        assert fresh.tuning_for(2, 3, 100_000) == (1024, 32, 0, 0)
        assert fresh.tuning_for(3, 3, 100_000) == (4096, 32, 0, 0)
        assert fresh.tuning_for(4, 6, 2_000_000) == (4096, 32, 0, 0)
        b = _mixed_tuning_batch()
        shared = nat.context()
        old = shared.get_tuning()
        try:
            shared.set_tuning(1024, 32)
            ref = b.run(export=True)
        finally:
            shared.set_tuning(*old)
        db = b.device(0, export=True)
        db.ctx = fresh   # the same batch, launched through the fresh context
        db.launch()
        got = db.results()
        np.testing.assert_array_equal(got[0], ref[0])
        for i in range(len(got[0])):
            np.testing.assert_array_equal(got[1][i], ref[1][i])
            np.testing.assert_array_equal(got[2][i], ref[2][i])
        fresh.set_tuning(1024, 32)
        assert fresh.tuning_for(3, 3, 100_000) == (1024, 32, 0, 0)
    finally:
        fresh.lib.enm_shutdown(fresh.ctx)


def test_plans_are_reused_across_arenas_of_one_shape():
    """Two batches of the same shape (same sizes and column offsets) but
    different data and device buffers share one cached plan; each must get
    its own results, eagerly and as graph replays (a graph captured with the
    first batch's pointers is captured again for the second)."""
    from ennemi_amd._batch import Batch
    nat = _nat()

    def make(seed):
        a = _data(seed, 5000, 3)
        b = Batch()
        b.add(nat.KIND_KSG, 3, _cols(a[:, :1]), _cols(a[:, 1:2]))
        b.add(nat.KIND_CMI, 4, _cols(a[:, :1]), _cols(a[:, 1:2]), _cols(a[:, 2:]))
        return b, a

    (b1, a1), (b2, a2) = make(501), make(502)
    w1 = [orc.estimate_single_mi(a1[:, :1], a1[:, 1:2], 3), orc.estimate_conditional_mi(a1[:, :1], a1[:, 1:2],
                                                                                          a1[:, 2:], 4)]
    w2 = [orc.estimate_single_mi(a2[:, :1], a2[:, 1:2], 3), orc.estimate_conditional_mi(a2[:, :1], a2[:, 1:2],
                                                                                          a2[:, 2:], 4)]
    for flags in (0, nat.F_GRAPH):
        d1, d2 = b1.device(flags), b2.device(flags)
        for _ in range(2):
            d1.launch()
            v1 = d1.results()
            d2.launch()
            v2 = d2.results()
            assert all(_close(v, w) for v, w in zip(v1, w1)), (flags, v1, w1)
            assert all(_close(v, w) for v, w in zip(v2, w2)), (flags, v2, w2)
