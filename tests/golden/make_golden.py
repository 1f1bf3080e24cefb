"""Write the golden vectors in this folder.

Run in the build container only (the reference is not present on the GPU
machines):  python tests/golden/make_golden.py [discrete]

Two kinds of vectors are written, both as small ``.npz`` fixtures:

* ``driver_cases.npz`` -- inputs and outputs of ennemi's public functions
  (estimate_mi, pairwise_mi, estimate_entropy, estimate_corr) and of its
  estimator-level functions, obtained by importing the reference package from
  its source tree.  These pin the composed driver path (lag windowing, masks,
  NaN drop, rescaling with the fixed-seed noise, task order, normalisation and
  the psi(0) broadcasting) bit for bit, which no closed form pins.
* ``discrete_cases.npz`` -- the discrete plug-in estimators (entropy, MI,
  conditional MI) on numeric, boolean, mixed-type and string values.
* ``knn_cases.npz`` -- per-point k-th-neighbour radii and ball counts from
  ``scipy.spatial.cKDTree`` called exactly as ennemi calls it, on inputs with
  exact ties, duplicates and negative radii.

Each case stores its inputs, so the tests never need the reference.
"""

from __future__ import annotations

import hashlib
import os
import sys
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = "/root/reference"


def _ref():
    sys.path.insert(0, REFERENCE)
    import ennemi  # noqa: F401  (the reference package, from its source tree)
    from ennemi import _entropy_estimators as est
    return ennemi, est


def driver_cases():
    ennemi, est = _ref()
    cases = {}

    def put(name, **arrays):
        for key, val in arrays.items():
            cases[f"{name}/{key}"] = np.asarray(val)

    rng = np.random.default_rng(1001)
    n = 400
    cov = np.array([[1.0, 0.6, 0.2], [0.6, 1.0, 0.3], [0.2, 0.3, 1.0]])
    base = rng.multivariate_normal([0, 0, 0], cov, size=n)
    y = base[:, 0].copy()
    x = np.column_stack([np.roll(base[:, 1], 2), base[:, 2] ** 3, rng.uniform(-1, 1, n)])
    z = rng.normal(size=(n, 2))

    lags = np.array([-3, 0, 2, 5])
    put("mi_lags", y=y, x=x, lag=lags, k=3,
        out=ennemi.estimate_mi(y, x, lag=lags, k=3))
    put("mi_k1_norm", y=y, x=x, lag=np.array([0, 1]), k=1,
        out=ennemi.estimate_mi(y, x, lag=[0, 1], k=1, normalize=True))
    put("corr", y=y, x=x, lag=np.array([-1, 1]), k=4,
        out=ennemi.estimate_corr(y, x, lag=[-1, 1], k=4))
    put("mi_cond1", y=y, x=x, lag=lags, k=3, cond=z[:, 0], cond_lag=2,
        out=ennemi.estimate_mi(y, x, lag=lags, k=3, cond=z[:, 0], cond_lag=2))
    cond_lag2 = np.array([[0, 1], [1, 2], [2, 0], [4, 4]])
    put("mi_cond2", y=y, x=x, lag=lags, k=3, cond=z, cond_lag=cond_lag2,
        out=ennemi.estimate_mi(y, x, lag=lags, k=3, cond=z, cond_lag=cond_lag2))

    mask = rng.uniform(size=n) > 0.3
    put("mi_mask", y=y, x=x, lag=lags, k=3, mask=mask,
        out=ennemi.estimate_mi(y, x, lag=lags, k=3, mask=mask))

    xn = x.copy()
    xn[rng.choice(n, 40, replace=False), 0] = np.nan
    xn[rng.choice(n, 10, replace=False), 2] = np.nan
    yn = y.copy()
    yn[rng.choice(n, 7, replace=False)] = np.nan
    zn = z.copy()
    zn[rng.choice(n, 5, replace=False), 1] = np.nan
    put("mi_nan", y=yn, x=xn, lag=lags, k=3, mask=mask, cond=zn,
        out=ennemi.estimate_mi(yn, xn, lag=lags, k=3, mask=mask, cond=zn, drop_nan=True))
    put("mi_nan_uncond", y=yn, x=xn, lag=lags, k=2,
        out=ennemi.estimate_mi(yn, xn, lag=lags, k=2, drop_nan=True))

    yd = (base[:, 0] > 0.3).astype(np.int64) + (base[:, 0] > -0.5)
    put("mi_discrete_y", y=yd, x=x, lag=lags, k=3,
        out=ennemi.estimate_mi(yd, x, lag=lags, k=3, discrete_y=True))
    xd = np.column_stack([(x[:, 0] > 0).astype(np.int64), np.round(x[:, 2] * 2)])
    put("mi_discrete_x", y=y, x=xd, lag=lags, k=3,
        out=ennemi.estimate_mi(y, xd, lag=lags, k=3, discrete_x=True))
    put("mi_discrete_y_cond", y=yd, x=x, lag=lags, k=3, cond=z[:, 0],
        out=ennemi.estimate_mi(yd, x, lag=lags, k=3, discrete_y=True, cond=z[:, 0]))
    put("mi_discrete_x_cond", y=y, x=xd, lag=lags, k=3, cond=z,
        out=ennemi.estimate_mi(y, xd, lag=lags, k=3, discrete_x=True, cond=z))
    put("mi_discrete_both", y=yd, x=xd, lag=lags, k=3,
        out=ennemi.estimate_mi(yd, xd, lag=lags, k=3, discrete_x=True, discrete_y=True))

    # Exact duplicates without preprocessing: radius 0, count 0, psi(0) = inf.
    xdup = np.round(x, 1)
    ydup = np.round(y, 1)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        put("mi_dup_nopre", y=ydup, x=xdup, lag=np.array([0]), k=3,
            out=ennemi.estimate_mi(ydup, xdup, k=3, preprocess=False))
        put("mi_dup_nopre_cond", y=ydup, x=xdup, lag=np.array([0]), k=3, cond=np.round(z, 1),
            out=ennemi.estimate_mi(ydup, xdup, k=3, preprocess=False, cond=np.round(z, 1)))
    put("mi_ties_pre", y=ydup, x=xdup, lag=np.array([0, 1]), k=3,
        out=ennemi.estimate_mi(ydup, xdup, lag=[0, 1], k=3))
    put("mi_nopre", y=y, x=x, lag=np.array([0, 2]), k=5,
        out=ennemi.estimate_mi(y, x, lag=[0, 2], k=5, preprocess=False))

    # pairwise
    m = 300
    data = np.column_stack([rng.standard_t(3, m), rng.uniform(-1, 1, m),
                            rng.lognormal(0, 1, m), np.zeros(m)])
    data[:, 3] = np.sin(data[:, 0]) + 0.3 * rng.normal(size=m)
    data_t = data.copy()
    data_t[:, 1] = np.round(data_t[:, 1], 2)
    put("pairwise", data=data_t, k=3, out=ennemi.pairwise_mi(data_t, k=3))
    put("pairwise_norm", data=data_t, k=2, out=ennemi.pairwise_mi(data_t, k=2, normalize=True))
    pmask = rng.uniform(size=m) > 0.2
    pc = rng.normal(size=m)
    put("pairwise_cond_mask", data=data_t, k=3, cond=pc, mask=pmask,
        out=ennemi.pairwise_mi(data_t, k=3, cond=pc, mask=pmask))
    pd_mix = data_t.copy()
    pd_mix[:, 1] = np.round(pd_mix[:, 1] * 2)
    disc = np.array([False, True, False, False])
    put("pairwise_mixed", data=pd_mix, k=3, discrete=disc,
        out=ennemi.pairwise_mi(pd_mix, k=3, discrete=disc))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        put("pairwise_nopre", data=data_t, k=3,
            out=ennemi.pairwise_mi(data_t, k=3, preprocess=False))
    pdn = data_t.copy()
    pdn[rng.choice(m, 15, replace=False), 2] = np.nan
    put("pairwise_nan", data=pdn, k=3, out=ennemi.pairwise_mi(pdn, k=3, drop_nan=True))

    # entropy
    ent = rng.multivariate_normal([0, 0, 0], cov, size=500)
    put("entropy_cols", x=ent, k=3, out=ennemi.estimate_entropy(ent, k=3))
    put("entropy_multidim", x=ent, k=4, out=ennemi.estimate_entropy(ent, k=4, multidim=True))
    emask = rng.uniform(size=500) > 0.25
    put("entropy_mask", x=ent, k=2, mask=emask, out=ennemi.estimate_entropy(ent, k=2, mask=emask))
    put("entropy_cond", x=ent[:, :2], k=3, cond=ent[:, 2],
        out=ennemi.estimate_entropy(ent[:, :2], k=3, cond=ent[:, 2]))
    put("entropy_cond_multidim", x=ent[:, :2], k=3, cond=ent[:, 2],
        out=ennemi.estimate_entropy(ent[:, :2], k=3, cond=ent[:, 2], multidim=True))
    entn = ent.copy()
    entn[rng.choice(500, 20, replace=False), 1] = np.nan
    put("entropy_nan_multidim", x=entn, k=3,
        out=ennemi.estimate_entropy(entn, k=3, multidim=True, drop_nan=True))
    put("entropy_nan_cols", x=entn, k=3, out=ennemi.estimate_entropy(entn, k=3, drop_nan=True))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        edup = np.round(ent[:, 0], 1)
        put("entropy_dup", x=edup, k=3, out=ennemi.estimate_entropy(edup, k=3))
    put("entropy_discrete", x=np.round(ent[:, 0]), k=3,
        out=ennemi.estimate_entropy(np.round(ent[:, 0]), k=3, discrete=True))

    # estimator level (no preprocessing, multi-dimensional x)
    g = rng.multivariate_normal([0, 0, 0, 0], np.eye(4) + 0.4, size=600)
    put("est_single_mi", x=g[:, 0], y=g[:, 1], k=3, out=est._estimate_single_mi(g[:, 0], g[:, 1], 3))
    put("est_cmi", x=g[:, 0], y=g[:, 1], cond=g[:, 2:], k=4,
        out=est._estimate_conditional_mi(g[:, 0], g[:, 1], g[:, 2:], 4))
    cls = (np.floor((g[:, 3] + 3) * 1.3)).astype(np.int64)
    put("est_ross_multid", x=g[:, :3], y=cls, k=3, out=est._estimate_semidiscrete_mi(g[:, :3], cls, 3))
    tiny = cls.copy()
    tiny[:2] = 99  # a class with fewer than k+1 points: radius +inf, count N
    put("est_ross_tiny_class", x=g[:, :2], y=tiny, k=3,
        out=est._estimate_semidiscrete_mi(g[:, :2], tiny, 3))
    put("est_cross", x=g[:, 0], y=cls, cond=g[:, 1:3], k=2,
        out=est._estimate_conditional_semidiscrete_mi(g[:, 0], cls, g[:, 1:3], 2))
    put("est_entropy_4d", x=g, k=3, out=est._estimate_single_entropy(g, 3))
    put("psi", x=np.arange(0, 50), out=np.array([est._psi(i) for i in range(0, 50)]))

    # CPU-runnable headline configuration: KSG at N=10,000, rho=0.6, k=3, normalize.
    rng_c1 = np.random.default_rng(101)
    d1 = rng_c1.multivariate_normal([0, 0], [[1, 0.6], [0.6, 1]], size=10_000)
    cases["c1/sha256"] = np.asarray(hashlib.sha256(d1.tobytes()).hexdigest())
    cases["c1/out"] = np.asarray(ennemi.estimate_mi(d1[:, 1], d1[:, 0], k=3, normalize=True))
    cases["c1/out_raw"] = np.asarray(ennemi.estimate_mi(d1[:, 1], d1[:, 0], k=3))
    return cases


def knn_cases():
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(2002)
    cases = {}
    inputs = {
        "normal2": rng.normal(size=(700, 2)),
        "rounded2": np.round(rng.normal(size=(700, 2)), 1),
        "rounded_noise2": np.round(rng.normal(size=(700, 2)), 1) + 1e-10 * rng.normal(size=(700, 2)),
        "normal1": rng.normal(size=(600, 1)),
        "rounded1": np.round(rng.normal(size=(600, 1)), 2),
        "normal4": rng.normal(size=(500, 4)),
        "heavy3": rng.standard_t(2, size=(500, 3)) * 50.0,
        "dup5": np.repeat(rng.normal(size=(100, 5)), 3, axis=0),
    }
    for name, pts in inputs.items():
        for k in (1, 3, 7):
            eps = cKDTree(pts).query(pts, k=[k + 1], p=np.inf)[0].flatten()
            r = eps - 1e-12
            cnt = cKDTree(pts).query_ball_point(pts, r, p=np.inf, return_length=True)
            cases[f"{name}/k{k}/eps"] = eps
            cases[f"{name}/k{k}/count"] = np.asarray(cnt, dtype=np.int64)
            # At the unshifted radius exact ties sit on the sphere, which pins
            # the inclusive (<=) boundary.
            cnt0 = cKDTree(pts).query_ball_point(pts, eps, p=np.inf, return_length=True)
            cases[f"{name}/k{k}/count_at_eps"] = np.asarray(cnt0, dtype=np.int64)
        cases[f"{name}/pts"] = pts
    return cases


def discrete_cases():
    """The discrete plug-in estimators: entropy, MI and conditional MI of
    discrete variables, through the public functions and the estimator-level
    ones, with numeric, mixed-type, boolean and string values."""
    ennemi, est = _ref()
    cases = {}

    def put(name, **arrays):
        for key, val in arrays.items():
            cases[f"{name}/{key}"] = np.asarray(val)

    rng = np.random.default_rng(3003)
    n = 2000
    z1 = rng.integers(0, 4, n)
    z2 = rng.integers(0, 3, n)
    x = (z1 + rng.integers(0, 3, n)) % 5
    y = (x + z2 + rng.integers(0, 2, n)) % 4
    xf = x * 0.5 - 1.0
    lags = np.array([0, 1, 3])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        put("disc_cond1", y=y, x=np.column_stack([x, xf]), lag=lags, cond=z1,
            out=ennemi.estimate_mi(y, np.column_stack([x, xf]), lag=lags, cond=z1,
                                   discrete_x=True, discrete_y=True))
        zz = np.column_stack([z1, z2])
        put("disc_cond2", y=y, x=x, lag=lags, cond=zz,
            out=ennemi.estimate_mi(y, x, lag=lags, cond=zz, discrete_x=True, discrete_y=True))
        put("est_cond_disc", x=xf, y=y, cond=zz,
            out=est._estimate_conditional_discrete_mi(xf, y, zz))
        put("est_disc_mi_mixed", x=xf, y=y, out=est._estimate_discrete_mi(xf, y))
        xb = x > 2
        put("est_disc_mi_bool", x=xb, y=y, out=est._estimate_discrete_mi(xb, y))
        xs = np.array(["a", "bb", "c", "dd", "e"])[x]
        put("est_disc_mi_str", x=xs, y=y, out=est._estimate_discrete_mi(xs, y))
        put("est_disc_entropy_2d", x=np.column_stack([x, z1, z2]),
            out=est._estimate_discrete_entropy(np.column_stack([x, z1, z2])))
        put("entropy_discrete_multidim", x=np.column_stack([xf, z2]),
            out=ennemi.estimate_entropy(np.column_stack([xf, z2]), discrete=True, multidim=True))
        put("entropy_discrete_cond", x=x, cond=z1,
            out=ennemi.estimate_entropy(x, cond=z1, discrete=True))
        # Many distinct values: the warnings fire, the estimate is still defined.
        xm = rng.integers(0, 900, n)
        put("est_disc_mi_many", x=xm, y=y, out=est._estimate_discrete_mi(xm, y))
        put("est_cond_disc_many", x=xm, y=y, cond=z1,
            out=est._estimate_conditional_discrete_mi(xm, y, z1))
    return cases


def main():
    if sys.argv[1:] == ["discrete"]:
        np.savez_compressed(os.path.join(HERE, "discrete_cases.npz"), **discrete_cases())
    else:
        np.savez_compressed(os.path.join(HERE, "knn_cases.npz"), **knn_cases())
        np.savez_compressed(os.path.join(HERE, "driver_cases.npz"), **driver_cases())
        np.savez_compressed(os.path.join(HERE, "discrete_cases.npz"), **discrete_cases())
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
