"""Seeded synthetic inputs of the benchmark configurations.

C1  KSG, N = 10,000, rho = 0.6 (seed 101).
C2  Frenzel-Pompe CMI, N = 100,000, 3-D Gaussian condition (seed 202), and the
    Kozachenko-Leonenko entropy of a 4-D Gaussian with a seeded SPD covariance.
C3  masked lag scan: 8 coupled AR(1) series, lags -48..48, 0-30 % NaNs per x
    column (seed 303); the headline uses length 100,096 (N ~ 1e5).
C4  pairwise_mi over 64 mixed columns x 50,000 rows, 10 % of the columns
    rounded to 2 decimals (seed 404).
C5a CMI at N = 2,000,000 with a 2-D Gaussian condition, k = 5 (seed 505).
C5b Ross MI, N = 1,000,000, x 3-D Gaussian mixture over 16 Zipf(1.1) classes
    (seed 506).
"""

from __future__ import annotations

import numpy as np


def c1(n: int = 10_000):
    rng = np.random.default_rng(101)
    d = rng.multivariate_normal([0, 0], [[1, 0.6], [0.6, 1]], size=n)
    return d[:, 0], d[:, 1]


def c2_cmi(n: int = 100_000):
    rng = np.random.default_rng(202)
    z = rng.normal(size=(n, 3))
    a = rng.normal(size=3)
    b = rng.normal(size=3)
    x = z @ a + 0.5 * rng.normal(size=n)
    y = z @ b + 0.5 * rng.normal(size=n)
    return x, y, z


def c2_entropy(n: int = 100_000):
    rng = np.random.default_rng(202)
    a = rng.normal(size=(4, 4))
    cov = a @ a.T + 0.5 * np.eye(4)
    x = rng.multivariate_normal(np.zeros(4), cov, size=n)
    return x, cov


def lag_scan(length: int, seed: int = 303):
    """Eight coupled AR(1) series (coefficients 0.3-0.9, coupling 0.2,
    burn-in 1000); y = series 0, x = all eight series with NaNs at a random
    0-30 % of the rows of each column."""
    rng = np.random.default_rng(seed)
    phi = np.linspace(0.3, 0.9, 8)
    burn = 1000
    total = length + burn
    noise = rng.normal(size=(total, 8))
    s = np.zeros((total, 8))
    for t in range(1, total):
        prev = s[t - 1]
        cur = phi * prev + noise[t]
        cur[:7] += 0.2 * prev[1:]
        s[t] = cur
    s = s[burn:]
    y = s[:, 0].copy()
    x = s.copy()
    for j in range(8):
        frac = rng.uniform(0.0, 0.3)
        rows = rng.choice(length, int(frac * length), replace=False)
        x[rows, j] = np.nan
    return y, x


def c4(n: int = 50_000, nvar: int = 64):
    rng = np.random.default_rng(404)
    data = np.empty((n, nvar))
    q = nvar // 4
    data[:, :q] = rng.standard_t(3, size=(n, q))
    data[:, q:2 * q] = rng.uniform(-1, 1, size=(n, q))
    data[:, 2 * q:3 * q] = rng.lognormal(0, 1, size=(n, q))
    # Columns 48-63: sin(col[c - 48]) (even c) or col[c - 40]^2 (odd c), plus noise.
    for c in range(3 * q, nvar):
        if c % 2 == 0:
            data[:, c] = np.sin(data[:, c - 3 * q]) + 0.3 * rng.normal(size=n)
        else:
            data[:, c] = data[:, c - 3 * q + q // 2] ** 2 + 0.3 * rng.normal(size=n)
    for c in range(0, nvar, 10):
        data[:, c] = np.round(data[:, c], 2)
    return data


def c5a(n: int = 2_000_000):
    rng = np.random.default_rng(505)
    z = rng.multivariate_normal([0, 0], [[1, 0.5], [0.5, 1]], size=n)
    a = rng.normal(size=2)
    b = rng.normal(size=2)
    x = z @ a + 0.5 * rng.normal(size=n)
    y = z @ b + 0.5 * rng.normal(size=n)
    return x, y, z


def c5b(n: int = 1_000_000, classes: int = 16):
    rng = np.random.default_rng(506)
    p = (np.arange(classes) + 1.0) ** -1.1
    p /= p.sum()
    y = rng.choice(classes, size=n, p=p)
    c = np.arange(classes)
    mu = np.column_stack([np.cos(2 * np.pi * c / classes), np.sin(2 * np.pi * c / classes), c / 8])
    x = mu[y] + rng.normal(size=(n, 3))
    return x, y
