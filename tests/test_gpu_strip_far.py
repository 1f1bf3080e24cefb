"""The strip kNN's far list (two-column grid orders).

A query whose radius still reaches past the two staged neighbour strips is
listed and finished by knn_far_kernel; when the list is full the rest are
finished in place.  Narrow strips (64 points) on heavy-tailed data send most
queries past their neighbours, so both the list and its overflow run; the
radii, counts and values must be those of the oracle (cKDTree) and of the
default strips, in every k class of the register lists.
"""

import numpy as np
import pytest

from oracle import ennemi_oracle as orc
from tests.test_gpu_parity import _batch_one, _close, _cols, _nat

pytestmark = pytest.mark.gpu


def _heavy(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.standard_t(2, size=n)
    y = 0.5 * x + rng.standard_t(3, size=n)
    return np.column_stack([x, y])


@pytest.mark.parametrize("strip", [64, 128, 1024])
@pytest.mark.parametrize("k", [3, 6, 12, 25])
def test_far_list_matches_oracle(strip, k):
    nat = _nat()
    a = _heavy(1000 + k, 6000)
    ctx = nat.context()
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(strip, old[1])
        v, eps, c = _batch_one(nat.KIND_KSG, k, _cols(a[:, :1]), _cols(a[:, 1:]))
    finally:
        ctx.set_tuning(*old)
    w, ref = orc.estimate_single_mi(a[:, :1], a[:, 1:], k, detail=True)
    np.testing.assert_array_equal(eps, ref["eps"])
    np.testing.assert_array_equal(c[0], ref["nx"])
    np.testing.assert_array_equal(c[1], ref["ny"])
    assert _close(v, w), (v, w)
