"""The inclusive ball-count boundary on the GPU.

ennemi counts the points within eps - 1e-12 with an inclusive test
(query_ball_point, ennemi/_entropy_estimators.py:104-110), and SciPy's own
test pins that boundary.  These inputs put, for every query of a group, a
candidate at a distance of exactly r = fl(1 - 1e-12) while the query's
k-th-neighbour radius is exactly 1, so the count radius fl(eps - 1e-12) is r
itself: a count kernel that used a strict '<' (or dropped the 1e-12) gets
these counts wrong.  Each group is, on the first axis, a query at 1 with a
candidate at 1 - r (computed exactly: the difference is r) or a query at 0
with a candidate at r; the groups are 4 apart on another axis, so they do
not interact in the joint space.  Covered: the 1-D count path (KSG
marginals), the multi-D grid path and the multi-D key-window path (CMI
spaces, default tuning and strip 0), FP32-filter and FP64 modes.
"""

import numpy as np
import pytest

from oracle import ennemi_oracle as orc

pytestmark = pytest.mark.gpu

R = 1.0 - 1e-12


def boundary_groups(m: int, seed: int):
    """(x, y, z1, z2) of 4 m points: per group a query whose radius (k = 3)
    is exactly 1 and a candidate at exactly r = fl(1 - 1e-12) from it on x."""
    rng = np.random.default_rng(seed)
    rows = []
    for g in range(m):
        yg = 4.0 * g
        if rng.random() < 0.5:
            q, c, far = 1.0, 1.0 - R, 0.0
        else:
            q, c, far = 0.0, R, 1.0
        rows += [(q, 0.0, yg, 0.0), (c, 0.0, yg, 0.0), (far, 0.0, yg, 0.0), (q, 0.0, yg + 0.5, 0.0)]
    pts = np.array(rows)
    return pts[rng.permutation(len(pts))]


def _boundary_hits(pts, eps):
    """Queries with a neighbour at exactly fl(eps - 1e-12)."""
    r = eps - 1e-12
    return sum(int(np.any(np.max(np.abs(pts - pts[i]), axis=1) == r[i])) for i in range(len(pts)))


def _strict(pts, r):
    return np.array([(np.max(np.abs(pts - pts[i]), axis=1) < r[i]).sum() for i in range(len(pts))])


def test_boundary_1d_counts():
    import ennemi_amd._native as nat
    from ennemi_amd._batch import Batch
    pts = boundary_groups(300, 1)
    x = np.ascontiguousarray(pts[:, 0])
    y = np.ascontiguousarray(pts[:, 2])
    for flags in (0, nat.F_FP64):
        b = Batch()
        b.add(nat.KIND_KSG, 3, [x], [y])
        vals, eps, cnts = b.run(flags=flags, export=True)
        w, ref = orc.estimate_single_mi(x, y, 3, detail=True)
        np.testing.assert_array_equal(eps[0], ref["eps"])
        assert _boundary_hits(x[:, None], ref["eps"]) >= 300
        assert (_strict(x[:, None], ref["eps"] - 1e-12) != ref["nx"]).any()
        np.testing.assert_array_equal(cnts[0][0], ref["nx"])
        np.testing.assert_array_equal(cnts[0][1], ref["ny"])
        assert abs(vals[0] - w) <= 1e-12


@pytest.mark.parametrize("strip", [None, 0])
def test_boundary_multid_counts(strip):
    import ennemi_amd._native as nat
    from ennemi_amd._batch import Batch
    pts = boundary_groups(250, 2)
    cols = [np.ascontiguousarray(pts[:, c]) for c in range(4)]
    ctx = nat.context()
    old = ctx.get_tuning()
    try:
        if strip is not None:
            ctx.set_tuning(strip, old[1])
        X, Y, Z = pts[:, :1], pts[:, 1:2], pts[:, 2:]
        w, ref = orc.estimate_conditional_mi(X, Y, Z, 3, detail=True)
        xz = pts[:, [0, 2, 3]]
        assert _boundary_hits(xz, ref["eps"]) >= 250
        assert (_strict(xz, ref["eps"] - 1e-12) != ref["nxz"]).any()
        for flags in (0, nat.F_FP64):
            b = Batch()
            b.add(nat.KIND_CMI, 3, [cols[0]], [cols[1]], cols[2:])
            vals, eps, cnts = b.run(flags=flags, export=True)
            np.testing.assert_array_equal(eps[0], ref["eps"])
            for s, key in enumerate(("nxz", "nyz", "nz")):
                np.testing.assert_array_equal(cnts[0][s], ref[key], err_msg=f"{key} flags={flags}")
            assert abs(vals[0] - w) <= 1e-12 or (np.isinf(w) and vals[0] == w)
    finally:
        ctx.set_tuning(*old)



def boundary_groups_on_y(m: int, seed: int, layout: str):
    """(x0, x1, y) of 4 m points for a KSG with a two-column x, with
    candidates exactly at the count radius on x1, the grid's second column.
    "ties": radius E = 2^20 (fl(E - 1e-12) = E), per group a query at
    x1 = 4 E g and candidates at x1 +- E; every x0 is 0, so every strip
    lies within every query's radius on x0.  "blocks": radius 1, count
    radius r = fl(1 - 1e-12), as in boundary_groups on x1, with sub-groups
    4 apart on y; blocks of 16 sub-groups (64 points) share x0 = 4 b, so a
    strip of 64 points is one block (within every radius on x0) and wider
    strips hold several (not within)."""
    rng = np.random.default_rng(seed)
    rows = []
    E = 2.0 ** 20
    for g in range(m):
        if layout == "ties":
            base = 4.0 * E * g
            rows += [(0.0, base, 0.0), (0.0, base + E, 0.0), (0.0, base - E, 0.0), (0.0, base, 0.5 * E)]
        else:
            u, yg = 4.0 * (g // 16), 4.0 * (g % 16)
            if rng.random() < 0.5:
                q, c, far = 1.0, 1.0 - R, 0.0
            else:
                q, c, far = 0.0, R, 1.0
            rows += [(u, q, yg), (u, c, yg), (u, far, yg), (u, q, yg + 0.5)]
    pts = np.array(rows)
    return pts[rng.permutation(len(pts))]


@pytest.mark.parametrize("strip", [None, 64, 128])
@pytest.mark.parametrize("layout", ["ties", "blocks"])
def test_boundary_2d_grid_counts(strip, layout):
    """A two-column count space on a grid (the x space of a KSG with a
    two-column x): strips whose whole key range lies within a query's
    radius are counted by binary searches of the y column, the others
    scanned; both must include a candidate at exactly the count radius."""
    import ennemi_amd._native as nat
    from ennemi_amd._batch import Batch
    pts = boundary_groups_on_y(400, 3, layout)
    x, y = pts[:, :2], pts[:, 2]
    w, ref = orc.estimate_single_mi(x, y, 3, detail=True)
    assert _boundary_hits(x, ref["eps"]) >= 400
    assert (_strict(x, ref["eps"] - 1e-12) != ref["nx"]).any()
    ctx = nat.context()
    old = ctx.get_tuning()
    try:
        if strip is not None:
            ctx.set_tuning(strip, old[1])
        for flags in (0, nat.F_FP64):
            b = Batch()
            b.add(nat.KIND_KSG, 3, [np.ascontiguousarray(x[:, 0]), np.ascontiguousarray(x[:, 1])],
                  [np.ascontiguousarray(y)])
            vals, eps, cnts = b.run(flags=flags, export=True)
            np.testing.assert_array_equal(eps[0], ref["eps"])
            np.testing.assert_array_equal(cnts[0][0], ref["nx"], err_msg=f"nx flags={flags}")
            np.testing.assert_array_equal(cnts[0][1], ref["ny"], err_msg=f"ny flags={flags}")
            assert abs(vals[0] - w) <= 1e-12
    finally:
        ctx.set_tuning(*old)
