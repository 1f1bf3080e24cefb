"""enm_run_batch_static, the entry point that allocates no device memory.

Its results must equal enm_run_batch's on the same batch: values, radii and
counts bitwise (the same kernels run on the same plan; only the workspace
differs).  Two calls in a row, the second smaller, check that nothing of
the first call's contents in the shared static array leaks into the second.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _nat():
    import ennemi_amd._native as nat
    return nat


def _mixed_batch(scale: int, seed: int):
    """KSG, CMI, Ross (4 classes), conditional Ross and KL instances."""
    from ennemi_amd._batch import Batch
    nat = _nat()
    rng = np.random.default_rng(seed)
    b = Batch()
    n = 6000 * scale
    cov = np.array([[1.0, 0.5, 0.2], [0.5, 1.0, 0.3], [0.2, 0.3, 1.0]])
    d = rng.multivariate_normal([0, 0, 0], cov, size=n)
    x, y, z = (np.ascontiguousarray(d[:, c]) for c in range(3))
    b.add(nat.KIND_KSG, 3, [x], [y])
    b.add(nat.KIND_KSG, 5, [x[: n // 2]], [np.round(y[: n // 2], 1)])
    b.add(nat.KIND_CMI, 3, [x], [y], [z])
    sizes = [n // 8, n // 4, n // 8, n // 2]
    xr = np.ascontiguousarray(rng.normal(size=n) + np.repeat(np.arange(4.0), sizes))
    b.add(nat.KIND_ROSS, 3, [xr], class_sizes=sizes)
    b.add(nat.KIND_CROSS, 4, [xr], z=[z], class_sizes=sizes)
    w = rng.normal(size=(n // 3, 2))
    b.add(nat.KIND_KL, 2, [np.ascontiguousarray(w[:, 0]), np.ascontiguousarray(w[:, 1])])
    return b


def _run_static(db, out, eps, cnt):
    db.ctx.run_static(db.inst_arr, db.n_inst, db.seg_arr, db.arena.data_ptr(), out.data_ptr(),
                      eps.data_ptr(), cnt.data_ptr(), db.flags, db.stream.cuda_stream)


def _check_against_run_batch(b, flags):
    import torch
    db = b.device(flags, export=True)
    db.launch()
    ref_vals, ref_eps, ref_cnt = db.results()
    ref_eps_all = db.eps.cpu().numpy().copy()
    ref_cnt_all = db.counts.cpu().numpy().copy()
    # Fresh outputs, poisoned so that an unwritten slot shows.
    out = torch.full_like(db.out, float("nan"))
    eps = torch.full_like(db.eps, -7.0)
    cnt = torch.full_like(db.counts, -7)
    with db.ctx.lock:
        _run_static(db, out, eps, cnt)
    db.stream.synchronize()
    vals = out.cpu().numpy()
    np.testing.assert_array_equal(vals, ref_vals)
    np.testing.assert_array_equal(eps.cpu().numpy(), ref_eps_all)
    np.testing.assert_array_equal(cnt.cpu().numpy(), ref_cnt_all)
    return vals


def test_static_entry_point_matches_run_batch_twice():
    nat = _nat()
    big = _mixed_batch(scale=4, seed=11)
    small = _mixed_batch(scale=1, seed=12)
    for flags in (0, nat.F_FP64):
        print(f"static vs run_batch: big batch, flags={flags}", flush=True)
        _check_against_run_batch(big, flags)
        print(f"static vs run_batch: smaller batch, flags={flags}", flush=True)
        _check_against_run_batch(small, flags)


def test_static_entry_point_rejects_oversized_batch_and_graph():
    import torch
    from ennemi_amd._batch import Batch
    nat = _nat()
    rng = np.random.default_rng(13)
    # About 3 million KSG points need more than the 256 MiB workspace.
    n = 3_000_000
    b = Batch()
    b.add(nat.KIND_KSG, 3, [rng.normal(size=n)], [rng.normal(size=n)])
    db = b.device(0, export=False)
    out = torch.empty_like(db.out)
    with db.ctx.lock:
        rc = db.ctx.lib.enm_run_batch_static(db.ctx.ctx, db.inst_arr, db.n_inst, db.seg_arr,
                                             db.arena.data_ptr(), out.data_ptr(), None, None, 0,
                                             db.stream.cuda_stream)
        msg = db.ctx.error()
    assert rc == -3, (rc, msg)
    assert "holds at most" in msg and str(nat.STATIC_WORKSPACE_BYTES) in msg, msg
    small = _mixed_batch(scale=1, seed=14)
    dbs = small.device(0, export=False)
    with dbs.ctx.lock:
        rc = dbs.ctx.lib.enm_run_batch_static(dbs.ctx.ctx, dbs.inst_arr, dbs.n_inst, dbs.seg_arr,
                                              dbs.arena.data_ptr(), dbs.out.data_ptr(), None, None,
                                              nat.F_GRAPH, dbs.stream.cuda_stream)
        msg = dbs.ctx.error()
    assert rc == -1 and "ENM_F_GRAPH" in msg, (rc, msg)
    # The context still works after the rejections.
    _check_against_run_batch(small, 0)
