"""Batch of estimator instances, packed into one float64 arena in HBM.

A ``Batch`` collects already validated and rescaled columns (NumPy float64)
and estimator instances that reference them.  Columns passed as the same
NumPy object are stored once, so instances that share a column (the x role
of one variable across ``pairwise_mi`` pairs, say) also share its device sort.
``Batch.run`` uploads the arena through PyTorch-ROCm, runs every instance in
one ``enm_run_batch`` call and returns one float64 value per instance.
"""

from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat


class Batch:
    def __init__(self) -> None:
        self._cols: List[np.ndarray] = []
        self._orig: List[np.ndarray] = []
        self._col_index: Dict[int, int] = {}
        self._inst: List[tuple] = []
        self._seg: List[int] = []

    def __len__(self) -> int:
        return len(self._inst)

    def _col(self, a: np.ndarray) -> int:
        # The originals are kept alive, so an id() cannot be reused meanwhile.
        key = id(a)
        idx = self._col_index.get(key)
        if idx is None:
            c = np.ascontiguousarray(a, dtype=np.float64)
            if c.ndim != 1:
                raise ValueError("columns must be one-dimensional")
            idx = len(self._cols)
            self._cols.append(c)
            self._orig.append(a)
            self._col_index[key] = idx
        return idx

    def add(self, kind: int, k: int, x: Sequence[np.ndarray], y: Sequence[np.ndarray] = (),
            z: Sequence[np.ndarray] = (), class_sizes: Optional[Sequence[int]] = None) -> int:
        """Add one instance; returns its index.  For ROSS/CROSS the points must
        be grouped by class, in class order, with the given class sizes."""
        cols = [self._col(c) for c in list(x) + list(y) + list(z)]
        if len(cols) > nat.MAX_COLS:
            raise ValueError(f"at most {nat.MAX_COLS} dimensions per instance")
        n = len(self._cols[cols[0]])
        for c in cols:
            if len(self._cols[c]) != n:
                raise ValueError("all columns of an instance must have the same length")
        seg_index = len(self._seg)
        if class_sizes is None:
            n_segs = 1
        else:
            n_segs = len(class_sizes)
            self._seg.extend(int(s) for s in class_sizes)
        self._inst.append((kind, int(k), n, len(x), len(y), len(z), n_segs, seg_index, cols))
        return len(self._inst) - 1

    def pack(self):
        """Host arena (float64) and the ctypes instance / class-size tables."""
        offsets = np.zeros(len(self._cols) + 1, dtype=np.int64)
        for i, c in enumerate(self._cols):
            offsets[i + 1] = offsets[i] + len(c)
        arena = np.empty(int(offsets[-1]), dtype=np.float64)
        for i, c in enumerate(self._cols):
            arena[offsets[i]:offsets[i + 1]] = c
        inst_arr = _instance_table(self._inst, offsets)
        seg_arr = (ctypes.c_int64 * max(1, len(self._seg)))(*self._seg)
        return arena, inst_arr, seg_arr

    def point_offsets(self) -> np.ndarray:
        n = np.array([t[2] for t in self._inst], dtype=np.int64)
        return np.concatenate([[0], np.cumsum(n)])

    def device(self, flags: int = 0, export: bool = False):
        """Upload the arena; return a DeviceBatch ready to launch."""
        return DeviceBatch(self, flags, export)

    def run(self, flags: int = 0, export: bool = False):
        """Run every instance.  Returns values, or (values, eps, counts) with
        per-instance lists of radii (n,) and counts (3, n) when export=True."""
        db = self.device(flags, export)
        db.launch()
        return db.results()


def _instance_table(inst: List[tuple], col_offsets: Optional[np.ndarray]):
    """The ctypes enm_instance table of (kind, k, n, dx, dy, dz, n_segs,
    seg_index, columns) tuples; a column is an arena offset, or an index
    into col_offsets."""
    n_inst = len(inst)
    inst_arr = (nat.EnmInstance * n_inst)()
    if n_inst == 0:
        return inst_arr
    tab = np.ctypeslib.as_array(inst_arr)   # the same memory, field by field
    head = np.array([t[:8] for t in inst], dtype=np.int64).reshape(n_inst, 8)
    for f, name in enumerate(("kind", "k", "n", "dx", "dy", "dz", "n_segs", "seg_index")):
        tab[name] = head[:, f]
    width = max(len(t[8]) for t in inst)
    cols = np.zeros((n_inst, width), dtype=np.int64)
    for i, t in enumerate(inst):
        cols[i, :len(t[8])] = t[8]
    if col_offsets is not None:
        cols = np.where(np.arange(width) < np.array([len(t[8]) for t in inst])[:, None], col_offsets[cols], 0)
    tab["col"][:, :width] = cols
    return inst_arr


class DeviceBatch:
    """A packed batch resident in HBM, launchable many times (bench, graphs)."""

    def __init__(self, batch: Batch, flags: int, export: bool, hints: bool = True):
        import torch
        self.batch = batch
        # Sort hints of a device-prepared arena (enm_run_batch_hinted).
        self.hint = getattr(batch, "device_hint", None) if hints else None
        self.ctx = nat.context()
        self.flags = flags
        arena, self.inst_arr, self.seg_arr = batch.pack()
        self.n_inst = len(batch)
        dev = torch.device("cuda", self.ctx.device)
        self.stream = _stream()
        with torch.cuda.stream(self.stream):
            # A batch prepared on the device brings its arena along.
            self.arena = batch.device_arena if arena is None else torch.from_numpy(arena).to(dev)
            self.out = torch.empty(self.n_inst, dtype=torch.float64, device=dev)
            self.offs = batch.point_offsets()
            total = int(self.offs[-1])
            self.eps = torch.empty(total, dtype=torch.float64, device=dev) if export else None
            self.counts = torch.empty(3 * total, dtype=torch.int32, device=dev) if export else None
        self.stream.synchronize()

    def launch(self) -> None:
        """Enqueue one run on the batch's stream (returns immediately)."""
        with self.ctx.lock:
            self.ctx.run(self.inst_arr, self.n_inst, self.seg_arr, self.arena.data_ptr(),
                         self.out.data_ptr(), self.eps.data_ptr() if self.eps is not None else 0,
                         self.counts.data_ptr() if self.counts is not None else 0,
                         self.flags, self.stream.cuda_stream,
                         self.hint.data_ptr() if self.hint is not None else 0)

    def stats(self) -> dict:
        with self.ctx.lock:
            return self.ctx.stats()

    def autotune(self) -> Tuple[int, int]:
        """Time the library's tile shapes on this batch and keep the fastest
        (in the context, for every later batch); returns (strip, waves/CU)."""
        with self.ctx.lock:
            self.ctx.autotune(self.inst_arr, self.n_inst, self.seg_arr, self.arena.data_ptr(),
                              self.out.data_ptr(), self.flags, self.stream.cuda_stream)
            self.stream.synchronize()
            return self.ctx.get_tuning()

    def results(self):
        self.stream.synchronize()
        values = self.out.cpu().numpy().copy()
        if self.eps is None:
            return values
        eps_all = self.eps.cpu().numpy()
        cnt_all = self.counts.cpu().numpy()
        eps, cnts = [], []
        for i in range(self.n_inst):
            a, b = int(self.offs[i]), int(self.offs[i + 1])
            n = b - a
            eps.append(eps_all[a:b].copy())
            cnts.append(cnt_all[3 * a:3 * b].reshape(3, n).copy())
        return values, eps, cnts


class ArenaBatch:
    """A batch whose float64 arena was written on the device (the columns
    of device-prepared estimate_mi tasks, enm_prep_write), with the sort
    hints of its columns when they were made (enm_prep_hints); instances
    name their columns by arena offset.  Same interface as Batch."""

    def __init__(self, device_arena, device_hint=None) -> None:
        self.device_arena = device_arena
        self.device_hint = device_hint
        self._inst: List[tuple] = []

    def __len__(self) -> int:
        return len(self._inst)

    def add(self, kind: int, k: int, n: int, offsets: Sequence[int], dx: int, dy: int, dz: int) -> int:
        self._inst.append((kind, int(k), int(n), dx, dy, dz, 1, 0, [int(o) for o in offsets]))
        return len(self._inst) - 1

    def pack(self):
        inst_arr = _instance_table(self._inst, None)
        seg_arr = (ctypes.c_int64 * 1)(0)
        return None, inst_arr, seg_arr

    def point_offsets(self) -> np.ndarray:
        n = np.array([t[2] for t in self._inst], dtype=np.int64)
        return np.concatenate([[0], np.cumsum(n)])

    def device(self, flags: int = 0, export: bool = False, hints: bool = True):
        return DeviceBatch(self, flags, export, hints)

    def run(self, flags: int = 0, export: bool = False, hints: bool = True):
        db = self.device(flags, export, hints)
        db.launch()
        return db.results()


_side_stream = None


def _stream():
    """A dedicated stream at default priority (hipGraph capture needs a
    non-legacy stream)."""
    global _side_stream
    import torch
    if _side_stream is None:
        _side_stream = torch.cuda.Stream()
    return _side_stream
