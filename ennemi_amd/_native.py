"""ctypes binding of libennemi_hip.so (declared in include/ennemi_hip.h).

PyTorch-ROCm is used only to own device buffers and name the stream; every
estimate is computed by the HIP kernels behind this binding.  There is no
CPU fallback: if the library or a gfx950 device is missing, the first call
raises ``NativeUnavailable``.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libennemi_hip.so")

KIND_KSG, KIND_CMI, KIND_ROSS, KIND_CROSS, KIND_KL = 0, 1, 2, 3, 4
MAX_COLS = 16

F_FP64 = 0x1
F_BRUTE = 0x2
F_GRAPH = 0x4
F_PER_INSTANCE = 0x8
F_TIMING = 0x10
STAGES = ("fill", "sort", "gather", "knn", "count", "combine")

EXPORTED_SYMBOLS = ("enm_init", "enm_shutdown", "enm_last_error", "enm_run_batch", "enm_run_batch_hinted",
                    "enm_run_batch_static", "enm_prep_count", "enm_prep_write", "enm_prep_hints",
                    "enm_unique_rows", "enm_get_stats",
                    "enm_autotune", "enm_set_tuning", "enm_set_tuning_ex", "enm_get_tuning",
                    "enm_get_tuning_for")
# Workspace of enm_run_batch_static, bytes (ENM_STATIC_WORKSPACE_BYTES).
STATIC_WORKSPACE_BYTES = 256 << 20


class NativeUnavailable(RuntimeError):
    """The HIP library or a gfx950 GPU is not available."""


class EnmInstance(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("k", ctypes.c_int32), ("n", ctypes.c_int64),
                ("dx", ctypes.c_int32), ("dy", ctypes.c_int32), ("dz", ctypes.c_int32),
                ("n_segs", ctypes.c_int32), ("seg_index", ctypes.c_int64),
                ("col", ctypes.c_int64 * MAX_COLS)]


PREP_MAXZ = MAX_COLS - 2
P_DROP_NAN = 0x1
P_RESCALE = 0x2
P_HINTS = 0x4


class EnmPrepTask(ctypes.Structure):
    _fields_ = [("x_col", ctypes.c_int64), ("y_col", ctypes.c_int64),
                ("x0", ctypes.c_int64), ("y0", ctypes.c_int64), ("w", ctypes.c_int64),
                ("dz", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("z_col", ctypes.c_int64 * PREP_MAXZ), ("z0", ctypes.c_int64 * PREP_MAXZ)]


class EnmRowset(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("ncol", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("col", ctypes.c_int64 * MAX_COLS)]


E_UNSUPPORTED = -4


class EnmTuning(ctypes.Structure):
    _fields_ = [("strip_points", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("list_slots", ctypes.c_int32), ("items_per_ticket", ctypes.c_int32)]

    def as_tuple(self) -> Tuple[int, int, int, int]:
        return (self.strip_points, self.waves_per_cu, self.list_slots, self.items_per_ticket)


class EnmStats(ctypes.Structure):
    _fields_ = [("knn_pairs", ctypes.c_uint64), ("knn_rechecks", ctypes.c_uint64),
                ("count_pairs", ctypes.c_uint64), ("count_rechecks", ctypes.c_uint64),
                ("count_queries_1d", ctypes.c_uint64), ("sort_keys", ctypes.c_uint64),
                ("n_launches", ctypes.c_uint32), ("graph_replayed", ctypes.c_uint32),
                ("stage_ms", ctypes.c_float * 6), ("total_ms", ctypes.c_float)]

    def as_dict(self) -> dict:
        d = {name: int(getattr(self, name)) for name, _ in self._fields_[:8]}
        d["stage_ms"] = {s: float(self.stage_ms[i]) for i, s in enumerate(STAGES)}
        d["total_ms"] = float(self.total_ms)
        return d


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the shared library and declare its C signatures (no GPU needed)."""
    if not os.path.exists(path):
        raise NativeUnavailable(
            f"{path} is missing; build it with `python -m ennemi_amd.build` "
            "(ennemi_amd has no CPU implementation)")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.enm_init.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.enm_init.restype = ctypes.c_int
    lib.enm_shutdown.argtypes = [vp]
    lib.enm_shutdown.restype = None
    lib.enm_last_error.argtypes = [vp]
    lib.enm_last_error.restype = ctypes.c_char_p
    lib.enm_run_batch.argtypes = [vp, ctypes.POINTER(EnmInstance), ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_int64), vp, vp, vp, vp,
                                  ctypes.c_uint32, vp]
    lib.enm_run_batch.restype = ctypes.c_int
    lib.enm_run_batch_hinted.argtypes = [vp, ctypes.POINTER(EnmInstance), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_int64), vp, vp, vp, vp, vp,
                                         ctypes.c_uint32, vp]
    lib.enm_run_batch_hinted.restype = ctypes.c_int
    lib.enm_run_batch_static.argtypes = lib.enm_run_batch.argtypes
    lib.enm_run_batch_static.restype = ctypes.c_int
    lib.enm_prep_count.argtypes = [vp, ctypes.POINTER(EnmPrepTask), ctypes.c_int32, vp, vp,
                                   ctypes.c_uint32, vp, vp, vp]
    lib.enm_prep_count.restype = ctypes.c_int
    lib.enm_prep_write.argtypes = [vp, ctypes.POINTER(EnmPrepTask), ctypes.c_int32, vp, vp, vp,
                                   ctypes.c_int64, ctypes.c_uint32, vp, vp, vp, vp]
    lib.enm_prep_write.restype = ctypes.c_int
    lib.enm_prep_hints.argtypes = [vp, ctypes.POINTER(EnmPrepTask), ctypes.c_int32, vp, vp, vp,
                                   ctypes.c_int64, ctypes.c_uint32, vp, vp, vp, vp]
    lib.enm_prep_hints.restype = ctypes.c_int
    lib.enm_unique_rows.argtypes = [vp, ctypes.POINTER(EnmRowset), ctypes.c_int32, vp, vp, vp, vp, vp]
    lib.enm_unique_rows.restype = ctypes.c_int
    lib.enm_get_stats.argtypes = [vp, ctypes.POINTER(EnmStats)]
    lib.enm_get_stats.restype = ctypes.c_int
    lib.enm_autotune.argtypes = [vp, ctypes.POINTER(EnmInstance), ctypes.c_int32,
                                 ctypes.POINTER(ctypes.c_int64), vp, vp, ctypes.c_uint32, vp]
    lib.enm_autotune.restype = ctypes.c_int
    lib.enm_set_tuning.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.enm_set_tuning.restype = ctypes.c_int
    lib.enm_get_tuning.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.enm_get_tuning.restype = ctypes.c_int
    lib.enm_set_tuning_ex.argtypes = [vp, ctypes.POINTER(EnmTuning)]
    lib.enm_set_tuning_ex.restype = ctypes.c_int
    lib.enm_get_tuning_for.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.POINTER(EnmTuning)]
    lib.enm_get_tuning_for.restype = ctypes.c_int
    return lib


class Context:
    """One library context on cuda:0 (a process-wide singleton by default)."""

    def __init__(self, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise NativeUnavailable("ennemi_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = load_library()
        self.device = device
        torch.cuda.init()
        with torch.cuda.device(device):
            ptr = ctypes.c_void_p()
            rc = self.lib.enm_init(device, ctypes.byref(ptr))
            if rc != 0:
                raise NativeUnavailable("enm_init failed: " + self.lib.enm_last_error(None).decode())
        self.ctx = ptr
        self.lock = threading.Lock()

    def error(self) -> str:
        return self.lib.enm_last_error(self.ctx).decode()

    def check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.error()}")

    def run(self, inst_arr, n_inst: int, seg_arr, arena_ptr: int, out_ptr: int,
            eps_ptr: int = 0, cnt_ptr: int = 0, flags: int = 0, stream_ptr: int = 0,
            hint_ptr: int = 0) -> None:
        """enm_run_batch, or enm_run_batch_hinted when hint_ptr is given."""
        if hint_ptr:
            rc = self.lib.enm_run_batch_hinted(self.ctx, inst_arr, n_inst, seg_arr, arena_ptr, hint_ptr, out_ptr,
                                               eps_ptr or None, cnt_ptr or None, flags, stream_ptr or None)
            self.check(rc, "enm_run_batch_hinted")
            return
        rc = self.lib.enm_run_batch(self.ctx, inst_arr, n_inst, seg_arr, arena_ptr, out_ptr,
                                    eps_ptr or None, cnt_ptr or None, flags, stream_ptr or None)
        self.check(rc, "enm_run_batch")

    def run_static(self, inst_arr, n_inst: int, seg_arr, arena_ptr: int, out_ptr: int,
                   eps_ptr: int = 0, cnt_ptr: int = 0, flags: int = 0, stream_ptr: int = 0) -> None:
        """enm_run_batch_static: the same batch with no device allocation
        (workspace in the library's static array)."""
        rc = self.lib.enm_run_batch_static(self.ctx, inst_arr, n_inst, seg_arr, arena_ptr, out_ptr,
                                           eps_ptr or None, cnt_ptr or None, flags, stream_ptr or None)
        self.check(rc, "enm_run_batch_static")

    def prep_count(self, tasks, n_tasks: int, raw_ptr: int, mask_ptr: int, flags: int,
                   h_n: np.ndarray, h_nan: np.ndarray, stream_ptr: int = 0) -> None:
        rc = self.lib.enm_prep_count(self.ctx, tasks, n_tasks, raw_ptr, mask_ptr or None, flags,
                                     h_n.ctypes.data, h_nan.ctypes.data, stream_ptr or None)
        self.check(rc, "enm_prep_count")

    def prep_write(self, tasks, n_tasks: int, raw_ptr: int, mask_ptr: int, noise_ptr: int,
                   noise_len: int, flags: int, arena_off: np.ndarray, arena_ptr: int,
                   const_ptr: int, stream_ptr: int = 0) -> None:
        rc = self.lib.enm_prep_write(self.ctx, tasks, n_tasks, raw_ptr, mask_ptr or None,
                                     noise_ptr or None, noise_len, flags, arena_off.ctypes.data,
                                     arena_ptr, const_ptr, stream_ptr or None)
        self.check(rc, "enm_prep_write")

    def prep_hints(self, tasks, n_tasks: int, raw_ptr: int, mask_ptr: int, noise_ptr: int,
                   noise_len: int, flags: int, arena_off: np.ndarray, arena_ptr: int,
                   hint_ptr: int, stream_ptr: int = 0) -> None:
        rc = self.lib.enm_prep_hints(self.ctx, tasks, n_tasks, raw_ptr, mask_ptr or None,
                                     noise_ptr or None, noise_len, flags, arena_off.ctypes.data,
                                     arena_ptr, hint_ptr, stream_ptr or None)
        self.check(rc, "enm_prep_hints")

    def unique_rows(self, sets, n_sets: int, values_ptr: int, codes_ptr: int, counts_ptr: int,
                    h_distinct: np.ndarray, stream_ptr: int = 0) -> int:
        """enm_unique_rows; returns its status (0, or E_UNSUPPORTED when a row
        set's keys would not be exact), raising on any other failure."""
        rc = self.lib.enm_unique_rows(self.ctx, sets, n_sets, values_ptr, codes_ptr, counts_ptr,
                                      h_distinct.ctypes.data, stream_ptr or None)
        if rc != E_UNSUPPORTED:
            self.check(rc, "enm_unique_rows")
        return rc

    def stats(self) -> dict:
        st = EnmStats()
        self.check(self.lib.enm_get_stats(self.ctx, ctypes.byref(st)), "enm_get_stats")
        return st.as_dict()

    def set_tuning(self, strip_points: int, waves_per_cu: int) -> None:
        self.check(self.lib.enm_set_tuning(self.ctx, strip_points, waves_per_cu), "enm_set_tuning")

    def get_tuning(self) -> Tuple[int, int]:
        q = ctypes.c_int32()
        w = ctypes.c_int32()
        self.check(self.lib.enm_get_tuning(self.ctx, ctypes.byref(q), ctypes.byref(w)), "enm_get_tuning")
        return q.value, w.value

    def set_tuning_ex(self, strip_points: int, waves_per_cu: int, list_slots: int = 0,
                      items_per_ticket: int = 0) -> None:
        t = EnmTuning(strip_points, waves_per_cu, list_slots, items_per_ticket)
        self.check(self.lib.enm_set_tuning_ex(self.ctx, ctypes.byref(t)), "enm_set_tuning_ex")

    def tuning_for(self, dims: int, k: int, n: int) -> Tuple[int, int, int, int]:
        t = EnmTuning()
        self.check(self.lib.enm_get_tuning_for(self.ctx, dims, k, n, ctypes.byref(t)), "enm_get_tuning_for")
        return t.as_tuple()

    def autotune(self, inst_arr, n_inst, seg_arr, arena_ptr, out_ptr, flags=0, stream_ptr=0) -> None:
        self.check(self.lib.enm_autotune(self.ctx, inst_arr, n_inst, seg_arr, arena_ptr, out_ptr,
                                         flags, stream_ptr or None), "enm_autotune")


_context: Optional[Context] = None
_context_lock = threading.Lock()


def context() -> Context:
    global _context
    with _context_lock:
        if _context is None:
            _context = Context(0)
        return _context
