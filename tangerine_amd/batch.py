"""Batched export of many trees in one persistent launch (tg_batch_*)."""

from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import check
from .export import Mesher


class _BorrowedMesher(Mesher):
    """A tree's mesher inside a batch (owned by the batch)."""

    def __init__(self, lib, handle, cells):
        self.lib = lib
        self.h = handle
        self._cells = tuple(int(c) for c in cells)

    def close(self) -> None:
        self.h = None


class Batch:
    def __init__(self):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.tg_batch_create(C.byref(h)), "tg_batch_create")
        self.h = h.value
        self._programs = []
        self._cells = []

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.tg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self, programs, grids, fast: bool = False) -> dict:
        """programs: Program objects; grids: (origin, step, cells) per tree."""
        n = len(programs)
        progs = (C.c_void_p * n)(*[p.handle for p in programs])
        gs = (_lib.TgGrid * n)(*[_lib.grid_struct(*g) for g in grids])
        check(self.lib.tg_batch_export(self.h, progs, gs, n, _lib.TG_FLAG_FAST if fast else 0), "tg_batch_export")
        self._programs = list(programs)  # the batch replays them: keep them alive
        self._cells = [tuple(g[2]) for g in grids]
        return self.stats()

    def replay(self) -> float:
        ms = C.c_double()
        check(self.lib.tg_batch_replay(self.h, C.byref(ms)), "tg_batch_replay")
        return ms.value

    def autotune(self) -> dict:
        r = _lib.TgBatchTuneResult()
        check(self.lib.tg_batch_autotune(self.h, C.byref(r)), "tg_batch_autotune")
        return {
            "points_per_lane": r.points_per_lane, "waves": r.waves, "workgroups_per_cu": r.workgroups_per_cu,
            "eval_ms": r.eval_ms,
            "candidates": [{"points_per_lane": r.candidate_points[i], "waves": r.candidate_waves[i],
                            "workgroups_per_cu": r.candidate_workgroups_per_cu[i], "ms": r.candidate_ms[i]}
                           for i in range(r.candidates)],
        }

    def set_tuning(self, points_per_lane: int, waves: int, workgroups_per_cu: int) -> None:
        check(self.lib.tg_batch_set_tuning(self.h, points_per_lane, waves, workgroups_per_cu), "tg_batch_set_tuning")

    def mesher(self, i: int) -> Mesher:
        h = self.lib.tg_batch_mesher(self.h, int(i))
        if not h:
            raise _lib.TangerineError(self.lib.tg_last_error().decode())
        return _BorrowedMesher(self.lib, h, self._cells[i])

    def stats(self) -> dict:
        st = _lib.TgBatchStats()
        check(self.lib.tg_batch_stats_get(self.h, C.byref(st)), "tg_batch_stats_get")
        return {k: getattr(st, k) for k, _ in _lib.TgBatchStats._fields_}
