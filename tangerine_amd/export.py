"""Grid export on the GPU: the MeshExportThread shape (tangerine/export.cpp:320-381)
behind tg_mesher_* (evaluate every grid sample, then naive surface nets, all on
the device)."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check


class Mesher:
    """Owns one device stream and the export buffers (tg_mesher)."""

    def __init__(self):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.tg_mesher_create(C.byref(h)), "tg_mesher_create")
        self.h = h.value
        self._cells = None

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.tg_mesher_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self, program, origin, step, cells, fast: bool = False, prune: bool = True) -> dict:
        g = _lib.grid_struct(origin, step, cells)
        flags = (_lib.TG_FLAG_FAST if fast else 0) | (0 if prune else _lib.TG_FLAG_NO_PRUNE)
        check(self.lib.tg_mesher_export(self.h, program.handle, C.byref(g), flags), "tg_mesher_export")
        self._cells = tuple(int(c) for c in cells)
        return self.stats()

    def export_culled(self, program, surrogate, origin, step, cells, fast: bool = False, prune: bool = True,
                      certify: bool = True) -> dict:
        """Export that evaluates exactly only the kept bricks: those the
        surrogate keeps and, with `certify` (always when surrogate is None),
        every skipped brick whose exact centre value allows a sign change."""
        g = _lib.grid_struct(origin, step, cells)
        flags = ((_lib.TG_FLAG_FAST if fast else 0) | (0 if prune else _lib.TG_FLAG_NO_PRUNE) |
                 (_lib.TG_FLAG_CERTIFY if certify else 0))
        check(self.lib.tg_mesher_export_culled(self.h, program.handle, C.byref(g),
                                               None if surrogate is None else surrogate.h, flags),
              "tg_mesher_export_culled")
        self._cells = tuple(int(c) for c in cells)
        return self.stats()

    def autotune(self, program, origin, step, cells, fast: bool = False) -> dict:
        """Time every (points per lane, waves) configuration of the block
        kernel on this grid and keep the fastest for later exports."""
        g = _lib.grid_struct(origin, step, cells)
        r = _lib.TgTuneResult()
        check(self.lib.tg_mesher_autotune(self.h, program.handle, C.byref(g), _lib.TG_FLAG_FAST if fast else 0,
                                          C.byref(r)), "tg_mesher_autotune")
        return {
            "points_per_lane": r.points_per_lane, "waves": r.waves, "workgroups_per_cu": r.workgroups_per_cu,
            "min_prune_len": r.min_prune_len, "eval_ms": r.eval_ms,
            "candidates": [{"points_per_lane": r.candidate_points[i], "waves": r.candidate_waves[i],
                            "workgroups_per_cu": r.candidate_workgroups_per_cu[i], "ms": r.candidate_ms[i]}
                           for i in range(r.candidates)],
            "min_prune_len_ms": {str(v): r.candidate_len_ms[i] for i, v in enumerate((0, 16, 48))},
        }

    def set_tuning(self, points_per_lane: int, waves: int, min_prune_len: int = 0, workgroups_per_cu: int = 0) -> None:
        """Block-evaluation launch: workgroups_per_cu 0 = a workgroup per
        block (LDS tables), 1..16 = persistent workgroups (global tables)."""
        check(self.lib.tg_mesher_set_tuning(self.h, int(points_per_lane), int(waves), int(workgroups_per_cu),
                                            int(min_prune_len)), "tg_mesher_set_tuning")

    def set_emission(self, mode: int) -> None:
        """Block-program emission after the pruning passes: 0 automatic (a
        wave per block on grids of at most 2048 blocks, or 4096 for programs
        of 1024 entries or more, else the walk), 1 a
        wave walking 16 blocks' entries, 2 a wave per block (same programs)."""
        check(self.lib.tg_mesher_set_emission(self.h, int(mode)), "tg_mesher_set_emission")

    def set_culled_tuning(self, points_per_lane: int, waves: int, workgroups_per_cu: int) -> None:
        """Launch shape of culled exports (kept apart from the plain export's)."""
        check(self.lib.tg_mesher_set_culled_tuning(self.h, int(points_per_lane), int(waves), int(workgroups_per_cu)),
              "tg_mesher_set_culled_tuning")

    def _bricks(self, fn, name) -> np.ndarray:
        n = C.c_uint64()
        check(fn(self.h, None, 0, C.byref(n)), name)
        out = np.zeros(n.value, dtype=np.uint8)
        if n.value:
            check(fn(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8)), n.value, C.byref(n)), name)
        return out

    def brick_flags(self) -> np.ndarray:
        """Bricks (8^3 cubes, x fastest) kept by the last culled export."""
        return self._bricks(self.lib.tg_mesher_brick_flags, "tg_mesher_brick_flags")

    def active_bricks(self) -> np.ndarray:
        """Bricks holding at least one active cube in the last export."""
        return self._bricks(self.lib.tg_mesher_active_bricks, "tg_mesher_active_bricks")

    def mesh_values(self, values, origin, step, cells) -> dict:
        g = _lib.grid_struct(origin, step, cells)
        v = np.ascontiguousarray(values, dtype=np.float32)
        check(self.lib.tg_mesher_mesh_values(self.h, v.ctypes.data_as(_lib._fp), C.byref(g)), "tg_mesher_mesh_values")
        self._cells = tuple(int(c) for c in cells)
        return self.stats()

    def stats(self) -> dict:
        st = _lib.TgExportStats()
        check(self.lib.tg_mesher_stats(self.h, C.byref(st)), "tg_mesher_stats")
        return {k: getattr(st, k) for k, _ in _lib.TgExportStats._fields_}

    def counts(self) -> tuple[int, int]:
        nv, nt = C.c_uint64(), C.c_uint64()
        check(self.lib.tg_mesher_counts(self.h, C.byref(nv), C.byref(nt)), "tg_mesher_counts")
        return nv.value, nt.value

    def mesh(self) -> tuple[np.ndarray, np.ndarray]:
        nv, nt = self.counts()
        v = np.zeros((nv, 3), dtype=np.float32)
        t = np.zeros((nt, 3), dtype=np.uint32)
        check(self.lib.tg_mesher_copy_mesh(self.h, v.ctypes.data_as(_lib._fp), t.ctypes.data_as(_lib._u32p)),
              "tg_mesher_copy_mesh")
        return v, t

    def write_stl(self, program, path, scale: float = 1.0, fast: bool = False) -> None:
        """Binary STL of the last export with gradient normals (WriteSTL)."""
        check(self.lib.tg_mesher_write_stl(self.h, program.handle, str(path).encode(), float(scale),
                                           _lib.TG_FLAG_FAST if fast else 0), "tg_mesher_write_stl")

    def write_ply(self, program, path, scale: float = 1.0, fast: bool = False) -> None:
        """Binary PLY of the last export with gradient normals (WritePLY)."""
        check(self.lib.tg_mesher_write_ply(self.h, program.handle, str(path).encode(), float(scale),
                                           _lib.TG_FLAG_FAST if fast else 0), "tg_mesher_write_ply")

    def export_sodapop(self, program, bounds, density: float = 20.0, shuffle: bool = True,
                       fast: bool = False) -> dict:
        """The interactive mesher's job (MeshingJob::NaiveSurfaceNets): its grid
        over the bounding cube of `bounds` (2x3), optional vertex shuffle,
        gradient normals (read with normals())."""
        b = (C.c_float * 6)(*[float(v) for v in np.asarray(bounds, dtype=np.float32).reshape(-1)])
        check(self.lib.tg_mesher_export_sodapop(self.h, program.handle, b, float(density), 1 if shuffle else 0,
                                                _lib.TG_FLAG_FAST if fast else 0), "tg_mesher_export_sodapop")
        g = self.grid_spec()
        self._cells = g[2]
        return self.stats()

    def export_points(self, program, lo, hi, step, refine_iterations: int = 0, fast: bool = False) -> dict:
        """Point cloud (PointCloudExportThread): centres of the cells of the grid
        lo + i step within the half-step diagonal of the surface, refined
        `refine_iterations` times along the gradient; read with mesh() (no
        triangles), written with write_ply()."""
        def f3(v):
            a = np.broadcast_to(np.asarray(v, dtype=np.float32), (3,))
            return (C.c_float * 3)(*[float(x) for x in a])
        self._cells = (0, 0, 0)  # no sample grid: grid() is refused by the library
        check(self.lib.tg_mesher_export_points(self.h, program.handle, f3(lo), f3(hi), f3(step), int(refine_iterations),
                                               _lib.TG_FLAG_FAST if fast else 0), "tg_mesher_export_points")
        return self.stats()

    def export_vox(self, program, bounds, grid_size: float, color_index: int = 0, path=None,
                   fast: bool = False) -> np.ndarray:
        """MagicaVoxel export (VoxExport): writes `path` when given and returns
        the kept cell indices (raster order) of the ceil(max - min) * grid_size
        voxel grid over `bounds` (2x3)."""
        b = (C.c_float * 6)(*[float(v) for v in np.asarray(bounds, dtype=np.float32).reshape(-1)])
        n = _lib._u64(0)
        self._cells = (0, 0, 0)
        check(self.lib.tg_mesher_export_vox(self.h, program.handle, b, float(grid_size), int(color_index),
                                            None if path is None else str(path).encode(),
                                            _lib.TG_FLAG_FAST if fast else 0, C.byref(n)), "tg_mesher_export_vox")
        out = np.zeros(n.value, dtype=np.uint32)
        if n.value:
            check(self.lib.tg_mesher_copy_vox_cells(self.h, out.ctypes.data_as(_lib._u32p)), "tg_mesher_copy_vox_cells")
        return out

    def grid_spec(self):
        """(origin, step, cells) of the last export."""
        g = _lib.TgGrid()
        check(self.lib.tg_mesher_grid_get(self.h, C.byref(g)), "tg_mesher_grid_get")
        return (g.x, g.y, g.z), (g.dx, g.dy, g.dz), (int(g.sx), int(g.sy), int(g.sz))

    def normals(self) -> np.ndarray:
        nv, _ = self.counts()
        out = np.zeros((nv, 3), dtype=np.float32)
        check(self.lib.tg_mesher_copy_normals(self.h, out.ctypes.data_as(_lib._fp)), "tg_mesher_copy_normals")
        return out

    def grid(self) -> np.ndarray:
        nx, ny, nz = (c + 1 for c in self._cells)
        out = np.zeros(nx * ny * nz, dtype=np.float32)
        check(self.lib.tg_mesher_copy_grid(self.h, out.ctypes.data_as(_lib._fp)), "tg_mesher_copy_grid")
        return out


def vox_write(path, size, cells, color_index: int = 0) -> None:
    """Writes cells (raster indices of a size[0] x size[1] x size[2] grid) as a
    MagicaVoxel file, in the order given (host only)."""
    s = np.ascontiguousarray(size, dtype=np.uint32).reshape(3)
    c = np.ascontiguousarray(cells, dtype=np.uint32).reshape(-1)
    check(_lib.load().tg_vox_write(str(path).encode(), s.ctypes.data_as(_lib._u32p), c.ctypes.data_as(_lib._u32p),
                                   c.size, int(color_index)), "tg_vox_write")


def export_grid_for_bounds(bounds: np.ndarray, step: float):
    """Grid of MeshExportThread (tangerine/export.cpp:324-337): the lower bound
    moved down by two steps, ceil((max - min) / step) cells per axis."""
    lo = np.asarray(bounds[0], dtype=np.float32) - np.float32(step) * np.float32(2.0)
    hi = np.asarray(bounds[1], dtype=np.float32)
    cells = tuple(int(math.ceil(float(np.float32((hi[i] - lo[i]) / np.float32(step))))) for i in range(3))
    return tuple(float(v) for v in lo), (float(np.float32(step)),) * 3, cells
