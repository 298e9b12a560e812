"""ctypes access to the CPU checkers (test infrastructure only).

  Oracle()      -> oracle/libtgoracle.so, this project's C++ restatement of
                   the reference path (tg_oracle.cpp).
  Reference()   -> oracle/_ref/libtgref.so, the reference's own sources built
                   by oracle/Makefile (ref_harness.cpp).  Present only where it
                   was built.

Both expose the same methods, so tests can run one check against either.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "libtgoracle.so"
REF_SO = HERE / "_ref" / "libtgref.so"

_f = C.c_float
_p = C.c_void_p
_u64 = C.c_uint64
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i32p = C.POINTER(C.c_int32)

_OPS = {"union": 0, "inter": 1, "diff": 2, "blend_union": 3, "blend_inter": 4, "blend_diff": 5}


def _fptr(a):
    return a.ctypes.data_as(_fp)


class _Cpu:
    prefix = ""

    def __init__(self, path: Path):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        self.lib = C.CDLL(str(path))
        p = self.prefix
        sig = {
            "sphere": (_p, [_f]), "ellipsoid": (_p, [_f, _f, _f]), "box": (_p, [_f, _f, _f]),
            "torus": (_p, [_f, _f]), "cylinder": (_p, [_f, _f]), "plane": (_p, [_f, _f, _f]),
            "cone": (_p, [_f, _f]), "coninder": (_p, [_f, _f, _f]),
            "op": (_p, [C.c_int, _f, _p, _p]), "flate": (_p, [_p, _f]),
            "move": (_p, [_p, _f, _f, _f]), "rotate": (_p, [_p, _f, _f, _f, _f]),
            "rotate_axis": (_p, [_p, C.c_int, _f]), "scale": (_p, [_p, _f]),
            "free": (None, [_p]), "leaf_count": (C.c_int, [_p]), "stack_size": (_u64, [_p]),
            "bounds": (None, [_p, _fp]), "eval": (None, [_p, _fp, _u64, _fp]),
            "compile": (_u64, [_p, _u32p, _u64]), "clip": (_p, [_p, _f, _f, _f, _f]),
            "gradient": (None, [_p, _fp, _u64, _fp]),
            "raymarch": (None, [_p, _fp, _fp, _u64, C.c_int, _f, _u8p, _fp, _fp]),
            "mesh_vertex_count": (_u64, [_p]), "mesh_face_count": (_u64, [_p]),
            "mesh_copy": (None, [_p, _fp, _u64p]), "mesh_free": (None, [_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(self.lib, p + name)
            fn.restype, fn.argtypes = res, args
        self._extra_signatures()

    def _extra_signatures(self):
        pass

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    # ---- scenes.Backend interface -----------------------------------------
    def brush(self, kind, *params):
        return self.fn(kind)(*[float(v) for v in params])

    def op(self, kind, t, a, b):
        return self.fn("op")(_OPS[kind], float(t), a, b)

    def flate(self, a, r):
        return self.fn("flate")(a, float(r))

    def move(self, a, x, y, z):
        return self.fn("move")(a, float(x), float(y), float(z))

    def rotate(self, a, qx, qy, qz, qw):
        return self.fn("rotate")(a, float(qx), float(qy), float(qz), float(qw))

    def rotate_axis(self, a, axis, deg):
        return self.fn("rotate_axis")(a, int(axis), float(deg))

    def scale(self, a, s):
        return self.fn("scale")(a, float(s))

    def free(self, a):
        if a:
            self.fn("free")(a)

    # ---- queries -------------------------------------------------------------
    def build(self, spec):
        import tangerine_amd.scenes as scenes  # spec builder only (no GPU use)
        return scenes.build(spec, self)

    def eval(self, h, points) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty(pts.shape[0], dtype=np.float32)
        self.fn("eval")(h, _fptr(pts), pts.shape[0], _fptr(out))
        return out

    def compile(self, h) -> np.ndarray:
        n = self.fn("compile")(h, None, 0)
        w = np.zeros(n, dtype=np.uint32)
        self.fn("compile")(h, w.ctypes.data_as(_u32p), n)
        return w

    def bounds(self, h) -> np.ndarray:
        out = np.zeros(6, dtype=np.float32)
        self.fn("bounds")(h, _fptr(out))
        return out.reshape(2, 3)

    def gradient(self, h, points) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(pts)
        self.fn("gradient")(h, _fptr(pts), pts.shape[0], _fptr(out))
        return out

    def raymarch(self, h, starts, dirs, iterations=100, eps=1e-3):
        s = np.ascontiguousarray(starts, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = s.shape[0]
        hit = np.zeros(n, dtype=np.uint8)
        t = np.zeros(n, dtype=np.float32)
        pos = np.zeros((n, 3), dtype=np.float32)
        self.fn("raymarch")(h, _fptr(s), _fptr(d), n, int(iterations), float(eps),
                            hit.ctypes.data_as(_u8p), _fptr(t), _fptr(pos))
        return hit.astype(bool), t, pos

    def _take_mesh(self, m):
        nv = self.fn("mesh_vertex_count")(m)
        nf = self.fn("mesh_face_count")(m)
        v = np.zeros((nv, 3), dtype=np.float32)
        f = np.zeros((nf, 3), dtype=np.uint64)
        self.fn("mesh_copy")(m, _fptr(v), f.ctypes.data_as(_u64p))
        self.fn("mesh_free")(m)
        return v, f


class Oracle(_Cpu):
    prefix = "tgo_"

    def __init__(self, path: Path = ORACLE_SO):
        super().__init__(path)

    def _extra_signatures(self):
        L = self.lib
        L.tgo_interp_words.restype = C.c_int
        L.tgo_interp_words.argtypes = [_u32p, _u64, _fp, _u64, _fp]
        L.tgo_grid_interp.restype = C.c_int
        L.tgo_grid_interp.argtypes = [_u32p, _u64, _fp, _fp, _u64p, _u64, _u64, C.c_int, _fp]
        L.tgo_gradient_central_words.restype = C.c_int
        L.tgo_gradient_central_words.argtypes = [_u32p, _u64, _fp, _u64, _f, _fp]
        L.tgo_gradient_tetra_words.restype = C.c_int
        L.tgo_gradient_tetra_words.argtypes = [_u32p, _u64, _fp, _u64, _f, _fp]
        L.tgo_raymarch_far_words.restype = C.c_int
        L.tgo_raymarch_far_words.argtypes = [_u32p, _u64, _fp, _fp, _u64, C.c_int, _f, _f, _u8p, _fp, _i32p]
        L.tgo_surface_nets.restype = _p
        L.tgo_surface_nets.argtypes = [_fp, _fp, _fp, _u64p]

    def interp_words(self, words, points) -> np.ndarray:
        w = np.ascontiguousarray(words, dtype=np.uint32)
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty(pts.shape[0], dtype=np.float32)
        rc = self.lib.tgo_interp_words(w.ctypes.data_as(_u32p), w.size, _fptr(pts), pts.shape[0], _fptr(out))
        if rc != 0:
            raise ValueError("malformed program")
        return out

    def grid_interp(self, words, origin, step, cells, k0=0, k1=None, threads=0) -> tuple[np.ndarray, int]:
        w = np.ascontiguousarray(words, dtype=np.uint32)
        o = np.array(origin, dtype=np.float32)
        s = np.array(step, dtype=np.float32)
        c = np.array(cells, dtype=np.uint64)
        k1 = int(c[2]) + 1 if k1 is None else k1
        out = np.empty((k1 - k0, int(c[1]) + 1, int(c[0]) + 1), dtype=np.float32)
        used = self.lib.tgo_grid_interp(w.ctypes.data_as(_u32p), w.size, _fptr(o), _fptr(s),
                                        c.ctypes.data_as(_u64p), k0, k1, int(threads), _fptr(out))
        if used < 0:
            raise ValueError("malformed program")
        return out, used

    def gradient_central(self, words, points, h=1e-3) -> np.ndarray:
        w = np.ascontiguousarray(words, dtype=np.uint32)
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(pts)
        if self.lib.tgo_gradient_central_words(w.ctypes.data_as(_u32p), w.size, _fptr(pts), pts.shape[0],
                                               float(h), _fptr(out)) != 0:
            raise ValueError("malformed program")
        return out

    def gradient_tetra(self, words, points, h=1e-4) -> np.ndarray:
        w = np.ascontiguousarray(words, dtype=np.uint32)
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(pts)
        if self.lib.tgo_gradient_tetra_words(w.ctypes.data_as(_u32p), w.size, _fptr(pts), pts.shape[0], float(h),
                                             _fptr(out)) != 0:
            raise ValueError("malformed program")
        return out

    def raymarch_far(self, words, starts, dirs, steps=256, eps=1e-4, far=40.0):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        s = np.ascontiguousarray(starts, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = s.shape[0]
        hit = np.zeros(n, dtype=np.uint8)
        t = np.zeros(n, dtype=np.float32)
        it = np.zeros(n, dtype=np.int32)
        if self.lib.tgo_raymarch_far_words(w.ctypes.data_as(_u32p), w.size, _fptr(s), _fptr(d), n, int(steps),
                                           float(eps), float(far), hit.ctypes.data_as(_u8p), _fptr(t),
                                           it.ctypes.data_as(_i32p)) != 0:
            raise ValueError("malformed program")
        return hit.astype(bool), t, it

    def surface_nets(self, values, origin, step, cells):
        v = np.ascontiguousarray(values, dtype=np.float32)
        o = np.array(origin, dtype=np.float32)
        s = np.array(step, dtype=np.float32)
        c = np.array(cells, dtype=np.uint64)
        m = self.lib.tgo_surface_nets(_fptr(v), _fptr(o), _fptr(s), c.ctypes.data_as(_u64p))
        return self._take_mesh(m)

    def point_cloud(self, words, lo, hi, step, iterations=0) -> np.ndarray:
        """PointCloudExportThread's points (raster order) over the whole program."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        f3 = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float32), (3,))) for v in (lo, hi, step)]
        L = self.lib
        L.tgo_point_cloud.restype = C.c_int64
        L.tgo_point_cloud.argtypes = [_u32p, _u64, _fp, _fp, _fp, C.c_int, _fp, _u64]
        n = L.tgo_point_cloud(w.ctypes.data_as(_u32p), w.size, *[_fptr(v) for v in f3], int(iterations), None, 0)
        if n < 0:
            raise ValueError("malformed program")
        out = np.zeros((n, 3), dtype=np.float32)
        L.tgo_point_cloud(w.ctypes.data_as(_u32p), w.size, *[_fptr(v) for v in f3], int(iterations), _fptr(out), n)
        return out

    def pivot_towards(self, words, max_dist, margin, max_angle_deg, pivot, heading, down, max_iterations=1 << 16):
        """LuaPivotTowards over the interpreter: (status, tail, heading)."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        v = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(3)) for a in (pivot, heading, down)]
        tail, head = np.zeros(3, np.float32), np.zeros(3, np.float32)
        L = self.lib
        L.tgo_pivot_towards_words.restype = C.c_int
        L.tgo_pivot_towards_words.argtypes = [_u32p, _u64, _f, _f, _f, _fp, _fp, _fp, C.c_int, _fp, _fp]
        st = L.tgo_pivot_towards_words(w.ctypes.data_as(_u32p), w.size, float(max_dist), float(margin),
                                       float(max_angle_deg), *[_fptr(a) for a in v], int(max_iterations),
                                       _fptr(tail), _fptr(head))
        return st, tail, head

    def pick(self, programs, transforms, start, direction):
        """Nearest hit over models (DeliverMouseButton): (model, travel, local position)."""
        words = [np.ascontiguousarray(p, dtype=np.uint32) for p in programs]
        offs = np.zeros(len(words) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([w.size for w in words])
        allw = np.ascontiguousarray(np.concatenate(words) if words else np.zeros(0, np.uint32))
        xf = np.ascontiguousarray(np.asarray(transforms, dtype=np.float32).reshape(-1, 8))
        s = np.ascontiguousarray(np.asarray(start, dtype=np.float32).reshape(3))
        d = np.ascontiguousarray(np.asarray(direction, dtype=np.float32).reshape(3))
        tr, pos = np.zeros(1, np.float32), np.zeros(3, np.float32)
        L = self.lib
        L.tgo_pick_words.restype = C.c_int
        L.tgo_pick_words.argtypes = [_u32p, _u64p, _fp, C.c_uint32, _fp, _fp, _fp, _fp]
        m = L.tgo_pick_words(allw.ctypes.data_as(_u32p), offs.ctypes.data_as(_u64p), _fptr(xf), len(words), _fptr(s),
                             _fptr(d), _fptr(tr), _fptr(pos))
        return m, float(tr[0]), pos

    def vox_cells(self, words, bounds, grid_size):
        """VoxExport's selection over the interpreter: (size, raster cell indices)."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        b = np.asarray(bounds, dtype=np.float32).reshape(-1)
        mn, mx = np.ascontiguousarray(b[:3]), np.ascontiguousarray(b[3:])
        size = np.zeros(3, dtype=np.int32)
        L = self.lib
        L.tgo_vox_cells.restype = C.c_int64
        L.tgo_vox_cells.argtypes = [_u32p, _u64, _fp, _fp, _f, C.POINTER(C.c_int32), _u32p, _u64]
        n = L.tgo_vox_cells(w.ctypes.data_as(_u32p), w.size, _fptr(mn), _fptr(mx), float(grid_size),
                            size.ctypes.data_as(C.POINTER(C.c_int32)), None, 0)
        if n < 0:
            return size, np.zeros(0, dtype=np.uint32)
        out = np.zeros(n, dtype=np.uint32)
        L.tgo_vox_cells(w.ctypes.data_as(_u32p), w.size, _fptr(mn), _fptr(mx), float(grid_size),
                        size.ctypes.data_as(C.POINTER(C.c_int32)), out.ctypes.data_as(_u32p), n)
        return size, out

    def mt19937_10000th(self) -> int:
        self.lib.tgo_mt19937_10000th.restype = C.c_uint32
        return int(self.lib.tgo_mt19937_10000th())

    def vertex_sequence(self, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.uint32)
        self.lib.tgo_vertex_sequence.argtypes = [C.c_uint64, _u32p]
        self.lib.tgo_vertex_sequence(int(n), out.ctypes.data_as(_u32p))
        return out


class Reference(_Cpu):
    prefix = "tgref_"

    def __init__(self, path: Path = REF_SO):
        super().__init__(path)

    def _extra_signatures(self):
        L = self.lib
        L.tgref_interp_eval.restype = None
        L.tgref_interp_eval.argtypes = [_p, _fp, _u64, _fp]
        L.tgref_octree_create.restype = _p
        L.tgref_octree_create.argtypes = [_p, _f, C.c_int, C.c_int, _f]
        L.tgref_octree_free.restype = None
        L.tgref_octree_free.argtypes = [_p]
        L.tgref_octree_eval.restype = None
        L.tgref_octree_eval.argtypes = [_p, _fp, _u64, C.c_int, _fp]
        L.tgref_pool_grid_eval.restype = C.c_int
        L.tgref_pool_grid_eval.argtypes = [_p, _fp, _fp, _u64p, _u64, _u64, _fp]
        L.tgref_surface_nets_values.restype = _p
        L.tgref_surface_nets_values.argtypes = [_fp, _fp, _fp, _u64p]
        L.tgref_export_threaded.restype = _p
        L.tgref_export_threaded.argtypes = [_p, _fp, _fp, _u64p, _u64, _u64, C.c_int]
        L.tgref_mesh_evals.restype = _u64
        L.tgref_mesh_evals.argtypes = [_p]
        L.tgref_threads_default.restype = C.c_int
        L.tgref_hardware_concurrency.restype = C.c_int

    def pivot_towards(self, h, max_dist, margin, max_angle_deg, pivot, heading, down):
        v = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(3)) for a in (pivot, heading, down)]
        tail, head = np.zeros(3, np.float32), np.zeros(3, np.float32)
        L = self.lib
        L.tgref_pivot_towards.restype = C.c_int
        L.tgref_pivot_towards.argtypes = [_p, _f, _f, _f, _fp, _fp, _fp, _fp, _fp]
        st = L.tgref_pivot_towards(h, float(max_dist), float(margin), float(max_angle_deg), *[_fptr(a) for a in v],
                                   _fptr(tail), _fptr(head))
        return st, tail, head

    def pick(self, handles, transforms, start, direction):
        hs = (C.c_void_p * len(handles))(*handles)
        xf = np.ascontiguousarray(np.asarray(transforms, dtype=np.float32).reshape(-1, 8))
        s = np.ascontiguousarray(np.asarray(start, dtype=np.float32).reshape(3))
        d = np.ascontiguousarray(np.asarray(direction, dtype=np.float32).reshape(3))
        tr, pos = np.zeros(1, np.float32), np.zeros(3, np.float32)
        L = self.lib
        L.tgref_pick.restype = C.c_int
        L.tgref_pick.argtypes = [C.c_void_p, _fp, C.c_uint32, _fp, _fp, _fp, _fp]
        m = L.tgref_pick(C.cast(hs, C.c_void_p), _fptr(xf), len(handles), _fptr(s), _fptr(d), _fptr(tr), _fptr(pos))
        return m, float(tr[0]), pos

    def vox_write(self, size, cells, color_index, path) -> None:
        s = np.ascontiguousarray(size, dtype=np.uint32).reshape(3)
        c = np.ascontiguousarray(cells, dtype=np.uint32).reshape(-1)
        L = self.lib
        L.tgref_vox_write.restype = None
        L.tgref_vox_write.argtypes = [_u32p, _u32p, _u64, C.c_int, C.c_char_p]
        L.tgref_vox_write(s.ctypes.data_as(_u32p), c.ctypes.data_as(_u32p), c.size, int(color_index), str(path).encode())

    def vox_export(self, h, grid_size, color_index, path) -> int:
        L = self.lib
        L.tgref_vox_export.restype = C.c_int64
        L.tgref_vox_export.argtypes = [_p, _f, C.c_int, C.c_char_p]
        return int(L.tgref_vox_export(h, float(grid_size), int(color_index), str(path).encode()))

    def point_cloud(self, h, lo, hi, step, iterations=0) -> np.ndarray:
        f3 = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float32), (3,))) for v in (lo, hi, step)]
        L = self.lib
        L.tgref_point_cloud.restype = C.c_int64
        L.tgref_point_cloud.argtypes = [_p, _fp, _fp, _fp, C.c_int, _fp, _u64]
        n = L.tgref_point_cloud(h, *[_fptr(v) for v in f3], int(iterations), None, 0)
        out = np.zeros((n, 3), dtype=np.float32)
        L.tgref_point_cloud(h, *[_fptr(v) for v in f3], int(iterations), _fptr(out), n)
        return out

    def interp_eval(self, h, points) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty(pts.shape[0], dtype=np.float32)
        self.lib.tgref_interp_eval(h, _fptr(pts), pts.shape[0], _fptr(out))
        return out

    def surface_nets(self, values, origin, step, cells):
        v = np.ascontiguousarray(values, dtype=np.float32)
        o = np.array(origin, dtype=np.float32)
        s = np.array(step, dtype=np.float32)
        c = np.array(cells, dtype=np.uint64)
        m = self.lib.tgref_surface_nets_values(_fptr(v), _fptr(o), _fptr(s), c.ctypes.data_as(_u64p))
        return self._take_mesh(m)

    def octree(self, h, target=0.25, coalesce=True, max_depth=-1, margin=0.0):
        return self.lib.tgref_octree_create(h, float(target), int(coalesce), int(max_depth), float(margin))

    def octree_eval(self, o, points, exact=True) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty(pts.shape[0], dtype=np.float32)
        self.lib.tgref_octree_eval(o, _fptr(pts), pts.shape[0], int(exact), _fptr(out))
        return out

    def pool_grid_eval(self, o, origin, step, cells, k0, k1):
        oo = np.array(origin, dtype=np.float32)
        s = np.array(step, dtype=np.float32)
        c = np.array(cells, dtype=np.uint64)
        out = np.empty((k1 - k0, int(c[1]) + 1, int(c[0]) + 1), dtype=np.float32)
        threads = self.lib.tgref_pool_grid_eval(o, _fptr(oo), _fptr(s), c.ctypes.data_as(_u64p), k0, k1, _fptr(out))
        return out, threads

    def export_threaded(self, o, origin, step, cells, k0=0, k1=None, threads=0):
        oo = np.array(origin, dtype=np.float32)
        s = np.array(step, dtype=np.float32)
        c = np.array(cells, dtype=np.uint64)
        k1 = int(c[2]) if k1 is None else k1
        m = self.lib.tgref_export_threaded(o, _fptr(oo), _fptr(s), c.ctypes.data_as(_u64p), k0, k1, int(threads))
        evals = int(self.lib.tgref_mesh_evals(m))
        v, f = self._take_mesh(m)
        return v, f, evals
