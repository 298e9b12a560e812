"""Python mirror of Tangerine's Lua `tangerine.sdf` surface on the GPU path.

Builders, set operators and transforms keep the Lua bindings' names, argument
meaning (diameters, threshold last) and copy-then-transform behaviour
(tangerine/lua_sdf.cpp:56-448, table at :704-761).  Queries run on the GPU
through libtangerine_gpu.so; there is no CPU evaluation path.

Out of scope (shading and application glue, not distance evaluation): paint,
paint_over, stencil, mask, set_bg, set_fixed_camera, random_seed, random,
shuffle_sequence, instance, pivot_towards.
"""

from __future__ import annotations

import ctypes as C
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._lib import TG_FLAG_FAST, TangerineError, check

f32 = np.float32


def _vec3(args: Sequence) -> tuple[float, float, float]:
    """Lua GetVec3 (tangerine/lua_vec.cpp:33-50): a 3-vector or 3 numbers, as floats."""
    if len(args) == 1:
        v = args[0]
        if isinstance(v, SDF):
            raise TypeError("expected a vector")
        x, y, z = v
    elif len(args) == 3:
        x, y, z = args
    else:
        raise TypeError("expected a vector or three numbers")
    return f32(x), f32(y), f32(z)


class SDF:
    """One immutable CSG tree (a tg_node handle)."""

    __slots__ = ("_h", "_program")

    def __init__(self, handle: int):
        if not handle:
            raise TangerineError(_lib.last_error() or "null tree")
        self._h = handle
        self._program = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().tg_node_release(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self) -> int:
        return self._h

    # ---- transforms (each returns a transformed copy) --------------------
    def move(self, *args) -> "SDF":
        x, y, z = _vec3(args)
        return SDF(_lib.load().tg_move(self._h, x, y, z))

    def move_x(self, x) -> "SDF":
        return self.move(x, 0.0, 0.0)

    def move_y(self, y) -> "SDF":
        return self.move(0.0, y, 0.0)

    def move_z(self, z) -> "SDF":
        return self.move(0.0, 0.0, z)

    def rotate(self, x, y, z, w) -> "SDF":
        return SDF(_lib.load().tg_rotate(self._h, f32(x), f32(y), f32(z), f32(w)))

    def rotate_x(self, degrees) -> "SDF":
        return SDF(_lib.load().tg_rotate_x(self._h, f32(degrees)))

    def rotate_y(self, degrees) -> "SDF":
        return SDF(_lib.load().tg_rotate_y(self._h, f32(degrees)))

    def rotate_z(self, degrees) -> "SDF":
        return SDF(_lib.load().tg_rotate_z(self._h, f32(degrees)))

    def scale(self, s) -> "SDF":
        return SDF(_lib.load().tg_scale(self._h, f32(s)))

    def flate(self, diameter) -> "SDF":
        # LuaFlate: radius = diameter * .5 (lua_sdf.cpp:173-180)
        return SDF(_lib.load().tg_flate(self._h, f32(float(diameter) * 0.5)))

    # ---- set operators as methods (Lua `a:union(b)`) ----------------------
    def union(self, *others) -> "SDF":
        return union(self, *others)

    def inter(self, *others) -> "SDF":
        return inter(self, *others)

    def diff(self, *others) -> "SDF":
        return diff(self, *others)

    def blend_union(self, *args) -> "SDF":
        return blend_union(self, *args)

    def blend_inter(self, *args) -> "SDF":
        return blend_inter(self, *args)

    def blend_diff(self, *args) -> "SDF":
        return blend_diff(self, *args)

    # ---- host-side structure ---------------------------------------------
    def bounds(self) -> np.ndarray:
        out = (C.c_float * 6)()
        check(_lib.load().tg_node_bounds(self._h, out), "tg_node_bounds")
        return np.array(out[:], dtype=np.float32).reshape(2, 3)

    def has_finite_bounds(self) -> bool:
        return bool(_lib.load().tg_node_has_finite_bounds(self._h))

    def leaf_count(self) -> int:
        return int(_lib.load().tg_node_leaf_count(self._h))

    def stack_size(self) -> int:
        return int(_lib.load().tg_node_stack_size(self._h))

    def compile(self) -> np.ndarray:
        """Reference ProgramBuffer words (uint32), Stop included."""
        lib = _lib.load()
        n = C.c_uint64(0)
        check(lib.tg_node_compile(self._h, None, 0, C.byref(n)), "tg_node_compile")
        words = np.zeros(n.value, dtype=np.uint32)
        check(lib.tg_node_compile(self._h, words.ctypes.data_as(_lib._u32p), n.value, C.byref(n)),
              "tg_node_compile")
        return words

    def inspect(self) -> dict:
        """Program properties from the host decoder (no device needed)."""
        return inspect_words(self.compile())

    def program(self) -> "Program":
        if self._program is None:
            self._program = Program.from_node(self)
        return self._program

    # ---- queries (GPU) ----------------------------------------------------
    def eval(self, *args, fast: bool = False) -> float:
        """Lua `eval(p)` (lua_sdf.cpp:386-395) for one point."""
        p = np.array([_vec3(args)], dtype=np.float32)
        return float(self.program().eval_points(p, fast=fast)[0])

    def eval_points(self, points, fast: bool = False) -> np.ndarray:
        return self.program().eval_points(points, fast=fast)

    def gradient(self, *args, fast: bool = False) -> np.ndarray:
        """Lua `gradient(p)` (lua_sdf.cpp:398-407): SDFNode::Gradient's tetrahedral rule."""
        from .trace import gradient
        return gradient(self.program(), np.array([_vec3(args)], dtype=np.float32), h=1e-4, fast=fast)[0]

    def ray_cast(self, origin, direction, max_iterations: int = 100, epsilon: float = 0.001, fast: bool = False):
        """Lua `ray_cast(origin, dir, [iters], [eps])` (lua_sdf.cpp:410-448): hit position or None."""
        from .trace import raymarch
        hit, _, pos = raymarch(self.program(), [_vec3((origin,))], [_vec3((direction,))], max_iterations, epsilon,
                               fast=fast)
        return pos[0] if hit[0] else None

    def magnet(self, origin, target, max_iterations: int = 100, epsilon: float = 0.001, fast: bool = False):
        """Lua `magnet(origin, target, ...)`: ray_cast toward normalize(target - origin)."""
        o = np.array(_vec3((origin,)), dtype=np.float32)
        d = np.array(_vec3((target,)), dtype=np.float32) - o
        d = d * (np.float32(1.0) / np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
        return self.ray_cast(o, d, max_iterations, epsilon, fast)

    def pivot_towards(self, max_dist, margin, max_angle, pivot, heading, down, fast: bool = False):
        """Lua `pivot_towards(max_dist, margin, max_angle, pivot, heading, down)`
        (lua_sdf.cpp:452-559): (tail, heading).  A heading antiparallel to down
        raises (the reference picks a random axis from the Lua generator)."""
        from .trace import pivot_towards
        st, tails, heads = pivot_towards(self.program(), [_vec3((pivot,))], [_vec3((heading,))], [_vec3((down,))],
                                         max_dist, margin, max_angle, fast=fast)
        if st[0] != 0:
            raise ValueError(f"pivot_towards: status {int(st[0])} (2: heading antiparallel to down)")
        return tails[0], heads[0]


def _info_dict(info) -> dict:
    return {
        "instructions": int(info.instructions),
        "max_stack": int(info.max_stack),
        "brushes": int(info.brushes),
        "set_ops": int(info.set_ops),
        "lipschitz": bool(info.lipschitz),
        "flops_per_eval": float(info.flops_per_eval),
        "leaf_extent": float(info.leaf_extent),
    }


def inspect_words(words) -> dict:
    w = np.ascontiguousarray(words, dtype=np.uint32)
    info = _lib.TgProgramInfo()
    check(_lib.load().tg_words_inspect(w.ctypes.data_as(_lib._u32p), w.size, C.byref(info)), "tg_words_inspect")
    return _info_dict(info)


class Program:
    """A flattened tree resident in device memory (tg_program)."""

    __slots__ = ("_h", "info")

    def __init__(self, handle: int):
        self._h = handle
        info = _lib.TgProgramInfo()
        check(_lib.load().tg_program_info_get(handle, C.byref(info)), "tg_program_info_get")
        self.info = _info_dict(info)

    @classmethod
    def from_node(cls, node: SDF) -> "Program":
        h = C.c_void_p()
        check(_lib.load().tg_program_from_node(node.handle, C.byref(h)), "tg_program_from_node")
        return cls(h.value)

    @classmethod
    def from_words(cls, words) -> "Program":
        w = np.ascontiguousarray(words, dtype=np.uint32)
        h = C.c_void_p()
        check(_lib.load().tg_program_from_words(w.ctypes.data_as(_lib._u32p), w.size, C.byref(h)),
              "tg_program_from_words")
        return cls(h.value)

    @property
    def handle(self) -> int:
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().tg_program_destroy(h)
            except Exception:
                pass
            self._h = None

    def eval_points(self, points, fast: bool = False) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty(pts.shape[0], dtype=np.float32)
        flags = TG_FLAG_FAST if fast else 0
        check(_lib.load().tg_eval_points(self._h, pts.ctypes.data_as(_lib._fp), pts.shape[0],
                                         out.ctypes.data_as(_lib._fp), flags), "tg_eval_points")
        return out


# ---- builders (Lua names and argument meaning) ------------------------------

def _wrap(h) -> SDF:
    return SDF(h)


def sphere(diameter) -> SDF:
    return _wrap(_lib.load().tg_sphere(f32(float(diameter) * 0.5)))


def ellipsoid(*args) -> SDF:
    x, y, z = (f32(v * f32(0.5)) for v in _vec3(args))
    return _wrap(_lib.load().tg_ellipsoid(x, y, z))


def box(*args) -> SDF:
    x, y, z = (f32(v * f32(0.5)) for v in _vec3(args))
    return _wrap(_lib.load().tg_box(x, y, z))


def cube(size) -> SDF:
    e = f32(float(size) * 0.5)
    return _wrap(_lib.load().tg_box(e, e, e))


def torus(major_diameter, minor_diameter) -> SDF:
    major = f32(float(major_diameter) * 0.5)
    minor = f32(float(minor_diameter) * 0.5)
    return _wrap(_lib.load().tg_torus(f32(major - minor), minor))


def cylinder(diameter, height) -> SDF:
    return _wrap(_lib.load().tg_cylinder(f32(float(diameter) * 0.5), f32(float(height) * 0.5)))


def plane(*args) -> SDF:
    x, y, z = _vec3(args)
    return _wrap(_lib.load().tg_plane(x, y, z))


def cone(diameter, height) -> SDF:
    return _wrap(_lib.load().tg_cone(f32(float(diameter) * 0.5), f32(height)))


def coninder(*args) -> SDF:
    lo, hi, h = _vec3(args)
    return _wrap(_lib.load().tg_coninder(f32(lo * f32(0.5)), f32(hi * f32(0.5)), h))


cylicone = coninder


def _fold(fn_name: str, nodes: Iterable[SDF], threshold=None) -> SDF:
    nodes = list(nodes)
    if len(nodes) < 2:
        raise TypeError("set operators take at least two trees")
    fn = getattr(_lib.load(), fn_name)
    acc = nodes[0]
    for nxt in nodes[1:]:
        acc = _wrap(fn(acc.handle, nxt.handle) if threshold is None else fn(f32(threshold), acc.handle, nxt.handle))
    return acc


def union(*nodes: SDF) -> SDF:
    return _fold("tg_union", nodes)


def inter(*nodes: SDF) -> SDF:
    return _fold("tg_inter", nodes)


def diff(*nodes: SDF) -> SDF:
    return _fold("tg_diff", nodes)


def blend_union(*args) -> SDF:
    *nodes, t = args
    return _fold("tg_blend_union", nodes, t)


def blend_inter(*args) -> SDF:
    *nodes, t = args
    return _fold("tg_blend_inter", nodes, t)


def blend_diff(*args) -> SDF:
    *nodes, t = args
    return _fold("tg_blend_diff", nodes, t)
