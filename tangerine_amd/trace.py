"""Sphere tracing and gradients on the GPU (SDFNode::RayMarch and
SDFNode::Gradient batched; tangerine/sdf_evaluator.cpp:298-353)."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check

f32 = np.float32


def _farr(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), fov_y_degrees=60.0,
            width=1920, height=1080) -> _lib.TgCamera:
    cam = _lib.TgCamera()
    check(_lib.load().tg_camera_look_at(C.byref(cam), _farr(eye), _farr(target), _farr(up), float(fov_y_degrees),
                                        int(width), int(height)), "tg_camera_look_at")
    return cam


def seeded_eye(seed: int, radius: float = 12.0) -> tuple[float, float, float]:
    """Uniform point on the sphere of `radius` (normalised Gaussian, numpy seed)."""
    v = np.random.default_rng(seed).normal(size=3)
    v = v / np.linalg.norm(v)
    return tuple(float(f32(radius * x)) for x in v)


def camera_directions(cam: _lib.TgCamera) -> np.ndarray:
    """Per-pixel unit directions with the tracer's float32 operation order
    (row-major, row 0 at the top)."""
    w, h = int(cam.width), int(cam.height)
    px = np.arange(w, dtype=f32)[None, :].repeat(h, 0)
    py = np.arange(h, dtype=f32)[:, None].repeat(w, 1)
    inv_w = f32(2.0) / f32(w)
    inv_h = f32(2.0) / f32(h)
    sx = (px + f32(0.5)) * inv_w - f32(1.0)
    sy = f32(1.0) - (py + f32(0.5)) * inv_h
    x = sx * f32(cam.tan_half_x)
    y = sy * f32(cam.tan_half_y)
    fw, rt, up = ([f32(v) for v in getattr(cam, k)] for k in ("forward", "right", "up"))
    v = [(fw[i] + x * rt[i]) + y * up[i] for i in range(3)]
    inv = f32(1.0) / np.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
    return np.stack([v[0] * inv, v[1] * inv, v[2] * inv], axis=-1).reshape(-1, 3).astype(f32)


class Tracer:
    """Owns a device stream and image buffers (tg_tracer)."""

    def __init__(self):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.tg_tracer_create(C.byref(h)), "tg_tracer_create")
        self.h = h.value
        self.shape = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.tg_tracer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_passes(self, first_budget: int, budget: int, resume_waves: int = 65536) -> None:
        """Steps of the first marching pass and of each later one, and the waves
        a later pass spreads its rays over (first_budget <= 0: one pass)."""
        check(self.lib.tg_tracer_set_passes(self.h, int(first_budget), int(budget), int(resume_waves)),
              "tg_tracer_set_passes")

    def set_domain(self, lo, size: float) -> None:
        check(self.lib.tg_tracer_set_domain(self.h, _farr(lo), float(size)), "tg_tracer_set_domain")

    def render(self, program, cam, max_steps=256, eps=1e-4, far=40.0, normal_h=1e-3,
               fast=False, prune=True) -> dict:
        flags = (_lib.TG_FLAG_FAST if fast else 0) | (0 if prune else _lib.TG_FLAG_NO_PRUNE)
        check(self.lib.tg_tracer_render(self.h, program.handle, C.byref(cam), int(max_steps), float(eps), float(far),
                                        float(normal_h), flags), "tg_tracer_render")
        self.shape = (int(cam.height), int(cam.width))
        st = _lib.TgTraceStats()
        check(self.lib.tg_tracer_stats(self.h, C.byref(st)), "tg_tracer_stats")
        return {k: getattr(st, k) for k, _ in _lib.TgTraceStats._fields_}

    def image(self):
        n = self.shape[0] * self.shape[1]
        t = np.zeros(n, dtype=f32)
        nrm = np.zeros((n, 3), dtype=f32)
        steps = np.zeros(n, dtype=np.uint32)
        check(self.lib.tg_tracer_copy(self.h, t.ctypes.data_as(_lib._fp), nrm.ctypes.data_as(_lib._fp),
                                      steps.ctypes.data_as(_lib._u32p)), "tg_tracer_copy")
        return t, nrm, steps


def raymarch(program, starts, dirs, max_steps=100, eps=1e-3, far=math.inf, fast=False):
    """RayMarchTree batched: returns (hit, travel, position)."""
    s = np.ascontiguousarray(starts, dtype=f32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype=f32).reshape(-1, 3)
    n = s.shape[0]
    hit = np.zeros(n, dtype=np.uint8)
    t = np.zeros(n, dtype=f32)
    pos = np.zeros((n, 3), dtype=f32)
    check(_lib.load().tg_raymarch(program.handle, s.ctypes.data_as(_lib._fp), d.ctypes.data_as(_lib._fp), n,
                                  int(max_steps), float(eps), float(far), _lib.TG_FLAG_FAST if fast else 0,
                                  hit.ctypes.data_as(C.POINTER(C.c_uint8)), t.ctypes.data_as(_lib._fp),
                                  pos.ctypes.data_as(_lib._fp)), "tg_raymarch")
    return hit.astype(bool), t, pos


def gradient(program, points, h=1e-4, central=False, fast=False) -> np.ndarray:
    p = np.ascontiguousarray(points, dtype=f32).reshape(-1, 3)
    out = np.zeros_like(p)
    check(_lib.load().tg_gradient_points(program.handle, p.ctypes.data_as(_lib._fp), p.shape[0], float(h),
                                         1 if central else 0, out.ctypes.data_as(_lib._fp),
                                         _lib.TG_FLAG_FAST if fast else 0), "tg_gradient_points")
    return out


def pivot_towards(program, pivots, headings, downs, max_dist, margin, max_angle_deg, max_iterations=1 << 16,
                  fast=False):
    """Lua `pivot_towards` (lua_sdf.cpp:452-559) batched: (status, tails, headings).
    Status 0 done, 2 heading antiparallel to down (no output), 3 iteration cap."""
    p = np.ascontiguousarray(pivots, dtype=f32).reshape(-1, 3)
    h = np.ascontiguousarray(headings, dtype=f32).reshape(-1, 3)
    d = np.ascontiguousarray(downs, dtype=f32).reshape(-1, 3)
    n = p.shape[0]
    tails = np.zeros((n, 3), dtype=f32)
    heads = np.zeros((n, 3), dtype=f32)
    status = np.zeros(n, dtype=np.int32)
    check(_lib.load().tg_pivot_towards(program.handle, n, p.ctypes.data_as(_lib._fp), h.ctypes.data_as(_lib._fp),
                                       d.ctypes.data_as(_lib._fp), float(max_dist), float(margin),
                                       float(max_angle_deg), int(max_iterations), tails.ctypes.data_as(_lib._fp),
                                       heads.ctypes.data_as(_lib._fp), status.ctypes.data_as(C.POINTER(C.c_int32)),
                                       _lib.TG_FLAG_FAST if fast else 0), "tg_pivot_towards")
    return status, tails, heads


def pick(programs, transforms, starts, dirs, fast=False):
    """Mouse picking (DeliverMouseButton over SDFModel::RayMarch) for a batch of
    rays: (model index or -1, local travel, local hit position)."""
    progs = (C.c_void_p * max(len(programs), 1))(*[pr.handle for pr in programs])
    xf = np.ascontiguousarray(transforms, dtype=f32).reshape(-1, 8)
    s = np.ascontiguousarray(starts, dtype=f32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype=f32).reshape(-1, 3)
    n = s.shape[0]
    model = np.zeros(n, dtype=np.int32)
    travel = np.zeros(n, dtype=f32)
    pos = np.zeros((n, 3), dtype=f32)
    check(_lib.load().tg_pick(C.cast(progs, C.POINTER(_lib._p)), xf.ctypes.data_as(_lib._fp), len(programs),
                              s.ctypes.data_as(_lib._fp), d.ctypes.data_as(_lib._fp), n,
                              _lib.TG_FLAG_FAST if fast else 0, model.ctypes.data_as(C.POINTER(C.c_int32)),
                              travel.ctypes.data_as(_lib._fp), pos.ctypes.data_as(_lib._fp)), "tg_pick")
    return model, travel, pos
