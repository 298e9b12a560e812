"""Learned brick culling for grid export (tg_surrogate_*; no reference
counterpart).

A Fourier-feature MLP fitted on the device to the clamped signed distance of
a program over a cube domain.  ``Mesher.export_culled`` uses it to skip the
exact evaluation of bricks that it certifies, with a margin calibrated on
held-out samples, to lie farther from the surface than their half-diagonal.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

FREQ = 48
PARAMS = 10433


class Surrogate:
    def __init__(self, lo, size: float, sigma: float = 2.0, clamp: float = 0.5, seed: int = 1):
        self.lib = _lib.load()
        h = C.c_void_p()
        lo_a = (C.c_float * 3)(*[float(v) for v in lo])
        check(self.lib.tg_surrogate_create(C.byref(h), lo_a, float(size), float(sigma), float(clamp), int(seed)),
              "tg_surrogate_create")
        self.h = h.value
        self.lo = tuple(float(v) for v in lo)
        self.size = float(size)
        self.clamp = float(clamp)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.tg_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> tuple[np.ndarray, np.ndarray]:
        p = np.zeros(PARAMS, dtype=np.float32)
        f = np.zeros((FREQ, 3), dtype=np.float32)
        check(self.lib.tg_surrogate_params(self.h, p.ctypes.data_as(_lib._fp), f.ctypes.data_as(_lib._fp), 0),
              "tg_surrogate_params")
        return p, f

    def set_params(self, params=None, fourier=None) -> None:
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        f = None if fourier is None else np.ascontiguousarray(fourier, dtype=np.float32)
        check(self.lib.tg_surrogate_params(self.h, None if p is None else p.ctypes.data_as(_lib._fp),
                                           None if f is None else f.ctypes.data_as(_lib._fp), 1),
              "tg_surrogate_params")

    def predict(self, xyz) -> np.ndarray:
        x = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(x.shape[0], dtype=np.float32)
        check(self.lib.tg_surrogate_predict(self.h, x.ctypes.data_as(_lib._fp), x.shape[0],
                                            out.ctypes.data_as(_lib._fp)), "tg_surrogate_predict")
        return out

    def gradients(self, xyz, target) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(PARAMS, dtype=np.float32)
        loss = C.c_float()
        check(self.lib.tg_surrogate_gradients(self.h, x.ctypes.data_as(_lib._fp), t.ctypes.data_as(_lib._fp),
                                              x.shape[0], g.ctypes.data_as(_lib._fp), C.byref(loss)),
              "tg_surrogate_gradients")
        return loss.value, g

    def fit(self, program, samples: int = 1 << 20, calib_samples: int = 200_000, surface_sigma: float = 0.05,
            epochs: int = 20, fast: bool = False) -> dict:
        st = _lib.TgSurrogateStats()
        check(self.lib.tg_surrogate_fit(self.h, program.handle, int(samples), int(calib_samples),
                                        float(surface_sigma), int(epochs), _lib.TG_FLAG_FAST if fast else 0,
                                        C.byref(st)), "tg_surrogate_fit")
        return {k: getattr(st, k) for k, _ in _lib.TgSurrogateStats._fields_}

    def set_margin(self, margin: float) -> None:
        check(self.lib.tg_surrogate_set_margin(self.h, float(margin)), "tg_surrogate_set_margin")

    def loss_log(self) -> np.ndarray:
        n = C.c_uint64()
        check(self.lib.tg_surrogate_loss_log(self.h, None, 0, C.byref(n)), "tg_surrogate_loss_log")
        out = np.zeros(n.value, dtype=np.float32)
        if n.value:
            check(self.lib.tg_surrogate_loss_log(self.h, out.ctypes.data_as(_lib._fp), n.value, C.byref(n)),
                  "tg_surrogate_loss_log")
        return out
