"""tg_raymarch_static: the ray march without device allocation.

Bar: through the C ABI, two calls in a row (the second with fewer rays, so
stale staging from the first would show) give hit flags, travel and positions
bit-identical to tg_raymarch on the same rays, in exact and fast numerics; a
count above TG_RAYMARCH_STATIC_MAX is refused with TG_ERR_LIMIT before any
device work.
"""

import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TG_ERR_LIMIT = -5
STATIC_MAX = 65536


def _rays(seed, n):
    rng = np.random.default_rng(seed)
    starts = rng.uniform(-12.0, 12.0, size=(n, 3)).astype(np.float32)
    # aim roughly at the scene so most rays hit
    target = rng.uniform(-2.0, 2.0, size=(n, 3)).astype(np.float32)
    return starts, (target - starts).astype(np.float32)


def _march(lib, fn, prog, starts, dirs, flags):
    n = starts.shape[0]
    hit = np.full(n, 7, dtype=np.uint8)
    travel = np.full(n, np.nan, dtype=np.float32)
    pos = np.full((n, 3), np.nan, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    rc = getattr(lib, fn)(prog.handle, starts.ctypes.data_as(fp), dirs.ctypes.data_as(fp), n, 256, C.c_float(1e-4),
                          C.c_float(math.inf), flags, hit.ctypes.data_as(C.POINTER(C.c_uint8)),
                          travel.ctypes.data_as(fp), pos.ctypes.data_as(fp))
    assert rc == 0, (fn, rc, lib.tg_last_error())
    return hit, travel, pos


@pytest.mark.parametrize("fast", [False, True])
def test_static_raymarch_matches_allocating_one(gpu, fast):
    from tangerine_amd import _lib, scenes
    lib = _lib.load()
    prog = scenes.build_sdf(scenes.random_csg_spec(1, 256)).program()
    flags = _lib.TG_FLAG_FAST if fast else 0
    for seed, n in [(1, 50_000), (2, 3_000)]:
        starts, dirs = _rays(seed, n)
        ref = _march(lib, "tg_raymarch", prog, starts, dirs, flags)
        got = _march(lib, "tg_raymarch_static", prog, starts, dirs, flags)
        assert np.array_equal(got[0], ref[0]), (seed, n)
        assert np.array_equal(got[1].view(np.uint32), ref[1].view(np.uint32)), (seed, n)
        assert np.array_equal(got[2].view(np.uint32), ref[2].view(np.uint32)), (seed, n)
        assert 0.2 < ref[0].mean() < 1.0


def test_static_raymarch_refuses_too_many_rays(gpu):
    from tangerine_amd import _lib, sdf
    lib = _lib.load()
    prog = sdf.sphere(2.0).program()
    n = STATIC_MAX + 1
    starts = np.zeros((n, 3), dtype=np.float32)
    dirs = np.ones((n, 3), dtype=np.float32)
    hit = np.zeros(n, dtype=np.uint8)
    travel = np.zeros(n, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    rc = lib.tg_raymarch_static(prog.handle, starts.ctypes.data_as(fp), dirs.ctypes.data_as(fp), n, 10,
                                C.c_float(1e-3), C.c_float(math.inf), 0, hit.ctypes.data_as(C.POINTER(C.c_uint8)),
                                travel.ctypes.data_as(fp), None)
    assert rc == TG_ERR_LIMIT
    assert b"TG_RAYMARCH_STATIC_MAX" in lib.tg_last_error()
