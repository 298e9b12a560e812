"""Batched export: many trees in one persistent launch.

Bar: every tree's samples and mesh are bit-identical to the single-tree
pruned export of the same grid (exact and fast numerics), graph replays and
every tuning leave them unchanged, and the work counters add up.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


# (seed, primitives, cells) with uneven sizes and resolutions, non-cubic grid last
TREES = [(11, 8, 48), (12, 40, 64), (13, 200, 96), (14, 700, 80), (15, 2048, 128), (16, 16, 33), (17, 120, 57)]


def _case():
    from tangerine_amd import scenes
    progs, grids = [], []
    for seed, prims, cells in TREES:
        progs.append(scenes.build_sdf(scenes.random_csg_spec(seed, prims)).program())
        grids.append(scenes.grid_for(-6.0, 6.0, cells))
    # one non-cubic grid
    o, s, _ = scenes.grid_for(-6.0, 6.0, 64)
    grids[-1] = (o, s, (57, 40, 71))
    return progs, grids


@pytest.mark.parametrize("fast", [False, True])
def test_batch_matches_single_exports(gpu, fast):
    from tangerine_amd.batch import Batch
    from tangerine_amd.export import Mesher
    progs, grids = _case()
    single = Mesher()
    expect = []
    for p, g in zip(progs, grids):
        single.export(p, *g, fast=fast)
        expect.append((bits(single.grid()), *single.mesh()))
    single.close()
    b = Batch()
    st = b.export(progs, grids, fast=fast)
    assert st["trees"] == len(progs) and st["blocks"] > 0
    assert st["samples"] == sum(int(np.prod([c + 1 for c in g[2]])) for g in grids)
    for i, (g, v, t) in enumerate(expect):
        m = b.mesher(i)
        assert np.array_equal(bits(m.grid()), g), i
        v1, t1 = m.mesh()
        assert v1.tobytes() == v.tobytes() and t1.tobytes() == t.tobytes(), i
    # replays of the captured launch rewrite the same samples
    for _ in range(2):
        assert b.replay() > 0
    for i, (g, _, _) in enumerate(expect):
        assert np.array_equal(bits(b.mesher(i).grid()), g), i
    b.close()


def test_batch_exports_in_the_room_of_the_last(gpu):
    """A batch's later exports make each tree's mesh in the room its last
    mesh left: the same trees again (every mesh fits), then the trees in
    reverse order (meshes that outgrow their tree slot's room are made again
    the waiting way); every mesh equals its single-tree export byte for byte."""
    from tangerine_amd.batch import Batch
    from tangerine_amd.export import Mesher
    progs, grids = _case()
    single = Mesher()
    expect = []
    for p, g in zip(progs, grids):
        single.export(p, *g)
        expect.append(single.mesh())
    single.close()
    b = Batch()
    for order in [list(range(len(progs))), list(range(len(progs))), list(reversed(range(len(progs)))),
                  list(range(len(progs)))]:
        b.export([progs[i] for i in order], [grids[i] for i in order])
        for slot, i in enumerate(order):
            v, t = expect[i]
            v1, t1 = b.mesher(slot).mesh()
            assert v1.tobytes() == v.tobytes() and t1.tobytes() == t.tobytes(), (order, slot)
    b.close()


def test_batch_tunings_agree(gpu):
    from tangerine_amd.batch import Batch
    progs, grids = _case()
    b = Batch()
    b.export(progs, grids)
    ref = [bits(b.mesher(i).grid()) for i in range(len(progs))]
    r = b.autotune()
    assert r["candidates"] and len(r["candidates"]) == 18 and r["eval_ms"] > 0
    for pts, waves, wg in [(2, 2, 1), (8, 4, 4), (4, 2, 2)]:
        b.set_tuning(pts, waves, wg)
        b.replay()
        for i in range(len(progs)):
            assert np.array_equal(bits(b.mesher(i).grid()), ref[i]), (pts, waves, wg, i)
    b.close()


def test_batch_rejects_bad_input(gpu):
    from tangerine_amd import scenes, sdf
    from tangerine_amd._lib import TangerineError
    from tangerine_amd.batch import Batch
    b = Batch()
    ell = sdf.ellipsoid(1, 2, 3).program()
    with pytest.raises(TangerineError, match="Lipschitz"):
        b.export([ell], [scenes.grid_for(-2, 2, 16)])
    with pytest.raises(TangerineError, match="no batch"):
        b.replay()
    b.close()


def test_batch_owned_meshers_refuse_relayout(gpu):
    """A tree's mesher belongs to the batch: calls that would lay its buffers
    out again (the captured graph points at them) and destroy are refused;
    reads still work and the batch replays unchanged."""
    import ctypes as C
    from tangerine_amd import _lib, scenes
    from tangerine_amd.batch import Batch
    lib = _lib.load()
    progs = [scenes.build_sdf(scenes.random_csg_spec(s, 16)).program() for s in (21, 22)]
    grids = [scenes.grid_for(-6.0, 6.0, 40), scenes.grid_for(-6.0, 6.0, 24)]
    b = Batch()
    b.export(progs, grids)
    before = bits(b.mesher(0).grid()).copy()
    h = lib.tg_batch_mesher(b.h, 0)
    g = _lib.grid_struct(*grids[1])
    assert lib.tg_mesher_export(h, progs[1].handle, C.byref(g), 0) == -1
    assert b"belongs to a tg_batch" in lib.tg_last_error()
    assert lib.tg_mesher_set_tuning(h, 2, 4, 0, 0) == -1
    lib.tg_mesher_destroy(h)  # refused: the handle stays valid
    assert b"belongs to a tg_batch" in lib.tg_last_error()
    assert b.replay() > 0
    assert np.array_equal(bits(b.mesher(0).grid()), before)
    b.close()
