"""Parity at the full benchmark size: the seeded 2048-primitive tree on the
1024^3-cell grid over [-8,8]^3 (1025^3 = 1,076,890,625 samples), the grid the
headline export is timed on.

Bars:
  * 20,000 seeded random samples of the device grid are bit-identical to the
    oracle interpreter at the same coordinates (lo + i * step in float32);
  * the culled export with the exact certificate alone (no network) gives
    vertices and triangles byte-identical to the plain export, and keeps
    every brick that holds an active cube (size-independent properties of
    the whole 1024^3 mesh).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_full_size_samples_and_culled_mesh(gpu, oracle):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    node = scenes.build_sdf(scenes.random_csg_spec(1, 2048))
    prog, words = node.program(), node.compile()
    origin, step, cells = scenes.grid_for(-8.0, 8.0, 1024)
    m = Mesher()
    plain = m.export(prog, origin, step, cells)
    v0, t0 = m.mesh()
    active = m.active_bricks().astype(bool)
    assert v0.shape[0] > 1_000_000 and t0.shape[0] > 2_000_000
    n = cells[0] + 1
    grid = m.grid()
    assert grid.size == n ** 3
    idx = np.random.default_rng(7).integers(0, n, size=(20000, 3))
    pts = (np.float32(origin[0]) + idx.astype(np.float32) * np.float32(step[0])).astype(np.float32)
    got = grid[(idx[:, 2] * n + idx[:, 1]) * n + idx[:, 0]]
    del grid
    ref = oracle.interp_words(words, pts)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    culled = m.export_culled(prog, None, origin, step, cells)
    kept = m.brick_flags().astype(bool)
    v1, t1 = m.mesh()
    assert int(np.count_nonzero(active & ~kept)) == 0
    assert v0.tobytes() == v1.tobytes() and t0.tobytes() == t1.tobytes()
    assert culled["tiles_evaluated"] < 0.5 * plain["tiles_evaluated"]
    m.close()
