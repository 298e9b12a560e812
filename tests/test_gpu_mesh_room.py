"""Surface nets in the room an earlier mesh left.

After its first mesh, a mesher runs the surface-net stages of later exports
without waiting for their vertex and quad counts: the kernels write within
the room the last mesh left (a quarter more than it), and a mesh that
outgrows it is made again the waiting way.  Each export below must give
vertices and triangles byte-identical to a fresh mesher's export of the same
tree and grid: a larger mesh after a small one (outgrows the room), the same
mesh again (fits), a small mesh in a large room (fits, then the room
shrinks), and an empty mesh.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fresh(prog, grid):
    from tangerine_amd.export import Mesher
    m = Mesher()
    m.export(prog, *grid)
    v, t = m.mesh()
    m.close()
    return v.copy(), t.copy()


def test_mesh_room_sequence_matches_fresh_exports(gpu):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    small = scenes.build_sdf([("sphere", 1.0)]).program()
    big = scenes.build_sdf(scenes.random_csg_spec(3, 96)).program()
    far = scenes.build_sdf([("sphere", 0.5), ("move", 20.0, 0.0, 0.0)]).program()
    g_small = scenes.grid_for(-2.0, 2.0, 48)
    g_big = scenes.grid_for(-6.0, 6.0, 160)
    ref = {"small": _fresh(small, g_small), "big": _fresh(big, g_big), "far": _fresh(far, g_small)}
    assert ref["big"][0].shape[0] > 8 * ref["small"][0].shape[0]
    assert ref["far"][0].shape[0] == 0
    m = Mesher()
    for name, prog, grid in [("small", small, g_small), ("big", big, g_big), ("big", big, g_big),
                             ("small", small, g_small), ("small", small, g_small), ("far", far, g_small),
                             ("big", big, g_big)]:
        m.export(prog, *grid)
        v, t = m.mesh()
        rv, rt = ref[name]
        assert v.shape == rv.shape and t.shape == rt.shape, name
        assert np.array_equal(v.view(np.uint32), rv.view(np.uint32)), name
        assert np.array_equal(t, rt), name
    m.close()
