"""MagicaVoxel export on the device (VoxExport, tangerine/magica.cpp:27-70).

Bar: with exact numerics the kept cells (raster order) are identical to the
oracle's on config 1 and on random trees, and the written file equals the
reference's (its own vox::VoxWriter) byte for byte; fast numerics keep the
same cells within 0.5%.
"""

import numpy as np
import pytest

from conftest import golden_spec, load_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,prims,grid", [(0, 0, 24.0), (5, 16, 6.0), (2, 64, 10.0)])
def test_vox_cells_match_oracle(gpu, oracle, seed, prims, grid):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    spec = scenes.config1_spec() if prims == 0 else scenes.random_csg_spec(seed, prims)
    node = scenes.build_sdf(spec)
    prog, words = node.program(), node.compile()
    bounds = node.bounds()
    m = Mesher()
    cells = m.export_vox(prog, bounds, grid)
    size, ref = oracle.vox_cells(words, np.asarray(bounds, np.float32).reshape(-1), grid)
    assert len(ref) > 1000 and np.array_equal(cells, ref)
    fast = m.export_vox(prog, bounds, grid, fast=True)
    assert abs(len(fast) - len(ref)) <= 0.005 * len(ref)
    m.close()


def test_vox_file_matches_reference(gpu, tmp_path):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    g = load_golden("vox_config1_g16.npz")
    prog = scenes.build_sdf(golden_spec(g)).program()
    m = Mesher()
    cells = m.export_vox(prog, g["bounds"], float(g["grid_size"]), int(g["color_index"]), path=tmp_path / "a.vox")
    assert len(cells) == int(g["voxels"])
    assert (tmp_path / "a.vox").read_bytes() == g["vox_bytes"].tobytes()
    m.close()


def test_vox_edges(gpu, tmp_path):
    from tangerine_amd import sdf
    from tangerine_amd._lib import TangerineError
    from tangerine_amd.export import Mesher
    prog = sdf.sphere(2.0).program()
    m = Mesher()
    with pytest.raises(TangerineError, match="empty"):
        m.export_vox(prog, [[0, 0, 0], [0, 1, 1]], 4.0)
    # one voxel: the whole box is the cell, its centre at distance 1 from the surface
    cells = m.export_vox(prog, [[-1, -1, -1], [1, 1, 1]], 0.5, path=tmp_path / "one.vox")
    assert cells.size in (0, 1)
    m.close()
