"""Point-cloud export on the device (PointCloudExportThread,
tangerine/export.cpp:384-480).

Bar: with exact numerics the points are bit-identical to the oracle's
(raster order; the threaded reference appends them in thread order) for 0,
1 and 3 refinement steps, on the config-1 tree and on random trees with
rotations; the fast build keeps the same cells (selection within 1e-5 of the
half-step diagonal may flip, so counts agree within 0.5%) and refined points
lie within 4e-3 of the half-step diagonal of the exact ones for 99% of points
(the finite-difference gradient amplifies ulp differences ~1e4 times).  The PLY file is
the reference's point-cloud layout (no face element) with the oracle's
tetrahedral gradients as normals.
"""

import numpy as np
import pytest

from conftest import golden_spec, load_golden

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


CASES = [
    ("config1", lambda s: s.config1_spec(), (-2.0, -1.5, -1.5), (2.0, 1.5, 1.5), 0.05),
    ("rand16", lambda s: s.random_csg_spec(5, 16), (-6.0,) * 3, (6.0,) * 3, 0.15),
    ("rand256", lambda s: s.random_csg_spec(1, 256), (-6.0, -6.0, -6.0), (6.0, 6.0, 6.0), 0.2),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("iterations", [0, 1, 3])
def test_points_match_oracle(gpu, oracle, case, iterations):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    name, spec, lo, hi, step = case
    node = scenes.build_sdf(spec(scenes))
    prog = node.program()
    words = node.compile()
    m = Mesher()
    st = m.export_points(prog, lo, hi, step, iterations)
    v, t = m.mesh()
    ref = oracle.point_cloud(words, lo, hi, step, iterations)
    assert len(ref) > 500 and t.shape == (0, 3)
    assert v.shape == ref.shape and np.array_equal(bits(v), bits(ref)), name
    assert st["samples"] == int(np.prod(np.ceil((np.float32(hi) - np.float32(lo)) / np.float32(step))))
    m.close()


def test_points_match_reference_fixture(gpu):
    """The reference's own points for config 1 (translations only, where its
    tree-walk gradient equals the interpreter's), bit for bit."""
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    g = load_golden("cloud_config1.npz")
    prog = scenes.build_sdf(golden_spec(g)).program()
    m = Mesher()
    for it in g["iterations"]:
        m.export_points(prog, g["lo"], g["hi"], float(g["step"]), int(it))
        v, _ = m.mesh()
        assert np.array_equal(bits(v), bits(g[f"points_it{int(it)}"]))
    m.close()


def test_fast_points_close(gpu, oracle):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    node = scenes.build_sdf(scenes.random_csg_spec(5, 16))
    prog = node.program()
    m = Mesher()
    m.export_points(prog, (-6.0,) * 3, (6.0,) * 3, 0.15, 0)
    exact0, _ = m.mesh()
    m.export_points(prog, (-6.0,) * 3, (6.0,) * 3, 0.15, 0, fast=True)
    fast0, _ = m.mesh()
    assert abs(len(fast0) - len(exact0)) <= 0.005 * len(exact0)
    m.export_points(prog, (-6.0,) * 3, (6.0,) * 3, 0.15, 2)
    exact2, _ = m.mesh()
    m.export_points(prog, (-6.0,) * 3, (6.0,) * 3, 0.15, 2, fast=True)
    fast2, _ = m.mesh()
    if len(fast0) == len(exact0) and np.array_equal(fast0, exact0):
        # the h = 1e-4 tetrahedral gradient amplifies the fast build's ulp-level
        # distance differences ~1e4 times (as between the reference's tree
        # walk and its interpreter): the bar is that of the reference fixture
        diag = float(np.linalg.norm(np.full(3, 0.075, dtype=np.float32)))
        d = np.linalg.norm(fast2 - exact2, axis=1)
        assert np.quantile(d, 0.99) <= 4e-3 * diag and d.max() <= diag
    m.close()


def test_point_cloud_ply(gpu, oracle, tmp_path):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    node = scenes.build_sdf(scenes.config1_spec())
    prog = node.program()
    words = node.compile()
    m = Mesher()
    m.export_points(prog, (-2.0, -1.5, -1.5), (2.0, 1.5, 1.5), 0.1, 2)
    v, _ = m.mesh()
    m.write_ply(prog, tmp_path / "c.ply", 2.0)
    head = ("ply\nformat binary_little_endian 1.0\ncomment Created by Tangerine\n"
            f"element vertex {len(v)}\nproperty float x\nproperty float y\nproperty float z\n"
            "property float nx\nproperty float ny\nproperty float nz\nend_header\n").encode()
    rec = np.zeros(len(v), dtype=[("p", "<f4", 3), ("n", "<f4", 3)])
    rec["p"] = (v * np.float32(2.0)).astype(np.float32)
    rec["n"] = oracle.gradient_tetra(words, v, 1e-4)
    assert (tmp_path / "c.ply").read_bytes() == head + rec.tobytes()
    m.close()


def test_point_cloud_edges(gpu):
    from tangerine_amd import sdf
    from tangerine_amd._lib import TangerineError
    from tangerine_amd.export import Mesher
    prog = sdf.sphere(2.0).program()  # diameter 2 (the Lua API's argument)
    m = Mesher()
    # no cells (hi <= lo on an axis), and cells that miss the surface
    m.export_points(prog, (0.0, 0.0, 0.0), (0.0, 1.0, 1.0), 0.1, 2)
    assert m.counts() == (0, 0)
    m.export_points(prog, (5.0, 5.0, 5.0), (6.0, 6.0, 6.0), 0.1, 2)
    assert m.counts() == (0, 0)
    # one cell holding the surface
    m.export_points(prog, (0.875, -0.0625, -0.0625), (1.0, 0.0625, 0.0625), 0.125, 1)
    v, _ = m.mesh()
    assert v.shape == (1, 3) and abs(float(np.linalg.norm(v[0])) - 1.0) < 1e-5
    with pytest.raises(TangerineError, match="negative"):
        m.export_points(prog, (-1.0,) * 3, (1.0,) * 3, 0.1, -1)
    with pytest.raises(TangerineError, match="step"):
        m.export_points(prog, (-1.0,) * 3, (1.0,) * 3, 0.0, 0)
    with pytest.raises(TangerineError, match="culled|grid"):
        m.grid()
    m.close()
