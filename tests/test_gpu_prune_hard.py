"""Pruning and the brick certificate on trees chosen to break them.

The pruning decisions and the certificate use a rounding slack derived from
the coordinates and brush sizes (rounding_slack in tg_program.cpp); these
trees push rounding errors well past a fixed 2e-4: a uniform scale of 50, a
tree moved by ~1e3, translations spread over +-1e3, thin features with blend
thresholds of 1e-3.  Bars, per tree:
  * the pruned exact export is bit-identical to the unpruned exact export on
    every sample, and 20,000 seeded samples equal the oracle interpreter;
  * the pruned fast export is bit-identical to the unpruned fast export;
  * the certificate-culled mesh is byte-identical to the plain mesh and no
    brick holding an active cube is skipped.
And the zero term of a dropped blend: samples exactly on a kept surface
(-0.0 from the plane's products) under BlendInter / BlendDiff come out +0.0,
as the reference's double-precision zero term makes them.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _case(name):
    from tangerine_amd import scenes
    if name == "scale50":
        return scenes.random_csg_spec(7, 64) + [("scale", 50.0)], (-300.0, 300.0), 192
    if name == "moved1e3":
        spec = scenes.random_csg_spec(8, 64) + [("move", 1000.0, -700.0, 400.0)]
        return spec, None, 192
    if name == "spread1e3":
        return scenes.random_csg_spec(9, 128, translate=1000.0, size=(5.0, 30.0)), (-1040.0, 1040.0), 256
    if name == "thin":
        return scenes.random_csg_spec(10, 128, translate=0.5, size=(0.002, 0.02), blend=(1e-3, 2e-3)), \
            (-0.6, 0.6), 256
    raise KeyError(name)


def _grid(name, lohi, cells):
    from tangerine_amd import scenes
    if lohi is not None:
        return scenes.grid_for(lohi[0], lohi[1], cells)
    # a cube of edge 12 around (1000, -700, 400), on a binary step
    step = np.float32(12.0 / cells)
    origin = tuple(float(np.float32(c - 6.0)) for c in (1000.0, -700.0, 400.0))
    return origin, (float(step),) * 3, (cells,) * 3


@pytest.mark.parametrize("name", ["scale50", "moved1e3", "spread1e3", "thin"])
def test_adversarial_pruning_is_value_exact(gpu, oracle, name):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    spec, lohi, cells = _case(name)
    node = scenes.build_sdf(spec)
    prog, words = node.program(), node.compile()
    origin, step, cc = _grid(name, lohi, cells)
    m = Mesher()
    m.export(prog, origin, step, cc, prune=False)
    whole = bits(m.grid()).copy()
    vw, tw = m.mesh()
    st = m.export(prog, origin, step, cc)
    pruned = bits(m.grid())
    assert np.array_equal(pruned, whole), f"{np.count_nonzero(pruned != whole)} samples differ"
    active = m.active_bricks().astype(bool)
    assert active.sum() > 10
    n = cells + 1
    rng = np.random.default_rng(5)
    idx = rng.integers(0, n, size=(20000, 3))
    pts = (np.asarray(origin, np.float32)[None, :] + idx.astype(np.float32) * np.asarray(step, np.float32)[None, :])
    got = pruned[(idx[:, 2] * n + idx[:, 1]) * n + idx[:, 0]]
    assert np.array_equal(got, bits(oracle.interp_words(words, pts.astype(np.float32))))
    assert st["flops"] < prog.info["flops_per_eval"] * st["exact_evals"]
    m.export(prog, origin, step, cc, fast=True, prune=False)
    fw = bits(m.grid()).copy()
    m.export(prog, origin, step, cc, fast=True)
    assert np.array_equal(bits(m.grid()), fw)
    m.export(prog, origin, step, cc)
    m.export_culled(prog, None, origin, step, cc)
    kept = m.brick_flags().astype(bool)
    vc, tc = m.mesh()
    assert int(np.count_nonzero(active & ~kept)) == 0
    assert vc.tobytes() == vw.tobytes() and tc.tobytes() == tw.tobytes()
    m.close()


@pytest.mark.parametrize("which", ["inter_left", "diff_left", "diff_right"])
def test_dropped_blend_adds_positive_zero(gpu, oracle, which):
    from tangerine_amd import scenes, sdf
    from tangerine_amd.export import Mesher
    down = sdf.plane(0.0, 0.0, -1.0)  # -z: -0.0 at z = 0 where x < 0 and y < 0
    up = sdf.plane(0.0, 0.0, 1.0)     # +z: +0.0 at z = 0, negated to -0.0 by the difference
    if which == "inter_left":         # max(plane, big sphere interior): keeps the plane
        tree = sdf.blend_inter(down, sdf.sphere(20.0), 0.2)
    elif which == "diff_left":        # max(plane, -far sphere): keeps the plane
        tree = sdf.blend_diff(down, sdf.sphere(1.0).move(10.0, 10.0, 10.0), 0.2)
    else:                             # max(-(big sphere), -plane): keeps the negated plane
        tree = sdf.blend_diff(sdf.sphere(200.0), up, 0.2)
    prog, words = tree.program(), tree.compile()
    origin, step, cells = scenes.grid_for(-2.0, 2.0, 64)
    m = Mesher()
    m.export(prog, origin, step, cells)
    got = bits(m.grid())
    m.export(prog, origin, step, cells, prune=False)
    whole = bits(m.grid())
    m.close()
    n = 65
    # the z = 0 plane (k = 32) over x < 0, y < 0: the kept plane's products give -0.0
    k = 32
    ii, jj = np.meshgrid(np.arange(0, 32), np.arange(0, 32), indexing="ij")
    sel = ((k * n + jj) * n + ii).ravel()
    pts = np.stack([np.float32(-2.0) + ii.ravel().astype(np.float32) * np.float32(step[0]),
                    np.float32(-2.0) + jj.ravel().astype(np.float32) * np.float32(step[1]),
                    np.zeros(sel.size, np.float32)], axis=1)
    expect = bits(oracle.interp_words(words, pts))
    assert np.all(expect == 0), "the oracle's blend adds +0.0"
    assert np.array_equal(whole[sel], expect)
    assert np.array_equal(got[sel], expect)
    assert np.array_equal(got, whole)
