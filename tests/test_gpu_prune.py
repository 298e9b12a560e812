"""Per-brick pruning keeps every sample's value bit for bit.

The pruned export (level 1: 32^3-sample blocks, level 2: 8^3 tiles) must give
the same samples as the whole-program interpreter: bit-identical to the
oracle in exact numerics, and bit-identical to the unpruned device path in
fast numerics; and it must execute less work than the whole program."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def mesher(gpu):
    from tangerine_amd.export import Mesher
    m = Mesher()
    yield m
    m.close()


CASES = [
    # (seed, primitives, lo, hi, cells)
    (5, 16, -6.0, 6.0, 64),
    (2, 64, -6.0, 6.0, 100),
    (1, 256, -6.0, 6.0, 128),
    (3, 2048, -8.0, 8.0, 96),
]


@pytest.mark.parametrize("seed,prims,lo,hi,cells", CASES)
def test_pruned_exact_equals_oracle(mesher, oracle, seed, prims, lo, hi, cells):
    from tangerine_amd import scenes
    from tangerine_amd.sdf import Program
    spec = scenes.random_csg_spec(seed, prims)
    h = oracle.build(spec)
    words = oracle.compile(h)
    origin, step, cc = scenes.grid_for(lo, hi, cells)
    prog = Program.from_words(words)
    st = mesher.export(prog, origin, step, cc)
    vals = mesher.grid()
    expect, _ = oracle.grid_interp(words, origin, step, cc, threads=8)
    assert np.array_equal(bits(vals), bits(expect.ravel()))
    v, t = mesher.mesh()
    full = prog.info["flops_per_eval"] * st["exact_evals"]
    assert st["exact_evals"] >= (cells + 1) ** 3
    if prims >= 64:
        assert st["flops"] < 0.5 * full, (st["flops"], full)
    oracle.free(h)


@pytest.mark.parametrize("seed,prims,lo,hi,cells", CASES[1:3])
def test_pruned_fast_equals_unpruned_fast(mesher, seed, prims, lo, hi, cells):
    from tangerine_amd import scenes
    spec = scenes.random_csg_spec(seed, prims)
    prog = scenes.build_sdf(spec).program()
    origin, step, cc = scenes.grid_for(lo, hi, cells)
    mesher.export(prog, origin, step, cc, fast=True, prune=True)
    a = mesher.grid()
    va, ta = mesher.mesh()
    mesher.export(prog, origin, step, cc, fast=True, prune=False)
    b = mesher.grid()
    vb, tb = mesher.mesh()
    assert np.array_equal(bits(a), bits(b))
    assert np.array_equal(bits(va), bits(vb)) and np.array_equal(ta, tb)


def test_pruning_gated_off_for_non_lipschitz_fields(mesher, oracle):
    """An ellipsoid is not a distance bound: the export still matches exactly
    (pruning is disabled for such programs)."""
    from tangerine_amd import sdf
    model = sdf.union(sdf.ellipsoid(2.0, 1.0, 1.4).rotate_z(20), sdf.box(1, 1, 1).move(0.8, 0, 0),
                      sdf.sphere(0.8).move(-1, 0.3, 0))
    words = model.compile()
    assert not model.inspect()["lipschitz"]
    origin, step, cells = (-2.0, -2.0, -2.0), (0.0625,) * 3, (64, 64, 64)
    from tangerine_amd.sdf import Program
    st = mesher.export(Program.from_words(words), origin, step, cells)
    expect, _ = oracle.grid_interp(words, origin, step, cells, threads=8)
    assert np.array_equal(bits(mesher.grid()), bits(expect.ravel()))


def test_deep_stack_and_partial_blocks(mesher, oracle):
    """A 16-slot program on a grid whose tile counts are not multiples of the
    4-tile block edge."""
    chain = [("sphere", 1.5)]
    for i in range(15):
        chain = [("sphere", 0.2 + 0.05 * i), ("move", 0.15 * i - 1.0, 0.1 * (i % 3), 0.0)] + chain + [("diff",)]
    h = oracle.build(chain)
    words = oracle.compile(h)
    from tangerine_amd.sdf import Program
    prog = Program.from_words(words)
    assert prog.info["max_stack"] == 16
    origin, step, cells = (-2.5, -2.0, -2.2), (0.05, 0.0625, 0.04), (97, 41, 110)
    mesher.export(prog, origin, step, cells)
    expect, _ = oracle.grid_interp(words, origin, step, cells, threads=8)
    assert np.array_equal(bits(mesher.grid()), bits(expect.ravel()))
    oracle.free(h)


@pytest.mark.parametrize("length", [40, 90])
def test_long_chains(mesher, oracle, length):
    """A left-deep chain of `length` blended spheres: its entry heights (40
    and 90) fall on both sides of the 64 heights the level-parallel decision
    pass takes (beyond that the pruning kernel walks the program itself);
    both give the oracle's grid bit for bit."""
    rng = np.random.default_rng(length)
    chain = [("sphere", 0.6)]
    for i in range(length - 1):
        x, y, z = (float(v) for v in rng.uniform(-2.0, 2.0, 3))
        chain += [("sphere", float(rng.uniform(0.3, 0.9))), ("move", x, y, z),
                  ("blend_union", 0.2) if i % 3 else ("union",)]
    h = oracle.build(chain)
    words = oracle.compile(h)
    from tangerine_amd.sdf import Program
    prog = Program.from_words(words)
    origin, step, cells = (-3.0, -3.0, -3.0), (0.05,) * 3, (120, 120, 120)
    st = mesher.export(prog, origin, step, cells)
    expect, _ = oracle.grid_interp(words, origin, step, cells, threads=8)
    assert np.array_equal(bits(mesher.grid()), bits(expect.ravel()))
    assert st["exact_evals"] > 0
    oracle.free(h)


def test_every_tuning_gives_the_same_grid(mesher):
    """Every (points per lane, waves, workgroups per CU, shortest pruned block
    program) setting of the block evaluation writes the same bits (a
    workgroup per block with LDS tables, or persistent workgroups with global
    tables), and the autotuner's choice is one of the timed candidates."""
    from tangerine_amd import scenes
    prog = scenes.build_sdf(scenes.random_csg_spec(1, 256)).program()
    origin, step, cells = scenes.grid_for(-6.0, 6.0, 100)
    ref = None
    launches = ([(p, w, 0, 0) for p in (2, 4, 8) for w in (1, 2, 4)] +
                [(p, w, 0, g) for p in (2, 4, 8) for w, g in ((2, 8), (4, 4), (1, 16))] +
                [(4, 2, 16, 0), (2, 4, 1000, 0), (4, 2, 16, 8)])
    for pts, waves, min_len, wgs in launches:
        mesher.set_tuning(pts, waves, min_len, wgs)
        mesher.export(prog, origin, step, cells)
        g = bits(mesher.grid())
        if ref is None:
            ref = g
        assert np.array_equal(g, ref), (pts, waves, min_len, wgs)
    r = mesher.autotune(prog, origin, step, cells)
    assert len(r["candidates"]) == 24 and r["eval_ms"] > 0
    assert (r["points_per_lane"], r["waves"], r["workgroups_per_cu"]) in {
        (c["points_per_lane"], c["waves"], c["workgroups_per_cu"]) for c in r["candidates"]}
    mesher.export(prog, origin, step, cells)
    assert np.array_equal(bits(mesher.grid()), ref)
    mesher.set_tuning(4, 2, 0, 8)
