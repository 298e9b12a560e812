"""GPU parity for point and grid evaluation and the device surface nets.

Bars (written into each test):
  * exact numerics (default flags): bit-identical to the reference
    interpreter on the same program words;
  * fast numerics (TG_FLAG_FAST): |d_gpu - d_ref| <= 1e-5 * scene extent;
  * both against the reference tree walk SDFNode::Eval: |d| within
    1e-5 * extent, identical sign wherever |d| exceeds that bound;
  * surface nets on identical samples: identical vertex bits, identical
    vertex numbering and identical (canonical) triangle sets.
"""

import ctypes as C

import numpy as np
import pytest

from conftest import canonical_faces, golden_spec, grid_points, load_golden, sha

pytestmark = pytest.mark.gpu

TREES = [("points_config1.npz", 5.0), ("points_rand16_s5.npz", 12.0), ("points_rand256_s1.npz", 12.0),
         ("points_rand2048_s1.npz", 16.0)]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def program(words):
    from tangerine_amd.sdf import Program
    return Program.from_words(words)


@pytest.mark.parametrize("name,extent", TREES)
def test_points_exact_bit_identical(gpu, name, extent):
    g = load_golden(name)
    prog = program(g["words"])
    d = prog.eval_points(g["points"])
    assert np.array_equal(bits(d), bits(g["interp_eval"]))


@pytest.mark.parametrize("name,extent", TREES)
def test_points_fast_within_tolerance(gpu, name, extent):
    g = load_golden(name)
    tol = 1e-5 * extent
    d = program(g["words"]).eval_points(g["points"], fast=True)
    assert np.max(np.abs(d - g["interp_eval"])) <= tol


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("name,extent", TREES)
def test_points_vs_tree_walk(gpu, name, extent, fast):
    g = load_golden(name)
    tol = 1e-5 * extent
    d = program(g["words"]).eval_points(g["points"], fast=fast)
    ref = g["tree_eval"]
    assert np.max(np.abs(d - ref)) <= tol
    far = np.abs(ref) > tol
    assert np.array_equal(d[far] >= 0, ref[far] >= 0)


def test_every_opcode_bit_identical(gpu, oracle):
    """All 19 reference opcodes, plus deep (16-slot) stacks, on random points."""
    from tangerine_amd import scenes
    spec = [
        ("sphere", 0.7), ("ellipsoid", 0.9, 0.4, 0.6), ("move", 0.3, 0.0, 0.0), ("union",),
        ("box", 0.5, 0.6, 0.7), ("rotate_axis", 0, 30.0), ("blend_inter", 0.15),
        ("torus", 0.8, 0.2), ("scale", 1.3), ("blend_union", 0.2),
        ("cylinder", 0.3, 1.2), ("rotate", 0.2, 0.3, 0.1, 0.927), ("diff",),
        ("cone", 0.6, 1.4), ("move", 0.0, 0.5, -0.5), ("blend_diff", 0.1),
        ("coninder", 0.3, 0.6, 1.0), ("flate", 0.05), ("inter",),
        ("plane", 0.2, 0.9, 0.1), ("move", 0.0, -0.2, 0.0), ("union",),
    ]
    # right-deep chain of differences: needs 16 stack slots
    chain = [("sphere", 1.5)]
    for i in range(15):
        chain = [("sphere", 0.2 + 0.05 * i), ("move", 0.1 * i, 0.0, 0.0)] + chain + [("diff",)]
    rng = np.random.default_rng(4)
    pts = rng.uniform(-2.5, 2.5, size=(50000, 3)).astype(np.float32)
    for s in (spec, chain):
        h = oracle.build(s)
        words = oracle.compile(h)
        expect = oracle.interp_words(words, pts)
        node = scenes.build_sdf(s)
        assert np.array_equal(node.compile(), words)
        got = program(words).eval_points(pts)
        assert np.array_equal(bits(got), bits(expect))
        oracle.free(h)
    assert max(program(oracle.compile(oracle.build(chain))).info["max_stack"], 0) == 16


def test_point_edge_cases(gpu):
    from tangerine_amd import _lib, sdf
    prog = sdf.sphere(2).program()
    out = np.zeros(1, dtype=np.float32)
    assert prog.eval_points(np.zeros((0, 3), dtype=np.float32)).size == 0
    assert prog.eval_points(np.array([[0, 0, 0]], dtype=np.float32))[0] == np.float32(-1.0)
    # counts that are not a multiple of the block size
    pts = np.random.default_rng(0).uniform(-2, 2, size=(257, 3)).astype(np.float32)
    d = prog.eval_points(pts)
    np.testing.assert_allclose(d, np.linalg.norm(pts, axis=1) - 1.0, atol=1e-6)


def test_sqrt_special_arguments_bit_identical(gpu, oracle):
    """Square-root arguments at the edges of the range: points at a sphere's
    centre (argument 0), at tiny offsets (arguments 1e-40 .. 1e-30, where
    the correctly rounded sequence scales its input) and far away (1e36,
    near overflow), alone and mixed into waves of ordinary points, give the
    oracle's bits (glibc's correctly rounded sqrtf)."""
    from tangerine_amd import sdf
    node = sdf.sphere(2)
    prog, words = node.program(), node.compile()
    rng = np.random.default_rng(5)
    special = np.float32([[0, 0, 0], [1e-20, 0, 0], [0, 3e-19, 0], [1e-15, 1e-15, 1e-15], [2e-24, 0, 7e-25],
                          [1e18, 0, 0], [1e19, 1e19, 0], [-0.0, 0, 0], [5e-20, -5e-20, 0]])
    ordinary = rng.uniform(-3, 3, size=(4096, 3)).astype(np.float32)
    mixed = ordinary.copy()
    mixed[::97] = special[np.arange(len(mixed[::97])) % len(special)]
    for pts in (special, ordinary, mixed, np.repeat(special, 64, axis=0)):
        d = prog.eval_points(pts)
        ref = oracle.interp_words(words, pts)
        assert np.array_equal(d.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("name", ["points_config1.npz", "points_rand16_s5.npz", "points_rand256_s1.npz",
                                  "points_rand2048_s1.npz"])
def test_non_finite_points_match_oracle(gpu, oracle, name):
    """Infinite and NaN coordinates (where rays that march off to infinity
    end up): the device gives the oracle's value wherever that is a number and
    NaN wherever it is NaN, so no such point passes a d <= eps test that the
    reference's would fail."""
    from tangerine_amd.sdf import Program
    g = load_golden(name)
    prog = Program.from_words(g["words"])
    inf, nan = np.inf, np.nan
    pts = np.float32([[inf, 0, 0], [-inf, 1, 2], [0, inf, 0], [0, 0, -inf], [inf, inf, inf], [inf, -inf, 0],
                      [nan, 0, 0], [0, nan, 1], [1, 2, nan], [nan, nan, nan], [inf, nan, 0], [1e30, 1e30, 1e30],
                      [3e38, 0, 0], [-3e38, 3e38, 1]])
    pts = np.concatenate([pts, np.random.default_rng(1).uniform(-6, 6, size=(50, 3)).astype(np.float32)])
    d = prog.eval_points(pts)
    ref = oracle.interp_words(g["words"], pts)
    assert np.array_equal(np.isnan(d), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.array_equal(d[ok].view(np.uint32), ref[ok].view(np.uint32))


@pytest.mark.parametrize("scramble", [0, 1])
def test_sqrt_all_inputs_match_sqrtf(gpu, scramble):
    """All 2^32 float inputs: the exact build's square root (rsq and one
    corrected step; sqrtf itself for a wave holding an input outside
    [2^-96, FLT_MAX]) gives the bits of the compiler's correctly rounded
    sqrtf, with waves of consecutive bit patterns (every input of the main
    path meets it) and with every wave mixing tiny, zero, negative, infinite,
    NaN and ordinary inputs."""
    from tangerine_amd import _lib
    lib = _lib.load()
    n = C.c_uint64(1)
    _lib.check(lib.tg_check_sqrt(scramble, C.byref(n)), "tg_check_sqrt")
    assert n.value == 0


def _cone_divisors(spec, count):
    """dot(Q, Q) and Q.x of the first `count` cones of a spec, in float32 as
    SDF::Cone (tangerine/sdf_evaluator.cpp:1296-1305) and SDFMath::Cone
    (:229-236) form them: Q = height * (radius / height, -1)."""
    f32 = np.float32
    out = []
    for tok in spec:
        if tok[0] != "cone" or len(out) >= 2 * count:
            continue
        r, h = f32(tok[1]), f32(tok[2])
        qx, qy = h * (r / h), h * f32(-1.0)
        out += [qx * qx + qy * qy, qx]
    return out


@pytest.mark.gpu
def test_quotient_by_record_constant(gpu):
    """Division by a record constant (the cone's dot(Q, Q) and Q.x, the
    coninder's dot(K2, K2)) as RN(1/b) once and one multiply plus two fused
    multiply-adds per point: the correctly rounded a / b for every float a on
    that path (all 2^32 bit patterns per divisor), over the 2048-primitive
    tree's cone divisors, log-uniform divisors over [2^-32, 2^32] of both
    signs, and divisors at the ends of the range and of a binade."""
    from tangerine_amd import _lib, scenes
    f32 = np.float32
    rng = np.random.default_rng(11)
    divs = _cone_divisors(scenes.random_csg_spec(1, 2048), 48)
    divs += list(f32(np.exp2(rng.uniform(-32, 32, 64))) * rng.choice([f32(-1), f32(1)], 64))
    specials = [1.0, 2.0 ** -32, 2.0 ** 32, 1.5, 3.0, 0.1, 7.0, -1.0]
    divs += [f32(v) for v in specials]
    divs += [np.nextafter(f32(1), f32(2)), np.nextafter(f32(2), f32(1)), np.nextafter(f32(2.0 ** 32), f32(0))]
    arr = np.asarray(divs, dtype=np.float32)
    lib = _lib.load()
    bad, taken = C.c_uint64(1), C.c_uint64(0)
    _lib.check(lib.tg_check_quotient(arr.ctypes.data_as(_lib._fp), len(arr), C.byref(bad), C.byref(taken)),
               "tg_check_quotient")
    # 2^31 values per divisor: every |a| in [2^-64, 2^64)
    assert taken.value == len(arr) * 2 * 128 * 2 ** 23
    assert bad.value == 0


# ---- grid export ---------------------------------------------------------------

class Mesher:
    def __init__(self):
        from tangerine_amd import _lib
        self.lib = _lib.load()
        self._lib = _lib
        h = C.c_void_p()
        _lib.check(self.lib.tg_mesher_create(C.byref(h)), "tg_mesher_create")
        self.h = h.value

    def __del__(self):
        self.lib.tg_mesher_destroy(self.h)

    def export(self, prog, origin, step, cells, fast=False):
        g = self._lib.grid_struct(origin, step, cells)
        self._lib.check(self.lib.tg_mesher_export(self.h, prog.handle, C.byref(g), 1 if fast else 0), "export")

    def mesh_values(self, values, origin, step, cells):
        g = self._lib.grid_struct(origin, step, cells)
        v = np.ascontiguousarray(values, dtype=np.float32)
        self._lib.check(self.lib.tg_mesher_mesh_values(self.h, v.ctypes.data_as(self._lib._fp), C.byref(g)),
                        "mesh_values")

    def mesh(self):
        nv, nt = C.c_uint64(), C.c_uint64()
        self._lib.check(self.lib.tg_mesher_counts(self.h, C.byref(nv), C.byref(nt)), "counts")
        v = np.zeros((nv.value, 3), dtype=np.float32)
        t = np.zeros((nt.value, 3), dtype=np.uint32)
        self._lib.check(self.lib.tg_mesher_copy_mesh(self.h, v.ctypes.data_as(self._lib._fp),
                                                     t.ctypes.data_as(self._lib._u32p)), "copy mesh")
        return v, t

    def grid(self, cells):
        nx, ny, nz = (int(c) + 1 for c in cells)
        out = np.zeros(nx * ny * nz, dtype=np.float32)
        self._lib.check(self.lib.tg_mesher_copy_grid(self.h, out.ctypes.data_as(self._lib._fp)), "copy grid")
        return out


@pytest.fixture(scope="module")
def mesher(gpu):
    return Mesher()


def test_config1_export_bit_identical(mesher):
    """129^3 samples of config 1: grid, vertices and faces match the
    reference interpreter + serial surface nets hashes."""
    g = load_golden("grid_config1.npz")
    words = load_golden("points_config1.npz")["words"]
    prog = program(words)
    mesher.export(prog, g["origin"], g["step"], g["cells"])
    vals = mesher.grid(g["cells"])
    assert sha(vals) == str(g["interp_sha"])
    v, t = mesher.mesh()
    assert len(v) == int(g["mesh_vertices"]) and len(t) == int(g["mesh_faces"])
    assert sha(v) == str(g["mesh_vertex_sha"])
    assert sha(canonical_faces(t)) == str(g["mesh_face_sha"])


def test_config1_fast_export_within_tolerance(mesher, oracle):
    g = load_golden("grid_config1.npz")
    words = load_golden("points_config1.npz")["words"]
    mesher.export(program(words), g["origin"], g["step"], g["cells"], fast=True)
    vals = mesher.grid(g["cells"])
    exact, _ = oracle.grid_interp(words, g["origin"], g["step"], g["cells"], threads=8)
    tol = 1e-5 * 4.0
    assert np.max(np.abs(vals - exact.ravel())) <= tol
    # topology on the fast samples equals the reference mesher on those samples
    v, t = mesher.mesh()
    rv, rf = oracle.surface_nets(vals, g["origin"], g["step"], g["cells"])
    assert np.array_equal(bits(v), bits(rv))
    assert np.array_equal(canonical_faces(t), canonical_faces(rf))


@pytest.mark.parametrize("name", ["mesh_config1_c24.npz", "mesh_rand16_c40.npz"])
def test_surface_nets_on_identical_values(mesher, name):
    g = load_golden(name)
    mesher.mesh_values(g["values"], g["origin"], g["step"], g["cells"])
    v, t = mesher.mesh()
    assert np.array_equal(bits(v), bits(g["vertices"]))
    assert np.array_equal(canonical_faces(t), g["faces"])


@pytest.mark.parametrize("seed,prims,cells", [(1, 256, 96), (3, 64, 57)])
def test_random_tree_export_matches_oracle(mesher, oracle, seed, prims, cells):
    from tangerine_amd import scenes
    spec = scenes.random_csg_spec(seed, prims)
    origin, step, cc = scenes.grid_for(-6.0, 6.0, cells)
    h = oracle.build(spec)
    words = oracle.compile(h)
    mesher.export(program(words), origin, step, cc)
    vals = mesher.grid(cc)
    expect, _ = oracle.grid_interp(words, origin, step, cc, threads=8)
    assert np.array_equal(bits(vals), bits(expect.ravel()))
    v, t = mesher.mesh()
    rv, rf = oracle.surface_nets(vals, origin, step, cc)
    assert len(v) > 100
    assert np.array_equal(bits(v), bits(rv))
    assert np.array_equal(canonical_faces(t), canonical_faces(rf))
    oracle.free(h)


def test_non_cubic_and_tiny_grids(mesher, oracle):
    from tangerine_amd import sdf
    model = sdf.blend_union(sdf.sphere(2), sdf.box(1.6, 1, 1).move_x(0.6), 0.2).diff(sdf.cylinder(0.6, 3))
    words = model.compile()
    prog = program(words)
    for origin, step, cells in [((-2.0, -1.5, -1.25), (0.0625, 0.125, 0.1), (70, 25, 27)),
                                ((-0.5, -0.5, -0.5), (1.0, 1.0, 1.0), (1, 1, 1)),
                                ((-1.3, -1.1, -1.2), (0.05, 0.05, 0.05), (130, 9, 65))]:
        mesher.export(prog, origin, step, cells)
        vals = mesher.grid(cells)
        expect, _ = oracle.grid_interp(words, origin, step, cells, threads=8)
        assert np.array_equal(bits(vals), bits(expect.ravel()))
        v, t = mesher.mesh()
        rv, rf = oracle.surface_nets(vals, origin, step, cells)
        assert np.array_equal(bits(v), bits(rv))
        assert np.array_equal(canonical_faces(t), canonical_faces(rf))


def test_empty_scene_exports_empty_mesh(mesher):
    from tangerine_amd import sdf
    prog = sdf.sphere(0.5).move(10, 10, 10).program()
    mesher.export(prog, (-1, -1, -1), (0.125,) * 3, (16, 16, 16))
    v, t = mesher.mesh()
    assert len(v) == 0 and len(t) == 0


def test_export_rejects_bad_grids(mesher):
    from tangerine_amd import _lib, sdf
    prog = sdf.sphere(1).program()
    for origin, step, cells in [((0, 0, 0), (0.1,) * 3, (0, 4, 4)), ((0, 0, 0), (-0.1,) * 3, (4, 4, 4)),
                                ((0, 0, 0), (0.1,) * 3, (20000, 4, 4))]:
        g = _lib.grid_struct(origin, step, cells)
        assert mesher.lib.tg_mesher_export(mesher.h, prog.handle, C.byref(g), 0) < 0
        assert _lib.last_error()
