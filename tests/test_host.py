"""Host-side checks that need no GPU: the C ABI loads and exports every
declared symbol, the host tree and flattener reproduce the reference's
program words bit for bit, and malformed programs are rejected before any
device work."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, golden_spec, load_golden

HEADER = ROOT / "include" / "tangerine_gpu.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from tangerine_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    # the ctypes table covers the whole ABI
    assert set(names) <= set(_lib.SIGNATURES)


@pytest.mark.parametrize("name", ["points_config1.npz", "points_rand16_s5.npz", "points_rand256_s1.npz",
                                  "points_rand2048_s1.npz"])
def test_flattener_words_match_reference(name):
    from tangerine_amd import scenes
    g = load_golden(name)
    node = scenes.build_sdf(golden_spec(g))
    assert np.array_equal(node.compile(), g["words"])
    assert node.stack_size() == int(g["stack_size"])
    assert node.leaf_count() == int(g["leaf_count"])
    assert np.array_equal(node.bounds(), g["bounds"])


def test_lua_surface_config1_words():
    """The config-1 model written with the Lua-named builders (diameters,
    threshold last) flattens to the reference's 17 words."""
    from tangerine_amd import sdf
    model = sdf.blend_union(sdf.sphere(2), sdf.box(1.6, 1, 1).move_x(0.6), 0.2).diff(sdf.cylinder(0.6, 3))
    g = load_golden("points_config1.npz")
    assert np.array_equal(model.compile(), g["words"])
    assert model.compile().size == 17


def test_transforms_copy():
    from tangerine_amd import sdf
    a = sdf.sphere(1)
    b = a.move(1, 2, 3)
    assert a.compile().size == 3 and b.compile().size == 7  # Offset(4) + Sphere(2) + Stop
    c = b.rotate_x(30).scale(2)
    w = c.compile()
    assert w[0] == 17 and w[17] == 1 and w[19] == 18  # Matrix, Sphere, ScaleField


def test_rejects_malformed_programs():
    from tangerine_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()

    def words(*w):
        a = np.array(w, dtype=np.uint32)
        return a.ctypes.data_as(_lib._u32p), a.size, a

    one = np.float32(1.0).view(np.uint32)
    cases = [
        (9, 0),                       # Union on an empty stack
        (1, one),                     # no Stop
        (1, one, 1, one, 0),          # two values left at Stop
        (99, 0),                      # unknown opcode
        (1, one, 0, 1, one),          # Stop not last
        (17, 0, 0),                   # truncated matrix
    ]
    for c in cases:
        ptr, n, keep = words(*c)
        rc = lib.tg_program_from_words(ptr, n, C.byref(h))
        assert rc == -3, (c, rc, _lib.last_error())
        assert _lib.last_error()


def test_deep_stacks_are_reordered():
    """A right-leaning chain of differences needs one reference stack slot
    per level (StackSize 21 for 20 levels, 301 for 300); the decoder runs each
    operator's deeper operand first, so the device needs 2 slots at any
    length.  A balanced tree needs log2(brushes) + 1 slots in any order:
    32,768 brushes fit the 16 device slots as they are, 65,536 (17) are
    refused with a size error."""
    from tangerine_amd import _lib, sdf
    from oracle.interp64 import SPHERE, STOP, UNION
    for levels, ref in ((20, 21), (300, 301)):
        node = sdf.sphere(1)
        for i in range(levels):
            node = sdf.sphere(1).diff(node)  # right-deep
        assert node.stack_size() == ref
        info = sdf.inspect_words(node.compile())
        assert info["max_stack"] == 2 and info["brushes"] == levels + 1

    def balanced(k):
        one = np.array([SPHERE, np.float32(0.5).view(np.uint32)], dtype=np.uint32)
        w = one
        for _ in range(k):
            w = np.concatenate([w, w, np.array([UNION], dtype=np.uint32)])
        return np.concatenate([w, np.array([STOP], dtype=np.uint32)])

    assert sdf.inspect_words(balanced(15))["max_stack"] == 16
    w = balanced(16)
    info = _lib.TgProgramInfo()
    rc = _lib.load().tg_words_inspect(w.ctypes.data_as(_lib._u32p), w.size, C.byref(info))
    assert rc == -5 and "stack" in _lib.last_error()


def test_builder_argument_errors():
    from tangerine_amd import _lib
    lib = _lib.load()
    assert not lib.tg_plane(0.0, 0.0, 0.0)
    assert "plane" in _lib.last_error()
    assert not lib.tg_union(None, None)
    assert not lib.tg_sphere(float("nan"))


def test_random_specs_are_deterministic():
    from tangerine_amd import scenes
    a = scenes.random_csg_spec(42, 300)
    b = scenes.random_csg_spec(42, 300)
    assert a == b and a != scenes.random_csg_spec(43, 300)
    kinds = {t[0] for t in a}
    assert {"sphere", "box", "torus", "cylinder", "cone", "rotate", "move"} <= kinds
    assert kinds & {"union", "blend_union"} and kinds & {"diff", "blend_diff"} and kinds & {"inter", "blend_inter"}
    batch = scenes.batch_specs(1, 64)
    counts = [sum(1 for t in s if t[0] in scenes.PRIMITIVES) for s, _ in batch]
    assert min(counts) >= 8 and max(counts) <= 2048
    assert {r for _, r in batch} <= {64, 128, 192, 256, 384}


def test_specs_build_identically_in_all_backends(oracle):
    from tangerine_amd import scenes
    spec = scenes.random_csg_spec(9, 40)
    gpu_words = scenes.build_sdf(spec).compile()
    h = oracle.build(spec)
    assert np.array_equal(gpu_words, oracle.compile(h))
    oracle.free(h)


def test_lipschitz_gate_for_pruning():
    """Per-cell pruning needs 1-Lipschitz fields; the decoder clears the flag
    for ellipsoids, non-unit rotations and unnormalised plane normals, and keeps
    it through rigid transforms and uniform scales."""
    from tangerine_amd import sdf, scenes
    assert scenes.build_sdf(scenes.random_csg_spec(3, 64)).inspect()["lipschitz"]
    assert sdf.sphere(1).rotate_x(30).scale(0.25).move(1, 2, 3).inspect()["lipschitz"]
    assert sdf.box(1, 2, 3).scale(3.0).inspect()["lipschitz"]
    assert not sdf.ellipsoid(1, 2, 3).union(sdf.sphere(1)).inspect()["lipschitz"]
    # a short quaternion makes ToMatrix() shrink, so its inverse stretches space
    assert not sdf.sphere(1).rotate(0.3, 0.0, 0.0, 0.3).inspect()["lipschitz"]
    assert sdf.sphere(1).rotate(1, 1, 1, 1).inspect()["lipschitz"]  # inverse contracts: still a bound
    words = sdf.plane(0, 0, 1).compile()
    words[3] = np.float32(2.0).view(np.uint32)  # normal (0, 0, 2)
    assert not sdf.inspect_words(words)["lipschitz"]
    info = scenes.build_sdf(scenes.random_csg_spec(1, 256)).inspect()
    assert info["brushes"] == 256 and info["set_ops"] == 255 and info["max_stack"] == 9
