"""Block-program emission: the wave-per-block emission (a lane per entry)
against the sequential walk over 16 blocks' entries.

Both emit each block's pruned program from the same decision and liveness
passes; the programs must be the same entry for entry.  Seen from the
export: the same grid samples bit for bit, the same mesh bytes, and the
same program statistics (longest block program, its binary entries, the
stack high-water mark through the launch's LDS size) on trees that exercise
every emission case: leaves and unary operators, kept and dropped operands
of every set operator, differences that keep only their subtrahend
(negation), dropped blend operands (the +0.0 flag), swapped operand order,
deep stacks, and long programs on coarse grids.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


CASES = [
    ("random 256, fine", lambda s: s.random_csg_spec(1, 256), (-6.0, 6.0, 160)),
    ("random 2048, coarse", lambda s: s.random_csg_spec(2, 2048), (-6.0, 6.0, 64)),
    ("random 700, medium", lambda s: s.random_csg_spec(7, 700, blend=(0.001, 0.5)), (-6.0, 6.0, 128)),
]


@pytest.mark.parametrize("name,spec,grid", CASES, ids=[c[0] for c in CASES])
def test_emission_modes_agree(gpu, name, spec, grid):
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    prog = scenes.build_sdf(spec(scenes)).program()
    g = scenes.grid_for(*grid)
    out = {}
    for mode in (1, 2, 0):
        m = Mesher()
        m.set_emission(mode)
        st = m.export(prog, *g)
        v, t = m.mesh()
        out[mode] = (_bits(m.grid()), v.copy(), t.copy(),
                     (st["block_max_len"], st["block_max_binary"], st["block_lds_bytes"], st["flops"]))
        m.close()
    for mode in (2, 0):
        assert np.array_equal(out[mode][0], out[1][0]), (name, mode)
        assert out[mode][1].tobytes() == out[1][1].tobytes() and out[mode][2].tobytes() == out[1][2].tobytes(), (name, mode)
        assert out[mode][3] == out[1][3], (name, mode, out[mode][3], out[1][3])


def test_emission_modes_agree_on_operator_cases(gpu):
    """Hand-built trees: every set operator with operands far apart (one
    dropped), a difference whose minuend is far away (negation), blend
    intersections and differences dropping an operand (+0.0 flag), and a
    long right-leaning chain of differences (deep stack)."""
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    specs = []
    for op in ("union", "inter", "diff", "blend_union", "blend_inter", "blend_diff"):
        tok = (op, 0.2) if op.startswith("blend") else (op,)
        specs.append([("sphere", 1.0), ("move", -3.0, 0.0, 0.0), ("box", 0.5, 0.7, 0.9), ("move", 3.0, 0.5, 0.0),
                      tok, ("torus", 1.0, 0.3), ("move", 0.0, -3.0, 1.0), tok])
    # right-leaning: 40 operands pushed before the first operator (the
    # compiler reorders operands to fit the stack, so operand order swaps)
    chain = []
    for k in range(40):
        chain += [("sphere", 0.3 + 0.01 * k), ("move", -4.0 + 0.2 * k, 0.1 * (k % 7), 0.0)]
    chain += [("diff",) if k % 3 else ("blend_union", 0.1) for k in range(39)]
    specs.append(chain)
    g = scenes.grid_for(-5.0, 5.0, 96)
    for spec in specs:
        prog = scenes.build_sdf(spec).program()
        res = []
        for mode in (1, 2):
            m = Mesher()
            m.set_emission(mode)
            st = m.export(prog, *g)
            v, t = m.mesh()
            res.append((_bits(m.grid()), v.tobytes(), t.tobytes(), st["block_max_len"], st["block_lds_bytes"]))
            m.close()
        assert np.array_equal(res[0][0], res[1][0]), spec[:8]
        assert res[0][1:] == res[1][1:], spec[:8]
