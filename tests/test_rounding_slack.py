"""The rounding slack of value-exact pruning, measured against float64.

Pruning and the brick certificate drop an operand (skip a brick) only with a
margin of slack = 2 K u (pmax + leaf extent), u = 2^-24, K = 64, plus 1e-5 of
the operand magnitudes (tangerine_amd/csrc/tg_program.cpp, rounding_slack).
That is sound when every evaluation of the float32 interpreter errs by at
most (K / 2) u (|p| + leaf extent) + 5e-6 |value| from the program's own
function.  The oracle's float32 interpreter (bit-identical to the device's
exact build) is compared here with the float64 restatement (oracle/
interp64.py) on random trees and on adversarial ones: a uniform scale of 50,
a whole tree moved by ~1e3, translations spread over +-1e3, thin features
with blend thresholds of 1e-3.  The measured constant must stay below K / 2.
"""

import numpy as np
import pytest

from conftest import load_golden

U = 2.0 ** -24
K = 64.0


def _spec(name):
    from tangerine_amd import scenes
    if name == "rand256":
        return scenes.random_csg_spec(1, 256), 6.0, (0.0, 0.0, 0.0)
    if name == "rand2048":
        return scenes.random_csg_spec(1, 2048), 8.0, (0.0, 0.0, 0.0)
    if name == "scale50":
        return scenes.random_csg_spec(7, 64) + [("scale", 50.0)], 300.0, (0.0, 0.0, 0.0)
    if name == "moved1e3":
        return scenes.random_csg_spec(8, 64) + [("move", 1000.0, -700.0, 400.0)], 6.0, (1000.0, -700.0, 400.0)
    if name == "spread1e3":
        return scenes.random_csg_spec(9, 128, translate=1000.0, size=(5.0, 30.0)), 1000.0, (0.0, 0.0, 0.0)
    if name == "thin":
        return scenes.random_csg_spec(10, 128, translate=0.5, size=(0.002, 0.02), blend=(1e-3, 2e-3)), 0.6, \
            (0.0, 0.0, 0.0)
    raise KeyError(name)


def _points(spec, ext, centre, n, seed):
    """Uniform points over the scene cube plus points near every primitive."""
    rng = np.random.default_rng(seed)
    c = np.asarray(centre, dtype=np.float64)
    pts = [c + rng.uniform(-ext, ext, size=(n, 3))]
    moves = [np.array(t[1:4], dtype=np.float64) for t in spec if t[0] == "move"]
    if moves:
        anchors = np.stack(moves)[rng.integers(0, len(moves), n)]
        scale = next((t[1] for t in spec if t[0] == "scale"), 1.0)
        pts.append(anchors * scale + c * (0 if scale == 1.0 and not np.any(c) else 1) + rng.normal(
            scale=0.05 * ext, size=(n, 3)))
    return np.concatenate(pts).astype(np.float32)


def test_interp64_matches_reference_values():
    """On the golden fixtures the float64 interpreter agrees with the
    reference's float32 interpreter to within float32 rounding."""
    from oracle.interp64 import interp64
    for name in ("points_config1.npz", "points_rand16_s5.npz", "points_rand256_s1.npz"):
        g = load_golden(name)
        d64 = interp64(g["words"], g["points"])
        ref = g["interp_eval"].astype(np.float64)
        p = np.abs(g["points"].astype(np.float64)).max(axis=1)
        assert np.all(np.abs(d64 - ref) <= 64 * U * (p + 10.0) + 1e-6 * np.abs(ref)), name


def test_interp64_closed_forms():
    from oracle.interp64 import interp64
    from tangerine_amd import sdf
    pts = np.random.default_rng(3).uniform(-3, 3, size=(2000, 3)).astype(np.float32)
    p64 = pts.astype(np.float64)
    d = interp64(sdf.sphere(2.5).compile(), pts)
    assert np.allclose(d, np.linalg.norm(p64, axis=1) - np.float64(np.float32(1.25)), rtol=0, atol=1e-12)
    # blend of two identical spheres: min - T/4 exactly
    t = np.float32(0.4)
    d = interp64(sdf.blend_union(sdf.sphere(2.0), sdf.sphere(2.0), float(t)).compile(), pts)
    assert np.allclose(d, np.linalg.norm(p64, axis=1) - 1.0 - np.float64(t) / 4, rtol=0, atol=1e-12)
    # a planted mistake (max instead of min in the union) is caught
    d = interp64(sdf.union(sdf.sphere(1.0), sdf.sphere(2.0).move(0.5, 0, 0)).compile(), pts)
    wrong = np.maximum(np.linalg.norm(p64, axis=1) - 0.5,
                       np.linalg.norm(p64 - np.array([0.5, 0, 0]), axis=1) - 1.0)
    assert not np.allclose(d, wrong, atol=1e-6)


@pytest.mark.parametrize("name", ["rand256", "rand2048", "scale50", "moved1e3", "spread1e3", "thin"])
def test_float32_error_within_slack_constant(oracle, name):
    from oracle.interp64 import interp64
    from tangerine_amd import scenes
    from tangerine_amd.sdf import inspect_words
    spec, ext, centre = _spec(name)
    words = scenes.build_sdf(spec).compile()
    extent = inspect_words(words)["leaf_extent"]
    n = 3000 if name == "rand2048" else 20000
    pts = _points(spec, ext, centre, n, 11)
    d32 = oracle.interp_words(words, pts).astype(np.float64)
    d64 = interp64(words, pts)
    ok = np.isfinite(d64)
    assert ok.mean() > 0.99
    p = np.linalg.norm(pts.astype(np.float64), axis=1)
    excess = np.maximum(np.abs(d32 - d64) - 5e-6 * np.abs(d64), 0.0)[ok]
    ratio = excess / (U * (p[ok] + extent))
    worst = float(ratio.max())
    print(f"{name}: leaf extent {extent:.4g}, largest error constant {worst:.2f} (bound {K / 2:.0f}), "
          f"largest |error| {np.abs(d32 - d64)[ok].max():.3g}")
    assert worst <= K / 2
    if name in ("moved1e3", "spread1e3", "scale50"):
        # these are the cases a fixed 2e-4 slack would not cover
        assert 2 * K * U * (p.max() + extent) > 2e-4
