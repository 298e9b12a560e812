"""Learned brick culling: the surrogate's kernels against the CPU oracle, and
culled export checked against the plain export.

Bars:
  * forward: device predictions against the oracle's BF16-rounding variant,
    max |dy| <= 5e-3 and 99th percentile <= 5e-4 (hardware sin/exp and FP32
    accumulation order move a BF16 rounding now and then);
  * gradients of the mean squared error: per parameter block,
    ||g_dev - g_oracle|| <= 2e-2 ||g_oracle|| and loss within 1e-4 relative;
  * culled export (surrogate proposal + exact brick certificate, or the
    certificate alone): vertices and triangles byte-identical to the plain
    exact export of the same grid, no brick holding an active cube left out,
    and fewer tiles evaluated than the plain export; the surrogate alone
    (uncertified) is audited against the same plain export.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LO = (-6.0, -6.0, -6.0)
SIZE = 12.0


def _points(n, seed):
    rng = np.random.default_rng(seed)
    return (np.array(LO) + SIZE * rng.random((n, 3))).astype(np.float32)


def test_predict_matches_oracle(gpu):
    from oracle import mlp
    from tangerine_amd.surrogate import Surrogate
    s = Surrogate(LO, SIZE, sigma=2.0, clamp=0.5, seed=11)
    params, fourier = s.params()
    assert np.all(np.isfinite(params)) and np.abs(params[mlp.W1:mlp.B1]).max() > 0
    x = _points(4000, 1)   # not a multiple of 32: exercises the ragged last group
    y = s.predict(x)
    ref, _ = mlp.forward(params, fourier, LO, SIZE, x, rounded=True)
    err = np.abs(y - ref)
    assert err.max() <= 5e-3
    assert np.quantile(err, 0.99) <= 5e-4
    assert np.std(ref) > 1e-2     # a transposed operand would not hide behind a constant output
    s.close()


@pytest.mark.parametrize("n", [32, 256])
def test_gradients_match_oracle(gpu, n):
    from oracle import mlp
    from tangerine_amd.surrogate import Surrogate
    s = Surrogate(LO, SIZE, sigma=2.0, clamp=0.5, seed=5)
    params, fourier = s.params()
    x = _points(n, 2)
    target = np.random.default_rng(3).normal(0, 0.6, n).astype(np.float32)
    loss, g = s.gradients(x, target)
    ref_loss, ref_g = mlp.loss_and_grad(params, fourier, LO, SIZE, x, target, 0.5, rounded=True)
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    blocks = [(mlp.W1, mlp.B1), (mlp.B1, mlp.W2), (mlp.W2, mlp.B2), (mlp.B2, mlp.W3), (mlp.W3, mlp.B3),
              (mlp.B3, mlp.PARAMS)]
    for a, b in blocks:
        d = np.linalg.norm(g[a:b] - ref_g[a:b])
        assert d <= 2e-2 * np.linalg.norm(ref_g[a:b]) + 1e-7, (a, b, d)
    s.close()


def test_gradient_step_reduces_loss(gpu):
    """Set parameters from the host, take one plain gradient step, and check
    that the device loss falls as the first-order model predicts."""
    from tangerine_amd.surrogate import Surrogate
    s = Surrogate(LO, SIZE, seed=9)
    x = _points(1024, 4)
    target = np.clip(np.linalg.norm(x, axis=1) - 3.0, -0.5, 0.5).astype(np.float32)
    l0, g = s.gradients(x, target)
    p, _ = s.params()
    lr = 1e-2 / max(float(np.linalg.norm(g)), 1e-12)
    s.set_params(p - lr * g)
    l1, _ = s.gradients(x, target)
    assert l1 < l0
    assert l0 - l1 == pytest.approx(lr * float(g @ g), rel=0.2)
    s.close()


def _sphere_program():
    from tangerine_amd import sdf
    return sdf.sphere(6.0).diff(sdf.box(3, 3, 8)).program()


def test_culled_export_rejects_bad_inputs(gpu):
    from tangerine_amd import sdf
    from tangerine_amd._lib import TangerineError as TgError
    from tangerine_amd.export import Mesher
    from tangerine_amd.surrogate import Surrogate
    m = Mesher()
    s = Surrogate(LO, SIZE)
    prog = _sphere_program()
    with pytest.raises(TgError, match="not calibrated"):
        m.export_culled(prog, s, LO, (0.1,) * 3, (32, 32, 32))
    s.set_margin(0.05)
    ell = sdf.ellipsoid(2, 3, 4).program()
    with pytest.raises(TgError, match="Lipschitz"):
        m.export_culled(ell, s, LO, (0.1,) * 3, (32, 32, 32))
    m.export_culled(prog, s, LO, (0.1,) * 3, (32, 32, 32))
    with pytest.raises(TgError, match="culled"):
        m.grid()
    m.close()
    s.close()


def test_audit_catches_a_surrogate_that_misses(gpu):
    """Planted mistake: a surrogate that predicts 0.4 everywhere with margin 0
    skips the interior bricks; the audit against the plain export reports
    the missed bricks and the meshes differ."""
    from tangerine_amd.export import Mesher
    from tangerine_amd.surrogate import Surrogate
    prog = _sphere_program()
    m = Mesher()
    # 256 cells: brick half-diagonal 4 sqrt(3) 12/256 = 0.325 < 0.4 < clamp
    step = (SIZE / 256,) * 3
    m.export(prog, LO, step, (256, 256, 256))
    v_plain, t_plain = m.mesh()
    active = m.active_bricks()
    s = Surrogate(LO, SIZE, clamp=0.5)
    p, _ = s.params()
    p[:] = 0.0
    p[-1] = 0.4
    s.set_params(p)
    s.set_margin(0.0)
    st = m.export_culled(prog, s, LO, step, (256, 256, 256), certify=False)
    kept = m.brick_flags()
    missed = int(np.count_nonzero(active.astype(bool) & ~kept.astype(bool)))
    assert missed > 0
    assert st["bricks_flagged"] == int(kept.sum()) and st["bricks_total"] == kept.size
    v, t = m.mesh()
    assert v.shape != v_plain.shape or t.shape != t_plain.shape
    # the exact certificate puts every missed brick back
    st = m.export_culled(prog, s, LO, step, (256, 256, 256), certify=True)
    kept = m.brick_flags().astype(bool)
    assert int(np.count_nonzero(active.astype(bool) & ~kept)) == 0
    assert st["bricks_certified"] >= missed
    v, t = m.mesh()
    assert v.tobytes() == v_plain.tobytes() and t.tobytes() == t_plain.tobytes()
    m.close()
    s.close()


@pytest.mark.parametrize("fast", [False, True])
def test_certificate_only_export_is_identical(gpu, fast):
    """No network: every brick is kept or skipped by its exact centre value."""
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    prog = scenes.build_sdf(scenes.random_csg_spec(2, 64)).program()
    m = Mesher()
    origin, step, cells = scenes.grid_for(-6.0, 6.0, 200)
    plain = m.export(prog, origin, step, cells, fast=fast)
    v0, t0 = m.mesh()
    active = m.active_bricks().astype(bool)
    culled = m.export_culled(prog, None, origin, step, cells, fast=fast)
    kept = m.brick_flags().astype(bool)
    v1, t1 = m.mesh()
    assert active.sum() > 100
    assert int(np.count_nonzero(active & ~kept)) == 0
    assert v0.tobytes() == v1.tobytes() and t0.tobytes() == t1.tobytes()
    assert culled["tiles_evaluated"] < 0.5 * plain["tiles_evaluated"]
    assert culled["bricks_certified"] == culled["bricks_flagged"] == kept.sum()
    # the culled launch shape (its own tuning) changes no byte of the mesh
    for shape in [(8, 2, 8), (2, 1, 0), (4, 4, 6), (8, 4, 2)]:
        m.set_culled_tuning(*shape)
        m.export_culled(prog, None, origin, step, cells, fast=fast)
        v2, t2 = m.mesh()
        assert v0.tobytes() == v2.tobytes() and t0.tobytes() == t2.tobytes(), shape
    m.close()


def test_fit_then_culled_export_is_identical(gpu):
    """Fit on the 256-primitive tree over [-6,6]^3, then a 512^3 culled export
    (brick half-diagonal 0.16) against the plain exact export of the same
    grid: the surrogate must keep well under half of the bricks."""
    from tangerine_amd import scenes
    from tangerine_amd.export import Mesher
    from tangerine_amd.surrogate import Surrogate
    prog = scenes.build_sdf(scenes.random_csg_spec(1, 256)).program()
    s = Surrogate((-6.25,) * 3, 12.5, sigma=2.0, clamp=0.5, seed=1)
    fit = s.fit(prog, samples=1 << 21, calib_samples=200_000, surface_sigma=0.05, epochs=16)
    log = s.loss_log()
    assert log.size == fit["steps"] and np.all(np.isfinite(log))
    assert log[-16:].mean() < 0.1 * log[:16].mean()
    assert 0 < fit["mean_residual"] < 0.25 and fit["mean_residual"] < fit["margin"]
    assert fit["surface_vertices"] > 1000
    m = Mesher()
    origin, step, cells = scenes.grid_for(-6.0, 6.0, 512)
    plain = m.export(prog, origin, step, cells)
    v0, t0 = m.mesh()
    active = m.active_bricks().astype(bool)
    guess = m.export_culled(prog, s, origin, step, cells, certify=False)
    guess_missed = int(np.count_nonzero(active & ~m.brick_flags().astype(bool)))
    culled = m.export_culled(prog, s, origin, step, cells)
    kept = m.brick_flags().astype(bool)
    v1, t1 = m.mesh()
    assert culled["bricks_certified"] >= guess_missed
    # certified, the surrogate's bar is its 90th percentile, not its largest
    # near-surface prediction: it proposes no more bricks than uncertified
    assert culled["cull_threshold"] <= guess["cull_threshold"]
    assert culled["bricks_flagged"] - culled["bricks_certified"] <= guess["bricks_flagged"]
    assert int(np.count_nonzero(active & ~kept)) == 0
    assert v0.shape == v1.shape and t0.shape == t1.shape
    assert v0.tobytes() == v1.tobytes() and t0.tobytes() == t1.tobytes()
    assert culled["cull_threshold"] < 0.5
    assert guess["bricks_flagged"] < 0.5 * guess["bricks_total"]
    assert culled["tiles_evaluated"] < 0.6 * plain["tiles_evaluated"]
    print("kept bricks", culled["bricks_flagged"], "of", culled["bricks_total"], "certified", culled["bricks_certified"],
          "surrogate alone missed", guess_missed, "threshold", culled["cull_threshold"],
          "tiles", culled["tiles_evaluated"], "of", plain["tiles_evaluated"])
    m.close()
    s.close()
