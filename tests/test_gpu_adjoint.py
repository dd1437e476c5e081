"""GPU differentiable rollout (mpm_diff_* through the C-ABI) against the FP64
oracle rollout and its reverse pass (oracle/mpm_adjoint.c, itself pinned by
central differences in tests/test_adjoint.py).

Tolerances: the GPU forward is float32 with an FP64 stress.  Positions near
x = 128 carry a float32 spacing of 7.6e-6, so 200 position updates alone may
round by up to 200 * 3.8e-6 = 7.6e-4: the final positions must agree with
the FP64 forward to 1e-3 absolute.  Gradients must agree to a relative L2
error of 1e-3 per gradient block (dL/dx0, dL/dv0, dL/dlambda, dL/dmu).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def block(pkg):
    return pkg.scenes.elastic_block()


@pytest.fixture(scope="module")
def fp64(orc, block):
    state, lam, mu, config, target = block
    cfg = orc.diff_config(config)
    final = orc.diff_forward(state, lam, mu, cfg)
    _, bar = orc.com_loss(final, target)
    b0, gl, gm, conf = orc.diff_backward(state, lam, mu, cfg, bar)
    return final, bar, b0, gl, gm, conf


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run_gpu(pkg, block, bar=None, segment=20, cfg_over=None, graphs=True):
    state, lam, mu, config, target = block
    config = dict(config, **(cfg_over or {}))
    r = pkg.adjoint.ElasticRollout(len(lam), config, segment_steps=segment, device=0, graphs=graphs)
    try:
        final = r.forward(state, lam, mu)
        if bar is None:
            _, bar = pkg_com_bar(final, target)
        b0, gl, gm = r.backward(bar)
        return final, b0, gl, gm, r.status(), r.timing(), r.layout()
    finally:
        r.close()


def pkg_com_bar(final, target):
    d = final[:, :2].astype(np.float64).mean(axis=0) - target
    bar = np.zeros(final.shape, np.float32)
    bar[:, :2] = 2.0 * d / final.shape[0]
    return float(d @ d), bar


def test_forward_matches_fp64_rollout(pkg, block, fp64):
    final, *_ = run_gpu(pkg, block)
    ref = fp64[0]
    print("max |dx|", np.abs(final[:, :2] - ref[:, :2]).max(), "max |dv|", np.abs(final[:, 2:4] - ref[:, 2:4]).max())
    np.testing.assert_allclose(final[:, :2], ref[:, :2], rtol=0, atol=1e-3)
    np.testing.assert_allclose(final[:, 2:4], ref[:, 2:4], rtol=0, atol=2e-2)
    np.testing.assert_allclose(final[:, 8:12], ref[:, 8:12], rtol=0, atol=2e-3)


def test_gradients_match_fp64_adjoint(pkg, block, fp64):
    _, bar, b0, gl, gm, conf = fp64
    final, g0, ggl, ggm, status, timing, layout = run_gpu(pkg, block, bar=bar.astype(np.float32))
    errs = {"x0": rel_l2(g0[:, 0:2], b0[:, 0:2]), "v0": rel_l2(g0[:, 2:4], b0[:, 2:4]),
            "lambda": rel_l2(ggl, gl), "mu": rel_l2(ggm, gm)}
    print("relative L2 errors", errs, "timing", timing, "layout", layout, "confined", conf)
    assert status & ~pkg.native.MPM_DIFF_CONFINED == 0
    for k, e in errs.items():
        assert e <= 1e-3, (k, e)
    assert layout["segment_steps"] == 20 and layout["checkpoints"] == 10


def test_segment_length_and_graphs_do_not_change_gradients(pkg, block, fp64):
    bar = fp64[1].astype(np.float32)
    base = run_gpu(pkg, block, bar=bar, segment=20)
    for seg, graphs in ((7, True), (200, True), (20, False)):
        other = run_gpu(pkg, block, bar=bar, segment=seg, graphs=graphs)
        assert other[6]["segment_steps"] == seg
        for a, b in zip(base[1:4], other[1:4]):
            # float atomics make sums order-dependent: agreement to rounding
            assert rel_l2(b, a) < 1e-4


def test_free_flight_closed_form(pkg):
    state, lam, mu, config, target = pkg.scenes.elastic_block(side=32, origin=(100.0, 120.0), gravity=(0.0, -9.81))
    final, g0, gl, gm, status, *_ = run_gpu(pkg, (state, lam, mu, config, target))
    assert final[:, :2].min() > 10.0
    _, bar = pkg_com_bar(final, target)
    g = bar[0, :2].astype(np.float64)
    T, dt = config["steps"], config["dt"]
    np.testing.assert_allclose(g0[:, 0:2], np.broadcast_to(g, g0[:, 0:2].shape), rtol=1e-3, atol=1e-3 * np.abs(g).max())
    np.testing.assert_allclose(g0[:, 2:4], np.broadcast_to(T * dt * g, g0[:, 2:4].shape), rtol=1e-3,
                               atol=1e-3 * T * dt * np.abs(g).max())
    # no wall, no dependence on the stiffness (relative to the v0 gradient scale)
    scale = T * dt * np.abs(g).max()
    assert np.abs(gl).max() < 1e-3 * scale and np.abs(gm).max() < 1e-3 * scale


def test_autograd_function_matches_direct_calls(pkg, block, fp64):
    torch = pytest.importorskip("torch")
    state, lam, mu, config, target = block
    dev = torch.device("cuda:0")
    s = torch.tensor(state, dtype=torch.float32, device=dev, requires_grad=True)
    la = torch.tensor(lam, dtype=torch.float32, device=dev, requires_grad=True)
    m = torch.tensor(mu, dtype=torch.float32, device=dev, requires_grad=True)
    r = pkg.adjoint.ElasticRollout(len(lam), config, device=0)
    try:
        final = pkg.adjoint.rollout(r, s, la, m)
        loss = pkg.adjoint.com_loss(final, target)
        loss.backward()
    finally:
        r.close()
    _, bar, b0, gl, gm, _ = fp64
    assert rel_l2(s.grad[:, 2:4].cpu().numpy(), b0[:, 2:4]) <= 1e-3
    assert rel_l2(la.grad.cpu().numpy(), gl) <= 1e-3
    assert rel_l2(m.grad.cpu().numpy(), gm) <= 1e-3


def test_particle_outside_grid_is_reported(pkg, block):
    state, lam, mu, config, target = block
    bad = state.copy()
    bad[0, 0] = 1e6
    r = pkg.adjoint.ElasticRollout(len(lam), config, device=0)
    try:
        with pytest.raises(pkg.native.NativeError) as e:
            r.forward(bad, lam, mu)
        assert e.value.code == pkg.native.MPM_ESTATE
        assert r.status() & pkg.native.MPM_DIFF_STENCIL
        with pytest.raises(pkg.native.NativeError):
            r.backward(np.zeros_like(state, dtype=np.float32))
    finally:
        r.close()


def test_gpu_gradients_match_central_differences_on_the_full_block(pkg, orc):
    """The GPU adjoint of the full differentiable block (its own float32
    forward and loss) against central differences of the FP64 forward at the
    sampled components of tests/test_adjoint_c5.py: 1e-3 relative, 1e-8
    absolute at the differences' noise floor."""
    from test_adjoint_c5 import c5_fd_samples, pick
    (state, lam, mu, config, target, _), samples = c5_fd_samples(pkg, orc)
    final, g0, gl, gm, status, *_ = run_gpu(pkg, (state, lam, mu, config, target))
    assert status & ~pkg.native.MPM_DIFF_CONFINED == 0
    for label, fd, index in samples:
        g = pick((g0.astype(np.float64), gl.astype(np.float64), gm.astype(np.float64)), index)
        print(f"{label}: FD {fd:+.6e} GPU {g:+.6e}", flush=True)
        assert g == pytest.approx(fd, rel=1e-3, abs=1e-8), label

