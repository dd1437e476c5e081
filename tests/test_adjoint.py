"""The FP64 differentiable rollout of the oracle (oracle/mpm_adjoint.c),
pinned independently of its own code:

* one forward step equals the oracle's reference step (orc_step_f64) for a
  snow particle at plastic volume 1 whose singular values stay inside the
  snow yield band — the fixed-corotated stress is the snow stress at
  hardening factor 1 (reference/solver.ts:2171-2203);
* the reverse pass agrees with central finite differences (h = 1e-4) of the
  FP64 forward for a random linear functional of the whole final state:
  sampled components of the initial state, directional derivatives for
  lambda and mu, on a block hitting the floor (plus the centre-of-mass loss
  there); the substep path (solver.ts:2561-2592) and wall friction get their
  own cases;
* free flight: P2G/G2P conserve momentum exactly, so the centre of mass moves
  as a point mass and d com(T)/d v0_p = T dt / N, d com(T)/d x0_p = 1 / N,
  d com(T)/d lambda = d com(T)/d mu = 0.
"""
from __future__ import annotations

import numpy as np
import pytest


def small_block(pkg, side=8, steps=60, cells=48, origin=(18.0, 7.0), **kw):
    return pkg.scenes.elastic_block(side=side, steps=steps, cells=cells, origin=origin, **kw)


def loss_of(orc, cfg, state, lam, mu, weights):
    """A linear functional of the whole final state: its reverse seed is
    `weights` itself, so the check covers every output component."""
    return float((orc.diff_forward(state, lam, mu, cfg) * weights).sum())


def test_diff_step_matches_reference_snow_step(pkg, orc):
    state, lam, mu, config, _ = small_block(pkg, side=6)
    rng = np.random.default_rng(5)
    n = len(lam)
    # gentle motion and a pre-deformed F inside the snow yield band
    state[:, 2:4] = 0.3 + 0.01 * rng.standard_normal((n, 2))
    theta = 0.3
    R = np.array([[np.cos(theta), -np.sin(theta)], [np.sin(theta), np.cos(theta)]])
    F = R @ np.diag([1.004, 0.985])
    state[:, 8:12] = F.reshape(-1)
    state[:, 4:8] = 0.02 * rng.standard_normal((n, 4))
    lam = lam * (1 + 0.1 * rng.random(n))
    mu = mu * (1 + 0.1 * rng.random(n))
    cfg = orc.diff_config(config, steps=1)
    out = orc.diff_step(state, lam, mu, cfg)

    p = np.zeros(n, orc.PARTICLE_DTYPE)
    for k, name in enumerate(("x", "y", "vx", "vy", "c00", "c01", "c10", "c11", "f00", "f01", "f10", "f11")):
        p[name] = state[:, k]
    p["plastic_volume"] = 1.0
    p["lambda_"], p["mu"] = lam, mu
    p["density"], p["volume"] = 1.0, config["particleVolume"]
    p["model"] = orc.MODEL_SNOW
    rc = orc.make_config(config["dt"], config["cellsX"], config["cellsY"], config["cellSize"],
                         config["particleMass"], config["particleVolume"], config["gravity"], 0.0)
    ref, _ = orc.step_f64(p, rc)
    # the snow projection must not have acted (plastic volume stays 1)
    assert np.allclose(ref["plastic_volume"], 1.0, rtol=0, atol=1e-12)
    ref_state = np.stack([ref[k] for k in ("x", "y", "vx", "vy", "c00", "c01", "c10", "c11",
                                           "f00", "f01", "f10", "f11")], axis=1)
    np.testing.assert_allclose(out, ref_state, rtol=1e-12, atol=1e-12)


def check_fd(orc, cfg, state, lam, mu, samples, rng, rel=1e-4):
    n = len(lam)
    weights = rng.standard_normal((n, 12)) / n
    b0, gl, gm, _ = orc.diff_backward(state, lam, mu, cfg, weights)
    h = 1e-4
    for _ in range(samples):
        p = int(rng.integers(n))
        j = int(rng.integers(12))
        sp, sm = state.copy(), state.copy()
        sp[p, j] += h
        sm[p, j] -= h
        fd = (loss_of(orc, cfg, sp, lam, mu, weights) - loss_of(orc, cfg, sm, lam, mu, weights)) / (2 * h)
        assert fd == pytest.approx(b0[p, j], rel=rel, abs=1e-9), (p, j)
    for grad, which in ((gl, "lambda"), (gm, "mu")):
        d = rng.choice([-1.0, 1.0], size=n)
        if which == "lambda":
            fd = (loss_of(orc, cfg, state, lam + h * d, mu, weights) - loss_of(orc, cfg, state, lam - h * d, mu, weights))
        else:
            fd = (loss_of(orc, cfg, state, lam, mu + h * d, weights) - loss_of(orc, cfg, state, lam, mu - h * d, weights))
        fd /= 2 * h
        assert abs(fd) > 1e-7, "the case must exercise the stress"
        assert fd == pytest.approx(grad @ d, rel=1e-3), which
    return b0, gl, gm


def test_adjoint_matches_central_differences_floor_impact(pkg, orc):
    state, lam, mu, config, target = small_block(pkg, side=10, steps=120)
    cfg = orc.diff_config(config)
    final = orc.diff_forward(state, lam, mu, cfg)
    assert final[:, 1].min() < 8.0, "the block must reach the wall cushion"
    check_fd(orc, cfg, state, lam, mu, samples=8, rng=np.random.default_rng(11))
    # and the centre-of-mass loss of the differentiable block
    _, bar = orc.com_loss(final, target)
    b0, gl, gm, _ = orc.diff_backward(state, lam, mu, cfg, bar)
    h, rng = 1e-4, np.random.default_rng(21)
    com = lambda s, la, m: orc.com_loss(orc.diff_forward(s, la, m, cfg), target)[0]  # noqa: E731
    for p in rng.integers(len(lam), size=4):
        sp, sm = state.copy(), state.copy()
        sp[p, 3] += h
        sm[p, 3] -= h
        assert (com(sp, lam, mu) - com(sm, lam, mu)) / (2 * h) == pytest.approx(b0[p, 3], rel=1e-4)
    d = rng.choice([-1.0, 1.0], size=len(lam))
    fd = (com(state, lam, mu + h * d) - com(state, lam, mu - h * d)) / (2 * h)
    assert fd == pytest.approx(gm @ d, rel=1e-3)


def test_adjoint_matches_central_differences_with_substeps(pkg, orc):
    # dt 0.05 with a shearing velocity field: dt*|C| > 0.25, so the F update
    # takes several substeps (updateDeformationGradient).
    state, lam, mu, config, target = small_block(pkg, side=4, steps=6, cells=32, origin=(12.0, 12.0),
                                                 dt=0.05, lam=5.0, mu=3.0)
    state[:, 2] += 14.0 * (state[:, 1] - state[:, 1].mean())
    cfg = orc.diff_config(config)
    # confirm the substep path is taken on the first step
    one = orc.diff_step(state, lam, mu, orc.diff_config(config, steps=1))
    row = np.maximum(np.abs(one[:, 4]) + np.abs(one[:, 5]), np.abs(one[:, 6]) + np.abs(one[:, 7]))
    assert (config["dt"] * row > 0.25).any()
    check_fd(orc, cfg, state, lam, mu, samples=8, rng=np.random.default_rng(12))


def test_adjoint_matches_central_differences_wall_friction(pkg, orc):
    state, lam, mu, config, target = small_block(pkg, side=8, steps=80, origin=(4.0, 6.0))
    config["wallFriction"] = 0.4
    state[:, 2] -= 6.0  # toward the left wall as well
    cfg = orc.diff_config(config)
    check_fd(orc, cfg, state, lam, mu, samples=6, rng=np.random.default_rng(13))


def test_adjoint_free_flight_closed_form(pkg, orc):
    state, lam, mu, config, target = small_block(pkg, side=8, steps=40, cells=64, origin=(24.0, 40.0))
    cfg = orc.diff_config(config)
    final = orc.diff_forward(state, lam, mu, cfg)
    n, T, dt = len(lam), config["steps"], config["dt"]
    # the block stays clear of the wall cushions (8 cells)
    assert final[:, :2].min() > 10.0 and final[:, :2].max() < 54.0
    L, bar = orc.com_loss(final, target)
    b0, gl, gm, conf = orc.diff_backward(state, lam, mu, cfg, bar)
    assert conf == 0
    g = bar[0, :2]  # dL/dx_p(T) = 2 (com - target) / N, same for every p
    np.testing.assert_allclose(b0[:, 0:2], np.broadcast_to(g, (n, 2)), rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(b0[:, 2:4], np.broadcast_to(T * dt * g, (n, 2)), rtol=1e-9, atol=1e-14)
    scale = np.abs(g).max()
    assert np.abs(gl).max() < 1e-12 * scale and np.abs(gm).max() < 1e-12 * scale
