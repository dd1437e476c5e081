"""The FP64 reverse pass of the differentiable rollout (oracle/mpm_adjoint.c)
against central finite differences (h = 1e-4) of the FP64 forward on the
full differentiable block itself: 4,096 fixed-corotated particles (64 x 64
lattice, lambda 200, mu 100, density 1) in a 256 x 256 grid, launched into
the floor for 200 steps of 1e-3 s, with the loss the squared distance of the
final centre of mass to its target.  Sampled components of the initial
position and velocity, per-particle lambda and mu, and random directions in
lambda and mu; about 60 FP64 forwards of the whole block.

Bars: 1e-3 relative (the agreement the differentiable path is held to),
with 1e-8 absolute for components at the differences' noise floor (the
loss is ~1, so its rounding and the floor contact leave about 1e-9 in a
difference quotient; several per-particle mu gradients are ~1e-13 there).
The directional derivatives are the strong checks for lambda and mu."""
from __future__ import annotations

import numpy as np
import pytest

H = 1e-4


def c5_fd_samples(pkg, orc, seed=5):
    """The block, its FP64 gradients, and central differences of the loss at
    sampled components: a list of (label, fd, index) with index selecting the
    matching entry of (d/dstate0, d/dlambda, d/dmu)."""
    state, lam, mu, config, target = pkg.scenes.elastic_block()
    cfg = orc.diff_config(config)

    def loss(s, la, m):
        return orc.com_loss(orc.diff_forward(s, la, m, cfg), target)[0]

    rng = np.random.default_rng(seed)
    n = len(lam)
    samples = []
    for p in rng.choice(n, size=6, replace=False):
        for j in (2, 3):  # initial velocity
            sp, sm = state.copy(), state.copy()
            sp[p, j] += H
            sm[p, j] -= H
            samples.append((f"v0[{p},{j - 2}]", (loss(sp, lam, mu) - loss(sm, lam, mu)) / (2 * H), ("state", p, j)))
    for p in rng.choice(n, size=3, replace=False):
        for j in (0, 1):  # initial position
            sp, sm = state.copy(), state.copy()
            sp[p, j] += H
            sm[p, j] -= H
            samples.append((f"x0[{p},{j}]", (loss(sp, lam, mu) - loss(sm, lam, mu)) / (2 * H), ("state", p, j)))
    for which in ("lambda", "mu"):
        for p in rng.choice(n, size=4, replace=False):
            e = np.zeros(n)
            e[p] = H
            if which == "lambda":
                fd = (loss(state, lam + e, mu) - loss(state, lam - e, mu)) / (2 * H)
            else:
                fd = (loss(state, lam, mu + e) - loss(state, lam, mu - e)) / (2 * H)
            samples.append((f"{which}[{p}]", fd, (which, p, None)))
        d = rng.choice([-1.0, 1.0], size=n)
        if which == "lambda":
            fd = (loss(state, lam + H * d, mu) - loss(state, lam - H * d, mu)) / (2 * H)
        else:
            fd = (loss(state, lam, mu + H * d) - loss(state, lam, mu - H * d)) / (2 * H)
        samples.append((f"{which} . d", fd, (which, None, d)))
    return (state, lam, mu, config, target, cfg), samples


def pick(grads, index):
    b0, gl, gm = grads
    kind, p, j = index
    if kind == "state":
        return b0[p, j]
    g = gl if kind == "lambda" else gm
    return g[p] if p is not None else float(g @ j)


def test_fp64_adjoint_matches_central_differences_on_the_full_block(pkg, orc):
    (state, lam, mu, config, target, cfg), samples = c5_fd_samples(pkg, orc)
    final = orc.diff_forward(state, lam, mu, cfg)
    assert final[:, 1].min() < 8.0, "the block must reach the floor cushion"
    _, bar = orc.com_loss(final, target)
    b0, gl, gm, _ = orc.diff_backward(state, lam, mu, cfg, bar)
    for label, fd, index in samples:
        g = pick((b0, gl, gm), index)
        print(f"{label}: FD {fd:+.6e} adjoint {g:+.6e}", flush=True)
        assert g == pytest.approx(fd, rel=1e-3, abs=1e-8), label
