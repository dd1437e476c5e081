"""CPU: oracle.contact_near_threshold flags exactly the nodes whose water/oil
contact decision (src/lib/mpm/reference/solver.ts:1293-1409) lies within
float32 rounding of its threshold, on hand-built node sums whose margins are
known in closed form."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc

CELLS = 8
W = H = CELLS + 3


def _grid(wV, oV, wm=1.0, om=0.8, wq=(0.0, 0.0), oq=(0.0, 0.0)):
    g = np.zeros((W * H, len(orc.GRID_FIELDS)))
    f = {n: i for i, n in enumerate(orc.GRID_FIELDS)}
    g[:, f["wV"]] = np.asarray(wV, float).ravel()
    g[:, f["oV"]] = np.asarray(oV, float).ravel()
    g[:, f["wm"]] = wm
    g[:, f["om"]] = om
    g[:, f["wqx"]], g[:, f["wqy"]] = wq
    g[:, f["oqx"]], g[:, f["oqy"]] = oq
    return g


def _cfg(gravity=(0.0, 90.0)):
    return orc.make_config(1e-4, CELLS, CELLS, 1.0, 0.02, 1.0, gravity)


def _ramps(slope_w, slope_o):
    y = np.arange(H, dtype=float)[:, None] * np.ones((1, W))
    return 1.0 + slope_w * y, 1.0 + slope_o * y


def test_flip_margin_at_the_threshold_is_flagged():
    # oil volume rising with y, water falling: n = +y, align = 1.  With
    # densities wd = wm / wV, od = om / oV, contrast = (od - wd) / max: chosen so
    # contrast * align + 1e-3 is exactly 0 at the centre node, then moved
    # well away from it.
    wV, oV = _ramps(-0.01, 0.01)
    sx = sy = 5
    wd = 1.0 / wV[sy, sx]
    od = wd * (1.0 - 1e-3)  # contrast = -1e-3 at the centre
    om = np.full((H, W), od) * oV
    g = _grid(wV, oV, wm=1.0, om=om.ravel())
    near, contact = orc.contact_near_threshold(g, _cfg())
    assert contact > 0
    assert near[sy, sx], "a flip margin of 0 lies within rounding"
    od2 = wd * 0.9  # contrast = -0.1: far from the threshold
    g2 = _grid(wV, oV, wm=1.0, om=(np.full((H, W), od2) * oV).ravel())
    near2, _ = orc.contact_near_threshold(g2, _cfg())
    assert not near2[sy, sx]


def test_cancelling_gradients_are_flagged():
    # equal water and oil volume gradients: g = 0, the normal is undetermined
    wV, oV = _ramps(0.01, 0.01)
    near, _ = orc.contact_near_threshold(_grid(wV, oV), _cfg())
    interior = near[2:-2, 2:-2]
    assert interior.all()


def test_impulse_margin_at_the_threshold_is_flagged():
    # no flip (contrast 0: equal densities), n = +y; rel = (vo - vw) . n is set
    # so that rel equals minsep = 0.25 * overlap * dx / 0.05 at the centre
    wV, oV = _ramps(-0.01, 0.01)
    sx = sy = 5
    overlap = min(1.0, 2 * min(wV[sy, sx], oV[sy, sx]) / (wV[sy, sx] + oV[sy, sx]))
    minsep = 0.25 * overlap * 1.0 / 0.05
    wm = wV.copy()
    om = oV.copy()  # densities 1 and 1: contrast 0
    g = _grid(wV, oV, wm=wm.ravel(), om=om.ravel(), oq=(0.0, 0.0))
    f = {n: i for i, n in enumerate(orc.GRID_FIELDS)}
    g[:, f["oqy"]] = (minsep * om).ravel()  # oil moves up at exactly minsep, water at rest
    near, _ = orc.contact_near_threshold(g, _cfg(gravity=(0.0, 0.0)))
    assert near[sy, sx]
    g[:, f["oqy"]] = (3 * minsep * om).ravel()
    near2, _ = orc.contact_near_threshold(g, _cfg(gravity=(0.0, 0.0)))
    assert not near2[sy, sx]


def test_particles_touching_follow_the_stencil():
    mask = np.zeros((H, W), bool)
    mask[5, 5] = True
    # stencil base floor(x - 0.5) = 3 covers nodes 4..6
    assert orc.particles_touching(mask, np.array([4.0]), np.array([4.0]), 1.0)[0]
    # base 5 covers nodes 6..8
    assert not orc.particles_touching(mask, np.array([6.0]), np.array([6.0]), 1.0)[0]
