"""Pinning of the oracle's config-2 step (orc_step_bodies_f64): bonded
triangle bodies and fracture bonds (reference/solver.ts:1023-1874,
reference/solid-bonds.ts:68-108), by closed forms and invariants that do
not reuse the oracle's code:

* triangleDistanceSample: signed distances and normals of an equilateral
  triangle from elementary geometry;
* trianglePairContact: separating-axis results of two right triangles;
* a lone pebble: the step with bodies equals the plain step followed by the
  two-iteration edge solve (computed here), with velocity += displacement/dt;
* momentum: fluid-triangle contact, edge solve and pair contacts are
  internal, so with no gravity and no wall in reach the total momentum is
  conserved to rounding while the contact is active;
* fracture: the damage law on a stretched pair of bonded particles.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

SQ3 = math.sqrt(3.0)


def equilateral(cx, cy, side):
    return np.array([cx - side / 2, cy - side / (2 * SQ3), cx + side / 2, cy - side / (2 * SQ3), cx, cy + side / SQ3])


def make_particles(orc, rows):
    p = np.zeros(len(rows), orc.PARTICLE_DTYPE)
    for i, r in enumerate(rows):
        p[i]["x"], p[i]["y"] = r["pos"]
        p[i]["vx"], p[i]["vy"] = r.get("vel", (0.0, 0.0))
        p[i]["f00"] = p[i]["f11"] = 1.0
        p[i]["plastic_volume"] = 1.0
        p[i]["lambda_"], p[i]["mu"] = r.get("lam", 20.0), r.get("mu", 10.0)
        p[i]["density"] = r.get("density", 0.0)
        p[i]["volume"] = r.get("volume", 100.0)
        p[i]["param"] = r.get("param", 0.0)
        p[i]["model"] = r.get("model", orc.MODEL_NEO_HOOKEAN)
        p[i]["bonded_vertex"] = r.get("vertex", 0)
    return p


def test_triangle_sample_closed_forms(orc):
    side = 6.0
    t = equilateral(0.0, 0.0, side)
    inradius = side / (2 * SQ3)
    s = orc.triangle_sample(0.0, 0.0, t)              # centroid: inside, nearest edge at the inradius
    assert s[0] == pytest.approx(-inradius, rel=1e-12)
    s = orc.triangle_sample(0.0, -inradius - 2.0, t)  # below the bottom edge
    assert s[0] == pytest.approx(2.0, rel=1e-12)
    assert s[1] == pytest.approx(0.0, abs=1e-12) and s[2] == pytest.approx(-1.0, rel=1e-12)
    assert s[3] == pytest.approx(0.5) and s[4] == pytest.approx(0.5) and s[5] == 0.0
    top = (0.0, side / SQ3)
    s = orc.triangle_sample(top[0], top[1] + 3.0, t)  # above the apex: distance to the vertex
    assert s[0] == pytest.approx(3.0, rel=1e-12)
    assert s[1] == pytest.approx(0.0, abs=1e-12) and s[2] == pytest.approx(1.0, rel=1e-12)
    # reversed winding gives the same field
    r = t.reshape(3, 2)[::-1].reshape(-1)
    assert orc.triangle_sample(0.0, -inradius - 2.0, r)[0] == pytest.approx(2.0, rel=1e-12)
    with pytest.raises(orc.OracleError):                # degenerate
        orc.triangle_sample(0.0, 0.0, [0, 0, 1, 1, 2, 2])


def test_triangle_pair_contact_closed_forms(orc):
    a = [0.0, 0.0, 2.0, 0.0, 0.0, 2.0]
    b = [1.5, 0.0, 3.5, 0.0, 1.5, 2.0]                  # b overlaps a by 0.5 along x at the base
    c = orc.triangle_pair_contact(a, b, 0.1)
    # smallest overlap axis: the hypotenuse normal (1,1)/sqrt2 of a, overlap (2 - 1.5)/sqrt2;
    # the normal points from b towards a (centroid difference along -x)
    ov = 0.5 / math.sqrt(2.0)
    assert c is not None
    assert c[2] == pytest.approx(0.1 + ov, rel=1e-12)
    assert c[0] == pytest.approx(-1 / math.sqrt(2), rel=1e-12) and c[1] == pytest.approx(-1 / math.sqrt(2), rel=1e-12)
    far = [5.0, 0.0, 7.0, 0.0, 5.0, 2.0]
    assert orc.triangle_pair_contact(a, far, 0.1) is None
    # separated by 0.05 along x (the largest separating gap; the hypotenuse
    # axis separates by only 0.05/sqrt2): penetration = skin - 0.05, normal -x
    near = [2.05, 0.0, 4.05, 0.0, 2.05, 2.0]
    c = orc.triangle_pair_contact(a, near, 0.1)
    assert c is not None and c[2] == pytest.approx(0.1 - 0.05, rel=1e-9)
    assert c[0] == -1.0 and c[1] == pytest.approx(0.0, abs=1e-15)


def gate4_like_config(orc, cells=40, dt=1.0 / 240.0, gravity=(0.0, 0.0)):
    return orc.make_config(dt, cells, cells, 10.0, 4.0, 100.0, gravity, 0.0)


def test_lone_pebble_is_plain_step_plus_edge_solve(orc):
    side, rest = 132.0, 132.0
    t = equilateral(200.0, 200.0, side * 1.08)          # stretched: the edge solve acts
    vel = [(30.0, -12.0), (-5.0, 40.0), (7.0, 3.0)]
    rows = [{"pos": (t[2 * v], t[2 * v + 1]), "vel": vel[v], "vertex": v + 1} for v in range(3)]
    p = make_particles(orc, rows)
    cfg = gate4_like_config(orc)
    bc = orc.BodyConfig(rest, 0.35, 1.0 / 240.0)
    out, _, _ = orc.step_bodies_f64(p, np.array([3, 3, 3]), cfg, bc)
    plain_in = p.copy()
    plain_in["bonded_vertex"] = 0
    plain, _ = orc.step_f64(plain_in, cfg)
    pos = np.stack([plain["x"], plain["y"]], axis=1)
    orig = pos.copy()
    k = min(0.35 * (cfg.dt / (1 / 240.0)) ** 2, 1.0)
    for _ in range(2):
        for a, b in ((0, 1), (1, 2), (2, 0)):
            d = pos[b] - pos[a]
            dist = math.hypot(*d)
            corr = d * ((dist - rest) * k * 0.5 / dist)
            pos[a] += corr
            pos[b] -= corr
    np.testing.assert_allclose(np.stack([out["x"], out["y"]], axis=1), pos, rtol=0, atol=1e-9)
    vexp = np.stack([plain["vx"], plain["vy"]], axis=1) + (pos - orig) / cfg.dt
    np.testing.assert_allclose(np.stack([out["vx"], out["vy"]], axis=1), vexp, rtol=1e-12, atol=1e-9)
    # the edge solve conserves the body's momentum
    assert np.sum(out["vx"]) == pytest.approx(np.sum(plain["vx"]), abs=1e-9)
    assert np.sum(out["vy"]) == pytest.approx(np.sum(plain["vy"]), abs=1e-9)


def test_contacts_conserve_momentum(orc):
    """Two pebbles of different objects in contact, moving into each other,
    inside a water/oil blob: no gravity, walls far away."""
    rows = []
    t1 = equilateral(190.0, 200.0, 132.0)
    t2 = equilateral(300.0, 205.0, 132.0)                # overlaps t1
    for t, v, vel in ((t1, 0, (60.0, 5.0)), (t2, 3, (-70.0, 0.0))):
        for k in range(3):
            rows.append({"pos": (t[2 * k], t[2 * k + 1]), "vel": vel, "vertex": k + 1})
    rng = np.random.default_rng(3)
    for i in range(120):                                  # liquid around and through the pebbles
        model = orc.MODEL_WATER if i % 2 == 0 else orc.MODEL_OIL
        rows.append({"pos": (150 + 200 * rng.random(), 140 + 130 * rng.random()),
                     "vel": tuple(10 * rng.standard_normal(2)), "model": model, "lam": 1000.0, "mu": 0.0,
                     "density": 0.02 if model == orc.MODEL_WATER else 0.016, "volume": 400.0})
    p = make_particles(orc, rows)
    oid = np.array([3, 3, 3, 4, 4, 4] + [1] * 120)
    cfg = gate4_like_config(orc, cells=50)
    bc = orc.BodyConfig(132.0, 0.35, 1.0 / 240.0)
    out, _, d = orc.step_bodies_f64(p, oid, cfg, bc)
    plain_in = p.copy()
    plain_in["bonded_vertex"] = 0
    plain, _ = orc.step_f64(plain_in, cfg)
    # the contact code acted: the pebble velocities differ from the plain step
    assert np.max(np.abs(out["vx"][:6] - plain["vx"][:6])) > 1.0
    assert d.output_px == pytest.approx(d.input_px, rel=1e-10, abs=1e-8)
    assert d.output_py == pytest.approx(d.input_py, rel=1e-10, abs=1e-8)
    assert d.output_mass == pytest.approx(d.input_mass, rel=1e-14)


def test_fracture_damage_law(orc):
    rows = [{"pos": (200.0, 200.0), "vel": (-40.0, 0.0)}, {"pos": (220.0, 200.0), "vel": (40.0, 0.0)},
            {"pos": (240.0, 200.0), "vel": (40.0, 0.0)}]
    p = make_particles(orc, rows)
    p["param"][2] = 0.5                                   # prior damage above the new one is kept
    bonds = np.zeros(2, orc.BOND_DTYPE)
    bonds[0] = (0, 1, 20.0, 0.12, 0.0, 0, 0)
    bonds[1] = (1, 2, 20.0, 0.12, 0.0, 0, 0)
    cfg = gate4_like_config(orc, gravity=(0.0, 0.0))
    bc = orc.BodyConfig(1.0, 0.35, 1.0 / 240.0)
    out, bo, _ = orc.step_bodies_f64(p, np.zeros(3), cfg, bc, bonds)
    plain, _ = orc.step_f64(p, cfg)
    for k, (a, b) in enumerate(((0, 1), (1, 2))):
        d = math.hypot(plain["x"][b] - plain["x"][a], plain["y"][b] - plain["y"][a])
        strain = max(0.0, (d - 20.0) / 20.0)
        broken = strain >= 0.12
        assert bool(bo["broken"][k]) == broken
        assert bo["damage"][k] == pytest.approx(1.0 if broken else min(1.0, strain / 0.12), rel=1e-12)
    assert bo["damage"][0] > 0
    assert out["param"][0] == pytest.approx(bo["damage"][0], rel=1e-12)
    assert out["param"][1] == pytest.approx(max(bo["damage"][0], bo["damage"][1]), rel=1e-12)
    assert out["param"][2] == pytest.approx(max(0.5, bo["damage"][1]), rel=1e-12)
    # a broken bond stays broken with damage 1
    bonds2 = bo.copy()
    bonds2["broken"][0] = 1
    _, bo2, _ = orc.step_bodies_f64(out, np.zeros(3), cfg, bc, bonds2)
    assert bo2["broken"][0] == 1 and bo2["damage"][0] == 1.0


def test_bodies_validation(orc):
    t = equilateral(200.0, 200.0, 132.0)
    rows = [{"pos": (t[2 * v], t[2 * v + 1]), "vertex": v + 1} for v in range(3)]
    p = make_particles(orc, rows)
    cfg = gate4_like_config(orc)
    with pytest.raises(orc.OracleError):                  # vertices of two objects
        orc.step_bodies_f64(p, np.array([3, 3, 4]), cfg, orc.BodyConfig(132.0, 0.35, 1 / 240))
    with pytest.raises(orc.OracleError):                  # missing rest length
        orc.step_bodies_f64(p, np.array([3, 3, 3]), cfg, orc.BodyConfig(0.0, 0.35, 1 / 240))
    q = p.copy()
    q["bonded_vertex"][0] = 2                             # must start at vertex 0
    with pytest.raises(orc.OracleError):
        orc.step_bodies_f64(q, np.array([3, 3, 3]), cfg, orc.BodyConfig(132.0, 0.35, 1 / 240))
    bonds = np.zeros(1, orc.BOND_DTYPE)
    bonds[0] = (0, 1, 20.0, 0.12, 0.0, 0, 0)             # bond on bonded-triangle vertices
    with pytest.raises(orc.OracleError):
        orc.step_bodies_f64(p, np.array([3, 3, 3]), cfg, orc.BodyConfig(132.0, 0.35, 1 / 240), bonds)
