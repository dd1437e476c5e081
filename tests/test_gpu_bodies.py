"""GPU config-2 path (bonded triangle bodies, fluid-triangle contact, pair
contacts, fracture bonds) through the C-ABI against the FP64 oracle step
orc_step_bodies_f64 (itself pinned in tests/test_oracle_bodies.py).

Tolerances: float32 state against the FP64 step.  Per microstep the relative
L2 error of all velocities stays within 1e-4; positions within 4 float32
spacings of the largest coordinate (the state is stored in float32: 4.9e-4
at the Gate 4 scene's y ~ 5,900), plus, for liquid particles, 1e-4 of the
distance the projection out of the triangles moved them (see compare_step);
bond damages and particle damage within 1e-4.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_inputs(pkg, orc, solver, config, dt, gravity):
    rec, oid = solver.read_particles(packed=True)
    cfg = orc.make_config(float(np.float32(dt)), config["grid"]["cellsX"], config["grid"]["cellsY"],
                          config["grid"]["cellSize"], config["particleMass"], config["particleVolume"],
                          (gravity["x"], gravity["y"]), config.get("wallFriction", 0.0))
    bc = orc.BodyConfig(float(np.float32(config.get("bondRestLength", 0.0))),
                        float(np.float32(config.get("bondStiffness", 0.35))),
                        float(np.float32(config.get("bondReferenceDt", dt))))
    return rec, oid, cfg, bc


def oracle_bonds(orc, solver):
    b = solver.read_fracture_bonds()
    out = np.zeros(len(b), orc.BOND_DTYPE)
    for k, d in enumerate(b):
        out[k] = (d["firstParticle"], d["secondParticle"], d["restLength"], d["failureStrain"], d["damage"],
                  1 if d["broken"] else 0, 0)
    return out


def as_f64(recs):
    return recs.view(np.float32)[:, :12].astype(np.float64)


STATE = ("x", "y", "vx", "vy", "c00", "c01", "c10", "c11", "f00", "f01", "f10", "f11")


def compare_step(pkg, orc, solver, config, dt, gravity, tol_v=1e-4, explain=False):
    rec, oid, cfg, bc = oracle_inputs(pkg, orc, solver, config, dt, gravity)
    bonds_before = oracle_bonds(orc, solver) if solver.state.fracture_bonds is not None else None
    solver.step(dt, gravity)
    solver.check_numerical_status()
    got, _ = solver.read_particles(packed=True)
    exp, bonds_after, diag = orc.step_bodies_f64(orc.particles_from_packed(rec), oid, cfg, bc, bonds_before)
    want = np.stack([exp[k] for k in STATE], axis=1)
    have = as_f64(got)
    rel_v = float(np.linalg.norm(have[:, 2:4] - want[:, 2:4]) / np.linalg.norm(want[:, 2:4]))
    scale = max(1.0, float(np.max(np.abs(want[:, :2]))))
    max_x = float(np.max(np.abs(have[:, :2] - want[:, :2])))
    print(f"step vs FP64: rel L2 v {rel_v:.2e}, max |dx| {max_x:.2e} (coordinates up to {scale:.0f})")
    vertex = got[:, 19] != 0
    model = got[:, 13]
    for name, m in (("liquid", (model == 2) | (model == 3)), ("vertex", vertex), ("solid", (model == 0) & ~vertex)):
        if np.any(m):
            dv = np.abs(have[m, 2:4] - want[m, 2:4]).max()
            dx = np.abs(have[m, :2] - want[m, :2]).max()
            print(f"  {name}: max |dv| {dv:.2e} (|v| up to {np.abs(want[m, 2:4]).max():.1f}), max |dx| {dx:.2e}")
    # The projection out of the triangles moves a liquid particle by r - sd
    # (up to ~54 units in the first Gate 4 steps, where the liquid starts
    # inside the pebbles) along a normal interpolated across the field; the
    # float32 rounding of the position it starts from (half a spacing) turns
    # that normal, so the bar adds 1e-4 of each particle's projection
    # displacement (measured against the oracle's step without bodies).
    plain_in = orc.particles_from_packed(rec)
    plain_in["bonded_vertex"] = 0
    plain, _ = orc.step_f64(plain_in, cfg)
    proj = np.hypot(want[:, 0] - plain["x"], want[:, 1] - plain["y"])
    liquid = (got[:, 13] == 2) | (got[:, 13] == 3)
    err = np.abs(have[:, :2] - want[:, :2]).max(axis=1)
    ulp = 2.0 ** (math.floor(math.log2(scale)) - 23)  # float32 spacing of the largest coordinate
    excess = err - np.where(liquid, 1e-4 * proj, 0.0)
    if explain:
        for i in np.argsort(err)[::-1][:3]:
            print(f"  particle {i} model {got[i, 13]}: |dx| {err[i]:.2e}, projection displacement {proj[i]:.3e}")
    assert rel_v <= tol_v
    assert float(np.max(excess)) <= 4 * ulp + 1e-5
    if bonds_before is not None:
        b = oracle_bonds(orc, solver)
        np.testing.assert_allclose(b["damage"], bonds_after["damage"], rtol=0, atol=1e-4)
        assert np.array_equal(b["broken"], bonds_after["broken"])
        dmg = got.view(np.float32)[:, 18].astype(np.float64)
        np.testing.assert_allclose(dmg, exp["param"], rtol=0, atol=1e-4)
    return have, want, diag


def contact_scene(pkg):
    """Two pebbles of different objects overlapping each other inside a
    water/oil blob, and a stretched fracture solid nearby.  Speeds stay below
    ~30, so the fixed-point path's 2^24-scaled momentum words cannot overflow."""
    zero = {"m00": 0.0, "m01": 0.0, "m10": 0.0, "m11": 0.0}
    ident = {"m00": 1.0, "m01": 0.0, "m10": 0.0, "m11": 1.0}
    nh = {"lambda": 20.0, "mu": 10.0, "model": "neo-hookean"}
    rng = np.random.default_rng(5)
    parts = []
    for i in range(240):
        mat = pkg.scenes.GATE4_WATER if i % 2 == 0 else pkg.scenes.GATE4_OIL
        parts.append({"position": {"x": 120 + 180 * rng.random(), "y": 120 + 140 * rng.random()},
                      "velocity": {"x": 8 * rng.standard_normal(), "y": 8 * rng.standard_normal()},
                      "C": dict(zero), "F": dict(ident), "objectId": 1, "volume": 400.0, "material": dict(mat)})
    s3 = math.sqrt(3)
    for k, (cx, cy, vx) in enumerate(((170.0, 190.0, 25.0), (270.0, 195.0, -28.0))):
        for v, (x, y) in enumerate(((cx - 66, cy - 66 / s3), (cx + 66, cy - 66 / s3), (cx, cy + 132 / s3))):
            parts.append({"position": {"x": x, "y": y}, "velocity": {"x": vx, "y": 4.0 * v}, "C": dict(zero),
                          "F": dict(ident), "objectId": 2 + k, "bondedTriangleVertex": v, "volume": 100.0,
                          "material": dict(nh)})
    start = len(parts)
    for row in range(2):
        for col in range(3):
            parts.append({"position": {"x": 360.0 + col * 21.0, "y": 300.0 + row * 21.0},
                          "velocity": {"x": (col - 1) * 28.0, "y": (row - 0.5) * 20.0}, "C": dict(zero),
                          "F": dict(ident), "objectId": 4, "volume": 100.0, "material": dict(nh)})
    bonds = []
    for row in range(2):
        for col in range(3):
            a = start + row * 3 + col
            if col + 1 < 3:
                bonds.append({"firstParticle": a, "secondParticle": a + 1, "restLength": 20.0, "failureStrain": 0.06})
            if row + 1 < 2:
                bonds.append({"firstParticle": a, "secondParticle": a + 3, "restLength": 20.0, "failureStrain": 0.06})
    config = {"grid": {"cellsX": 50, "cellsY": 50, "cellSize": 10.0}, "particleMass": 4.0, "particleVolume": 100.0,
              "bondRestLength": 132.0, "bondReferenceDt": 1.0 / 240.0, "material": dict(nh), "fractureBonds": bonds,
              "capacity": len(parts)}
    return parts, config


@pytest.mark.parametrize("factory,steps", [("create_mpm_mls_gpu_solver", 6),
                                           ("create_experimental_mpm_mls_atomic_gpu_solver", 1)])
def test_contact_scene_matches_oracle_step_by_step(pkg, orc, factory, steps):
    # The contacts accelerate the pebble vertices past ~128 units/s after the
    # first step; at particleMass 4 a vertex's 2^24-scaled momentum word then
    # overflows int32 (fixed-point flag 65536, as in the reference), so the
    # fixed-point variant is compared over one step.
    parts, config = contact_scene(pkg)
    gravity = {"x": 0.0, "y": 0.0}
    dt = 1.0 / 240.0 / 8
    s = getattr(pkg, factory)(parts, config, device=0)
    try:
        for _ in range(steps):
            have, want, diag = compare_step(pkg, orc, s, config, dt, gravity)
        bonds = s.read_fracture_bonds()
        assert any(b["damage"] > 0 for b in bonds)
    finally:
        s.dispose()


def many_bodies_scene(pkg, nx=10, ny=7):
    """nx x ny pebbles (each its own object) on a lattice of pitch 115 < edge
    132, so neighbours in a row overlap and touch, in a sparse water/oil
    field: more bodies than the kernel caches pair tests for (64)."""
    zero = {"m00": 0.0, "m01": 0.0, "m10": 0.0, "m11": 0.0}
    ident = {"m00": 1.0, "m01": 0.0, "m10": 0.0, "m11": 1.0}
    nh = {"lambda": 20.0, "mu": 10.0, "model": "neo-hookean"}
    rng = np.random.default_rng(11)
    s3 = math.sqrt(3)
    parts = []
    for i in range(300):
        mat = pkg.scenes.GATE4_WATER if i % 2 == 0 else pkg.scenes.GATE4_OIL
        parts.append({"position": {"x": 60 + 1200 * rng.random(), "y": 60 + 900 * rng.random()},
                      "velocity": {"x": 6 * rng.standard_normal(), "y": 6 * rng.standard_normal()},
                      "C": dict(zero), "F": dict(ident), "objectId": 1, "volume": 400.0, "material": dict(mat)})
    k = 0
    for r in range(ny):
        for c in range(nx):
            cx, cy = 120.0 + 115.0 * c, 120.0 + 125.0 * r
            vx, vy = 10.0 * rng.standard_normal(), 10.0 * rng.standard_normal()
            for v, (x, y) in enumerate(((cx - 66, cy - 66 / s3), (cx + 66, cy - 66 / s3), (cx, cy + 132 / s3))):
                parts.append({"position": {"x": x, "y": y}, "velocity": {"x": vx, "y": vy + 2.0 * v},
                              "C": dict(zero), "F": dict(ident), "objectId": 2 + k, "bondedTriangleVertex": v,
                              "volume": 100.0, "material": dict(nh)})
            k += 1
    config = {"grid": {"cellsX": 132, "cellsY": 100, "cellSize": 10.0}, "particleMass": 4.0,
              "particleVolume": 100.0, "bondRestLength": 132.0, "bondReferenceDt": 1.0 / 240.0,
              "material": dict(nh), "capacity": len(parts)}
    return parts, config


def test_many_bodies_scene_matches_oracle_step_by_step(pkg, orc):
    # 70 bodies: the pair contacts run without the LDS pair cache
    parts, config = many_bodies_scene(pkg)
    gravity = {"x": 0.0, "y": 0.0}
    dt = 1.0 / 240.0 / 8
    s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        assert s.state.bonded_body_count == 70
        for _ in range(3):
            compare_step(pkg, orc, s, config, dt, gravity)
    finally:
        s.dispose()


def test_gate4_scene_microsteps_match_oracle(pkg, orc):
    parts, config, gravity = pkg.scenes.gate4_scene()
    dt = 1.0 / 240.0 / 16
    s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        for _ in range(3):
            compare_step(pkg, orc, s, config, dt, gravity, explain=True)
    finally:
        s.dispose()


def test_gate4_scene_macrosteps_run_clean(pkg):
    parts, config, gravity = pkg.scenes.gate4_scene()
    s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        for _ in range(4):  # four 1/240 s macrosteps of 16 microsteps (under the liquids' CFL limit)
            s.step_batch(16, 1.0 / 240.0 / 16, gravity)
        s.check_numerical_status()
        recs, oid = s.read_particles(packed=True)
        f = recs.view(np.float32)
        assert np.all(np.isfinite(f[:, :12]))
        # triangle vertices keep their codes and objects; bonds report a state
        assert np.array_equal(recs[3500:3650:3, 19], np.ones(50, np.uint32))
        assert len(s.read_fracture_bonds()) == 53
    finally:
        s.dispose()


def test_bodies_scene_cannot_be_batched(pkg):
    parts, config = contact_scene(pkg)
    with pytest.raises(pkg.native.NativeError) as e:
        pkg.MpmMlsGpuBatch([(parts, config)], device=0)
    assert e.value.code == pkg.native.MPM_EINVAL
