"""GPU: the interaction operations of the solver API (pick, outlet removal,
clone, drag through a selection), through the C-ABI against a float32
restatement of the reference's kernels in numpy, bit for bit.

* pick (pickParticle, kernels/shaders.ts:3082-3118): the reference scans
  every particle in index order and keeps the first one with the smallest
  float32 squared distance within the radius; the restatement is the same
  scan as an arg-min over (distance, index).
* outlet (removeParticlesInOutlet, :3041-3065): the water and oil particles
  inside the closed box become removed tombstones; the rest of the scene
  then steps bit for bit like a scene created without them.
* clone (cloneObject, :3120-3147): the copy of an object, translated and
  clamped to the domain, becomes a new object after the active slots; the
  scene then steps bit for bit like a scene created with the copy appended
  (bonded triangle pebbles included).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32
REMOVED = 0xFFFFFFFF
DT = 1.0 / 9600.0


def pick_restated(recs: np.ndarray, x: float, y: float, radius: float):
    """pickParticle in float32: index of the first particle (index order)
    with the smallest squared distance <= radius^2, or None."""
    f = recs.view(np.float32)
    live = recs[:, 13] != REMOVED
    ox = f[:, 0] - F32(x)
    oy = f[:, 1] - F32(y)
    d2 = ox * ox + oy * oy
    r2 = F32(radius) * F32(radius)
    ok = live & (d2 <= r2)
    if not ok.any():
        return None
    cand = np.flatnonzero(ok)
    best = d2[cand].min()
    return int(cand[d2[cand] == best][0])


def column(pkg, cells=200, seed=5):
    arr, config = pkg.scenes.water_oil_column(cells=cells, seed=seed)
    return arr, dict(config, capacity=2 * cells * cells)


@pytest.mark.parametrize("steps", [0, 12])
def test_pick_is_the_reference_scan(pkg, steps):
    arr, config = column(pkg, 400)
    s = pkg.create_mpm_mls_gpu_solver(arr, config, device=0)
    try:
        if steps:
            s.step_batch(steps, DT, pkg.scenes.WATER_OIL_GRAVITY)
        recs, ids = s.read_particles(packed=True)
        rng = np.random.default_rng(11)
        queries = [(float(x), float(y), r) for (x, y), r in
                   zip(rng.uniform(0, 400, (12, 2)), [0.05, 0.3, 0.7, 1.5, 3.0, 12.0] * 2)]
        queries += [(0.0, 0.0, 0.0), (400.0, 400.0, 0.0), (200.0, 200.0, 400.0)]
        hits = 0
        for x, y, r in queries:
            got = s.pick({"x": x, "y": y}, r)
            want = pick_restated(recs, x, y, r)
            if want is None:
                assert got is None, (x, y, r)
                continue
            hits += 1
            f = recs[want].view(np.float32)
            assert got["particleIndex"] == want, (x, y, r)
            assert (got["anchor"]["x"], got["anchor"]["y"]) == (float(f[0]), float(f[1]))
            assert got["objectId"] == int(ids[want])
        print(f"pick after {steps} microsteps: {hits} of {len(queries)} queries hit a particle")
        assert hits >= 8 and hits < len(queries)
    finally:
        s.dispose()


def test_pick_ties_take_the_lowest_index(pkg):
    """Three particles at one position: the reference's serial scan keeps the
    first, so the pick is the lowest index."""
    arr, config = column(pkg, 32)
    x = np.asarray(arr.x).copy()
    y = np.asarray(arr.y).copy()
    for i in (700, 300, 512):
        x[i], y[i] = 10.25, 11.75
    arr = pkg.layout.ParticleArrays(**{**arr.__dict__, "x": x, "y": y})
    s = pkg.create_mpm_mls_gpu_solver(arr, config, device=0)
    try:
        got = s.pick({"x": 10.3, "y": 11.7}, 1.0)
        assert got["particleIndex"] == 300
    finally:
        s.dispose()


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_outlet_removal_then_stepping(pkg, mode):
    arr, config = column(pkg, 120)
    # a few solids inside the outlet box stay
    model = np.asarray(arr.model).copy()
    model[[5440, 5441, 5442]] = 0
    arr = pkg.layout.ParticleArrays(**{**arr.__dict__, "model": model,
                                       "lambda_": np.where(model == 0, 400.0, np.asarray(arr.lambda_)),
                                       "mu": np.where(model == 0, 200.0, 0.0)})
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    g = pkg.scenes.WATER_OIL_GRAVITY
    box = {"minX": 20.0, "minY": 30.5, "maxX": 70.25, "maxY": 60.0}
    s = factory(arr, config, device=0)
    try:
        s.step_batch(6, DT, g)
        before, ids = s.read_particles(packed=True)
        f = before.view(np.float32)
        liquid = (before[:, 13] == 2) | (before[:, 13] == 3)
        inside = liquid & (f[:, 0] >= F32(box["minX"])) & (f[:, 1] >= F32(box["minY"])) & \
            (f[:, 0] <= F32(box["maxX"])) & (f[:, 1] <= F32(box["maxY"]))
        assert any(before[i, 13] == 0 and f[i, 0] >= 20 and f[i, 0] <= 70.25 and 30.5 <= f[i, 1] <= 60
                   for i in (5440, 5441, 5442))
        removed = s.remove_particles_in_outlet(box)
        assert removed == int(inside.sum()) > 0
        assert s.particle_count == len(arr) - removed
        after, _ = s.read_particles(packed=True)
        want = before.copy()
        want[inside, 13] = REMOVED
        assert np.array_equal(after, want)
        assert s.pick({"x": float(f[inside][0, 0]), "y": float(f[inside][0, 1])}, 0.0) is None
        assert len(s.read_particles()) == len(arr) - removed
        s.step_batch(6, DT, g)
        s.check_numerical_status()
        got, _ = s.read_particles(packed=True)
    finally:
        s.dispose()
    # the same state without the removed particles, created fresh
    keep = ~inside
    fresh = pkg.layout.unpack_particles(before[keep], ids[keep])
    t = factory(fresh, dict(config, capacity=int(keep.sum())), device=0)
    try:
        t.step_batch(6, DT, g)
        t.check_numerical_status()
        ref, _ = t.read_particles(packed=True)
    finally:
        t.dispose()
    assert np.array_equal(got[keep], ref)
    assert np.array_equal(got[inside], want[inside])  # tombstones keep their last state


REST = {"C": {"m00": 0.0, "m01": 0.0, "m10": 0.0, "m11": 0.0}, "F": {"m00": 1.0, "m01": 0.0, "m10": 0.0, "m11": 1.0}}


def blocks_scene(pkg):
    """Water (object 0) with a Neo-Hookean block (object 4) and a powder
    block (object 9) on a 64x64 grid."""
    parts = []
    for j in range(20):
        for i in range(40):
            parts.append({"position": {"x": 4.25 + 0.5 * i, "y": 2.25 + 0.5 * j}, "velocity": {"x": 0.3, "y": 0.0},
                          "material": {"model": "water", "lambda": 400.0, "mu": 0.0, "density": 1.0},
                          "volume": 0.25, "objectId": 0, **REST})
    for j in range(6):
        for i in range(6):
            parts.append({"position": {"x": 30.25 + 0.5 * i, "y": 30.25 + 0.5 * j}, "velocity": {"x": -1.0, "y": 0.5},
                          "material": {"model": "neo-hookean", "lambda": 400.0, "mu": 200.0},
                          "volume": 0.25, "objectId": 4, **REST})
    for j in range(4):
        for i in range(5):
            parts.append({"position": {"x": 12.25 + 0.5 * i, "y": 40.25 + 0.5 * j}, "velocity": {"x": 0.0, "y": 0.0},
                          "material": {"model": "powder", "lambda": 400.0, "mu": 200.0, "frictionCoefficient": 0.5},
                          "volume": 0.25, "objectId": 9, **REST})
    config = {"grid": {"cellsX": 64, "cellsY": 64, "cellSize": 1.0}, "particleMass": 0.25, "particleVolume": 0.25,
              "material": {"model": "water", "lambda": 400.0, "mu": 0.0, "density": 1.0}, "capacity": 2000}
    return parts, config


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_clone_then_stepping(pkg, mode):
    parts, config = blocks_scene(pkg)
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    g = {"x": 0.0, "y": -9.0}
    offset = {"x": 25.0, "y": 32.0}   # the top row of the copy is clamped at the top wall
    s = factory(parts, config, device=0)
    try:
        s.step_batch(5, 1e-3, g)
        sel = s.pick({"x": 31.5, "y": 31.5}, 1.0)
        assert sel["objectId"] == 4
        before, ids = s.read_particles(packed=True)
        assert s.clone(sel, offset) is True
        after, ids_after = s.read_particles(packed=True)
        n = len(parts)
        assert len(after) == n + 36 and s.particle_count == n + 36
        src = before[800:836]
        copy = src.copy()
        fs = src.view(np.float32)
        cx = np.minimum(np.maximum(fs[:, 0] + F32(offset["x"]), F32(0)), F32(64))
        cy = np.minimum(np.maximum(fs[:, 1] + F32(offset["y"]), F32(0)), F32(64))
        copy[:, 0], copy[:, 1] = cx.view(np.uint32), cy.view(np.uint32)
        assert np.array_equal(after[:n], before)
        assert np.array_equal(after[n:], copy)
        assert np.all(ids_after[n:] == 10) and (cy == F32(64)).any()  # next object id; clamped copies
        s.step_batch(5, 1e-3, g)
        s.check_numerical_status()
        got, _ = s.read_particles(packed=True)
    finally:
        s.dispose()
    fresh = pkg.layout.unpack_particles(np.concatenate([before, copy]), np.concatenate([ids, np.full(36, 10)]))
    t = factory(fresh, config, device=0)
    try:
        t.step_batch(5, 1e-3, g)
        t.check_numerical_status()
        ref, _ = t.read_particles(packed=True)
    finally:
        t.dispose()
    assert np.array_equal(got, ref)


def test_clone_of_a_bonded_pebble_adds_a_body(pkg):
    """Gate 4: cloning a pebble (a bonded triangle object) adds a body; the
    scene then steps like one created with the copy appended."""
    parts, config, gravity = pkg.scenes.gate4_scene()
    config = dict(config, capacity=len(parts) + 30)
    s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        s.step_batch(2, 1.0 / 2400.0, gravity)
        before, ids = s.read_particles(packed=True)
        vertex = np.flatnonzero(before[:, 19] == 1)[7]
        f = before[vertex].view(np.float32)
        sel = s.pick({"x": float(f[0]), "y": float(f[1])}, 0.0)
        assert sel["particleIndex"] == vertex
        rng = s._object_ranges[sel["objectId"]]
        assert rng["count"] == 3
        offset = {"x": 7.0, "y": 400.0}
        assert s.clone(sel, offset) is True
        after, ids_after = s.read_particles(packed=True)
        assert np.array_equal(after[-3:, 19], [1, 2, 3])
        s.step_batch(2, 1.0 / 2400.0, gravity)
        s.check_numerical_status()
        got, _ = s.read_particles(packed=True)
    finally:
        s.dispose()
    fresh = pkg.layout.unpack_particles(after, ids_after)
    cfg = dict(config, capacity=len(after))
    t = pkg.create_mpm_mls_gpu_solver(fresh, cfg, device=0)
    try:
        t.step_batch(2, 1.0 / 2400.0, gravity)
        t.check_numerical_status()
        ref, _ = t.read_particles(packed=True)
    finally:
        t.dispose()
    assert np.array_equal(got, ref)


def test_clone_is_refused_after_an_outlet_removal(pkg):
    parts, config = blocks_scene(pkg)
    s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        sel = s.pick({"x": 31.5, "y": 31.5}, 1.0)
        assert s.remove_particles_in_outlet({"minX": 0.0, "minY": 0.0, "maxX": 10.0, "maxY": 10.0}) > 0
        assert s.clone(sel, {"x": 1.0, "y": 1.0}) is False
    finally:
        s.dispose()


def test_drag_through_a_selection(pkg, orc):
    """beginDrag writes the reference's uniform (target, radius, stiffness 20,
    damping 8): one microstep equals the oracle's step with that drag."""
    parts, config = blocks_scene(pkg)
    s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        sel = s.pick({"x": 31.5, "y": 31.5}, 1.0)
        s.begin_drag(sel, {"x": 33.0, "y": 33.5}, 2.5)
        state = pkg.layout.validate_input(parts, config)
        s.step(1e-3, {"x": 0.0, "y": 0.0})
        s.check_numerical_status()
        got, _ = s.read_particles(packed=True)
        s.end_drag()
    finally:
        s.dispose()
    cfg = orc.make_config(float(F32(1e-3)), 64, 64, 1.0, 0.25, 0.25, (0.0, 0.0), 0.0,
                          drag=(33.0, 33.5, 2.5, 20.0, 8.0))
    exp, _ = orc.step_f64(orc.particles_from_packed(state.packed), cfg)
    want = np.stack([exp[k] for k in ("x", "y", "vx", "vy")], axis=1)
    have = got.view(np.float32)[:, :4].astype(np.float64)
    assert np.abs(have - want).max() <= 1e-5
    plain, _ = orc.step_f64(orc.particles_from_packed(state.packed),
                            orc.make_config(float(F32(1e-3)), 64, 64, 1.0, 0.25, 0.25))
    assert np.abs(plain["vx"] - exp["vx"]).max() > 1e-3  # the drag acted
