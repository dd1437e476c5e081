"""GPU: a solver restored from a snapshot file continues bit for bit like
the solver the snapshot was taken from: the Gate 4 scene (bonded pebbles and
a fracture solid whose bonds carry damage), and a scene after an outlet
removal and a clone (tombstones and a cloned object), in both P2G modes."""
from __future__ import annotations

import numpy as np
import pytest

from test_gpu_interaction import blocks_scene

pytestmark = pytest.mark.gpu


def continue_and_compare(pkg, s, path, steps, dt, gravity):
    pkg.snapshot.save(s, path)
    s.step_batch(steps, dt, gravity)
    s.check_numerical_status()
    want, wid = s.read_particles(packed=True)
    wbonds = s.read_fracture_bonds()
    t = pkg.snapshot.load(path, device=0)
    try:
        assert t.particle_count == s.particle_count and t._next_object_id == s._next_object_id
        t.step_batch(steps, dt, gravity)
        t.check_numerical_status()
        got, gid = t.read_particles(packed=True)
        gbonds = t.read_fracture_bonds()
    finally:
        t.dispose()
    assert np.array_equal(got, want) and np.array_equal(gid, wid)
    assert gbonds == wbonds
    return got


def test_gate4_snapshot_continues_bit_for_bit(pkg, tmp_path):
    """Float gather only: in fixed point the pebble vertices' 2^24-scaled
    momentum overflows int32 after the first microsteps of this scene, as in
    the reference."""
    parts, config, gravity = pkg.scenes.gate4_scene()
    # stretch the fracture solid's rest lengths so that its bonds strain and
    # gather damage before the snapshot
    bonds = [dict(b, restLength=b["restLength"] * (0.9 if k % 3 == 0 else 1.0))
             for k, b in enumerate(config["fractureBonds"])]
    config = dict(config, fractureBonds=bonds)
    s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        s.step_batch(24, 1.0 / 2400.0, gravity)
        s.check_numerical_status()
        damaged = [b for b in s.read_fracture_bonds() if b["damage"] > 0]
        print(f"Gate 4: {len(damaged)} fracture bonds carry damage at the snapshot", flush=True)
        assert damaged
        continue_and_compare(pkg, s, tmp_path / "gate4.mpm", 3, 1.0 / 2400.0, gravity)
    finally:
        s.dispose()


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_snapshot_after_outlet_and_clone(pkg, tmp_path, mode):
    parts, config = blocks_scene(pkg)
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    g = {"x": 0.0, "y": -9.0}
    s = factory(parts, config, device=0)
    try:
        s.step_batch(4, 1e-3, g)
        sel = s.pick({"x": 31.5, "y": 31.5}, 1.0)
        assert s.clone(sel, {"x": -12.0, "y": 5.0})
        assert s.remove_particles_in_outlet({"minX": 4.0, "minY": 2.0, "maxX": 9.0, "maxY": 6.0}) > 0
        s.step_batch(3, 1e-3, g)
        recs = s.read_particles(packed=True)[0]
        assert np.count_nonzero(recs[:, 13] == 0xFFFFFFFF) > 0
        got = continue_and_compare(pkg, s, tmp_path / "blocks.mpm", 4, 1e-3, g)
        assert np.count_nonzero(got[:, 13] == 0xFFFFFFFF) == np.count_nonzero(recs[:, 13] == 0xFFFFFFFF)
    finally:
        s.dispose()
