"""GPU: the active-tile queue (mpm_tuning.tile_queue) changes which blocks
run which tiles, never the results.  With the queue on and off the same
scenes step bit for bit alike: the SI Tait tank, a batch of mixed scenes
(solids, Tait, linear liquids on grids of different sizes) and a dense
column, in both P2G modes, with the FP64 diagnostics on (the diagnostic
grid sums of the tiles the queue passes over must read zero) and
through a failing microstep (the queue passes over a failed scene's tiles,
as the step kernels do)."""
from __future__ import annotations

import numpy as np
import pytest

from test_gpu_parity import _batch_scenes

pytestmark = pytest.mark.gpu

SHAPE = {"p2g_tile": 20, "g2p_tile": 8, "g2p_threads": 256, "bucket_capacity": 0}


def same_diagnostics(a, b):
    """The grid sums are per-tile sums in a fixed order and must be equal; the
    particle sums follow the particles' slot order, which the atomic slot
    reservation of the bucketing leaves free, so they agree to rounding."""
    for key in a:
        va, vb = a[key], b[key]
        vals = [(va, vb)] if not isinstance(va, dict) else [(va[k], vb[k]) for k in va]
        for x, y in vals:
            if key.startswith("grid"):
                assert x == y, key
            else:
                assert x == pytest.approx(y, rel=1e-12, abs=1e-12 * max(1.0, abs(y))), key


def factory(pkg, mode):
    return pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver


def run_single(pkg, mode, parts, config, steps, dt, gravity, queue, diag=True):
    s = factory(pkg, mode)(parts, config, device=0, tuning=dict(SHAPE, tile_queue=queue))
    try:
        assert s.ctx.get_tuning().as_dict()["tile_queue"] == queue
        if diag:
            s.set_diagnostics(True)
        s.step_batch(steps, dt, gravity)
        speed, done, status = s.ctx.read_compact()
        recs, ids = s.read_particles(packed=True)
        d = s.read_diagnostics() if diag else None
    finally:
        s.dispose()
    return recs, ids, d, done, list(status)


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_tait_tank_queue_on_equals_off(pkg, mode):
    parts, config, macro_dt, _ = pkg.scenes.tait_tank()
    dt = macro_dt / 224
    on = run_single(pkg, mode, parts, config, 48, dt, pkg.scenes.TAIT_TANK_GRAVITY, 1)
    off = run_single(pkg, mode, parts, config, 48, dt, pkg.scenes.TAIT_TANK_GRAVITY, 2)
    print(f"Tait tank {mode}: done {on[3]}/{off[3]}, status {on[4]}", flush=True)
    assert on[4][0] == 0 and on[3] == off[3] == 48
    assert np.array_equal(on[0], off[0]) and np.array_equal(on[1], off[1])
    same_diagnostics(on[2], off[2])


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_dense_column_queue_on_equals_off(pkg, mode):
    arr, config = pkg.scenes.water_oil_column(cells=300)
    on = run_single(pkg, mode, arr, config, 12, 1.0 / 9600.0, pkg.scenes.WATER_OIL_GRAVITY, 1)
    off = run_single(pkg, mode, arr, config, 12, 1.0 / 9600.0, pkg.scenes.WATER_OIL_GRAVITY, 2)
    assert on[4][0] == 0
    assert np.array_equal(on[0], off[0])
    same_diagnostics(on[2], off[2])


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_batch_queue_on_equals_off(pkg, mode):
    scenes, dts, gravity = _batch_scenes(pkg)
    p2g = pkg.native.MPM_P2G_GATHER if mode == "gather" else pkg.native.MPM_P2G_FIXED_POINT
    results = {}
    for queue in (1, 2):
        b = pkg.MpmMlsGpuBatch(scenes, p2g_mode=p2g, device=0, tuning=dict(SHAPE, tile_queue=queue))
        try:
            for i in range(len(b)):
                b.scene(i).set_diagnostics(True)
            b.step_batch(24, dts, gravity)
            out = []
            for i in range(len(b)):
                sc = b.scene(i)
                sc.check_numerical_status()
                out.append((sc.read_particles(packed=True)[0], sc.read_diagnostics()))
        finally:
            b.dispose()
        results[queue] = out
    print(f"batch of {len(scenes)} scenes, {mode}: compared with the queue on and off", flush=True)
    for i, ((ra, da), (rb, db)) in enumerate(zip(results[1], results[2])):
        assert np.array_equal(ra, rb), f"scene {i}"
        same_diagnostics(da, db)


def test_failing_microstep_with_the_queue(pkg):
    """A microstep that raises the CFL guard in a queued launch rolls back to
    the same committed state and the same flags as without the queue."""
    parts, config, macro_dt, _ = pkg.scenes.tait_tank()
    outs = []
    for queue in (1, 2):
        s = pkg.create_mpm_mls_gpu_solver(parts, config, device=0, tuning=dict(SHAPE, tile_queue=queue))
        try:
            s.step_batch(6, macro_dt / 224, pkg.scenes.TAIT_TANK_GRAVITY)
            # a velocity bound below the tank's speeds fails a later microstep
            s.step_batch(8, macro_dt / 224, pkg.scenes.TAIT_TANK_GRAVITY, cfl_velocity_bound=1e-3)
            speed, done, status = s.ctx.read_compact()
            recs, _ = s.ctx.read_particles()
        finally:
            s.dispose()
        outs.append((recs, done, list(status)))
    print(f"failing microstep: committed {outs[0][1]}, status {outs[0][2]}", flush=True)
    assert outs[0][2][0] != 0 and outs[0][1] < 8
    # the flags and the committed state; the particle index of status word 1
    # names whichever failing particle a block checked before the others
    # stopped, with or without the queue
    assert outs[0][1] == outs[1][1] and outs[0][2][0] == outs[1][2][0]
    assert np.array_equal(outs[0][0], outs[1][0])
