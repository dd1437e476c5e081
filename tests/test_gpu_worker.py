"""GPU: the worker protocol (reference/cpu-simulation.worker.ts:32-193) over
the MI355X solver on the Gate 4 code grid.  The worker's batches (runs of
microsteps with one gravity sent as one batch) must leave the particles bit
for bit where a solver stepped one microstep at a time with the schedule
computed here lands them, and its render snapshot must restate the
particles, liquid areas and bond flags read back from the GPU."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = {"x": 0.0, "y": 90.0}


def schedule(cfl, t0, batch, amp):
    micro = batch.microstepPlan.microDt
    out = []
    for macro in range(batch.macroStepCount):
        for m in range(batch.microstepPlan.stepCount):
            osc = cfl.mean_sinusoidal_acceleration(t0 + (macro * batch.macroDt + m * micro), micro, amp)
            out.append((micro, {"x": G["x"] + osc, "y": G["y"]}))
    return out


def test_worker_batches_match_single_microsteps(pkg):
    cfl = pkg.cfl
    w = pkg.worker.MpmSimulationWorker(device=0)
    parts, config, _ = pkg.scenes.gate4_scene(mpm_grid=False, liquid_area=400.0)
    twin = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    try:
        print("worker: initialize", flush=True)
        r = w.handle({"type": "initialize", "generation": 1, "liquidParticleDiameter": 20, "gravity": G})
        assert r["type"] == "ready", r
        plan = []
        t = 0.0
        first, _ = cfl.calculate_cfl_step_batch(
            cfl.derive_cfl_bounds_packed(twin.state.packed, twin.state.particle_mass_f32).velocityBound,
            90.0, 0.0, 0.0, 0.0, 0.0, accumulatorSeconds=0.0, elapsedSeconds=1 / 240, simulationSpeed=1.0,
            fixedStepSeconds=1 / 240, cellSize=10.0,
            waveSpeedBound=cfl.derive_cfl_bounds_packed(twin.state.packed, twin.state.particle_mass_f32).waveSpeedBound,
            maxMicrosteps=256, courantLimit=cfl.DEFAULT_MPM_COURANT_LIMIT, allowPartialBatch=True)
        plan += schedule(cfl, t, first, 0.0)
        t += first.simulatedSeconds

        print("worker: advance 0.05 s with oscillation 30", flush=True)
        r = w.handle({"type": "advance", "generation": 2, "accumulatorSeconds": 0.05, "maxMicrosteps": 256,
                      "gravity": G, "oscillationAmplitude": 30.0})
        assert r["type"] == "advanced", r
        assert r["simulatedSeconds"] == pytest.approx(12 / 240, rel=1e-12)
        # the twin plans the same batch from its own particles (the state
        # the worker's batch started from)
        for dt, g in plan:
            twin.step_batch(1, dt, g, bond_reference_dt=1 / 240)
        rec, _ = twin.ctx.read_particles()
        b = cfl.derive_cfl_bounds_packed(rec, twin.state.particle_mass_f32)
        batch, _ = cfl.calculate_cfl_step_batch(
            b.velocityBound, 90.0, 30.0, 0.0, 0.0, 0.0, accumulatorSeconds=0.05, elapsedSeconds=0.0,
            simulationSpeed=1.0, fixedStepSeconds=1 / 240, cellSize=10.0, waveSpeedBound=b.waveSpeedBound,
            maxMicrosteps=256, courantLimit=cfl.DEFAULT_MPM_COURANT_LIMIT, allowPartialBatch=True)
        assert batch.microStepCount == r["microStepCount"] and batch.simulatedSeconds == r["simulatedSeconds"]
        for dt, g in schedule(cfl, t, batch, 30.0):
            twin.step_batch(1, dt, g, bond_reference_dt=1 / 240)
        twin.check_numerical_status()

        got, gid = w.solver.read_particles(packed=True)
        want, wid = twin.read_particles(packed=True)
        print(f"worker: {r['microStepCount']} microsteps, {r['workerMilliseconds']:.2f} ms", flush=True)
        assert np.array_equal(got, want) and np.array_equal(gid, wid)

        snap = r["snapshot"]
        rec, _ = w.solver.ctx.read_particles()
        f = rec.view(np.float32)
        np.testing.assert_array_equal(snap["positions"].reshape(-1, 2), f[:, 0:2].astype(np.float64))
        liquid = (rec[:, 13] == 2) | (rec[:, 13] == 3)
        det = f[:, 8].astype(np.float64) * f[:, 11] - f[:, 9].astype(np.float64) * f[:, 10]
        np.testing.assert_array_equal(snap["currentLiquidAreas"][liquid],
                                      (f[liquid, 17].astype(np.float64) * det[liquid]).astype(np.float32))
        assert np.all(snap["currentLiquidAreas"][~liquid] == 0)
        assert np.all(np.isfinite(snap["positions"]))
        flags = [1 if bond["broken"] else 0 for bond in w.solver.read_fracture_bonds()]
        assert snap["brokenFractureBonds"].tolist() == flags

        print("worker: step", flush=True)
        r = w.handle({"type": "step", "generation": 3, "maxMicrosteps": 256, "gravity": G,
                      "oscillationAmplitude": 30.0})
        assert r["type"] == "stepped" and r["simulatedSeconds"] == pytest.approx(1 / 240)
        assert math.isclose(w.simulated_seconds, 14 / 240, rel_tol=1e-12)
    finally:
        twin.dispose()
        w.dispose()
