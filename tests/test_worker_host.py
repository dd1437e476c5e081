"""The worker protocol (reference/cpu-simulation.worker.ts:32-193) on the
host, driven with a stand-in solver that records the steps it is given:
request and response shapes, the backlog batches, the per-microstep
oscillation and the error answers.  The oscillation mean is pinned by
quadrature."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import integrate


class _Ctx:
    def __init__(self, state):
        self.state = state

    def read_particles(self):
        return self.state.packed.copy(), self.state.object_ids.copy()


class RecordingSolver:
    """Takes the place of MpmMlsGpuSolver: keeps the validated scene and
    records every step_batch call."""

    def __init__(self, layout, particles, config):
        self.layout = layout
        self.state = layout.validate_input(particles, config)
        self.ctx = _Ctx(self.state)
        self.calls = []
        self.disposed = False

    def step_batch(self, step_count, dt, gravity, bond_reference_dt=None):
        self.calls.append((step_count, dt, dict(gravity), bond_reference_dt))

    def check_numerical_status(self):
        pass

    def read_fracture_bonds(self):
        return self.layout.unpack_fracture_bonds(self.state.fracture_bonds)

    def dispose(self):
        self.disposed = True

    def microsteps(self):
        out = []
        for count, dt, g, ref in self.calls:
            out += [(dt, g, ref)] * count
        return out


@pytest.fixture()
def worker(pkg):
    from importlib import import_module
    layout = import_module(pkg.__name__ + ".layout")
    wmod = import_module(pkg.__name__ + ".worker")
    made = []

    def make(particles, config):
        made.append(RecordingSolver(layout, particles, config))
        return made[-1]

    w = wmod.MpmSimulationWorker(make_solver=make)
    w.made = made
    return w, wmod


def _cfl(pkg):
    from importlib import import_module
    return import_module(pkg.__name__ + ".cfl")


G = {"x": 0.0, "y": 90.0}


def test_mean_sinusoidal_acceleration_is_the_interval_mean(pkg):
    cfl = _cfl(pkg)
    w = 2 * math.pi * cfl.HORIZONTAL_OSCILLATION_FREQUENCY_HZ
    for start, dur, amp in ((0.0, 1 / 240, 30.0), (1.37, 0.25, -12.0), (5.0, 3.1, 7.0), (0.2, 1e-9, 4.0)):
        mean = integrate.quad(lambda t: amp * math.sin(w * t), start, start + dur, epsabs=0, epsrel=1e-13)[0] / dur
        assert cfl.mean_sinusoidal_acceleration(start, dur, amp) == pytest.approx(mean, rel=1e-9, abs=1e-12)
    assert cfl.mean_sinusoidal_acceleration(0.4, 0.0, 3.0) == pytest.approx(3.0 * math.sin(w * 0.4), rel=1e-15)
    assert cfl.mean_sinusoidal_acceleration(0.4, 0.1, 0.0) == 0.0
    assert cfl.mean_sinusoidal_acceleration(0.4, 0.1, 5.0, frequencyHz=0.0) == 0.0
    # a whole period averages to zero
    assert abs(cfl.mean_sinusoidal_acceleration(0.3, 1 / cfl.HORIZONTAL_OSCILLATION_FREQUENCY_HZ, 9.0)) < 1e-12
    for bad in ((math.nan, 0.1, 1.0), (0.0, -0.1, 1.0), (0.0, 0.1, math.inf)):
        with pytest.raises(ValueError):
            cfl.mean_sinusoidal_acceleration(*bad)


def test_initialize_runs_one_macrostep_and_answers_ready(worker, pkg):
    w, wmod = worker
    r = w.handle({"type": "initialize", "generation": 7, "liquidParticleDiameter": 20, "gravity": G})
    assert r["type"] == "ready" and r["generation"] == 7, r
    solver = w.made[-1]
    steps = solver.microsteps()
    assert len(steps) >= 1
    assert sum(dt for dt, _, _ in steps) == pytest.approx(1 / 240, rel=1e-12)
    assert all(g == G and ref == wmod.FIXED_STEP_SECONDS for _, g, ref in steps)
    assert len(solver.calls) == 1                         # one gravity, so one batch of steps
    assert r["remainderSeconds"] == 0.0
    snap = r["snapshot"]
    n = solver.state.particle_count
    assert snap["positions"].dtype == np.float64 and snap["positions"].shape == (2 * n,)
    assert snap["currentLiquidAreas"].dtype == np.float32 and snap["currentLiquidAreas"].shape == (n,)
    model = solver.state.packed[:, 13]
    liquid = (model == 2) | (model == 3)
    assert liquid.sum() == 3500
    np.testing.assert_array_equal(snap["currentLiquidAreas"][liquid], np.float32(400.0))
    np.testing.assert_array_equal(snap["currentLiquidAreas"][~liquid], 0.0)
    assert snap["brokenFractureBonds"].dtype == np.uint8 and snap["brokenFractureBonds"].shape == (53,)
    f = solver.state.packed.view(np.float32)
    np.testing.assert_array_equal(snap["positions"][0::2], f[:, 0].astype(np.float64))
    assert w.simulated_seconds == pytest.approx(1 / 240)


def test_liquid_particle_diameter_sets_the_liquid_area(worker):
    w, _ = worker
    r = w.handle({"type": "initialize", "generation": 1, "liquidParticleDiameter": 13, "gravity": G})
    areas = r["snapshot"]["currentLiquidAreas"]
    assert np.all((areas == 169.0) | (areas == 0.0)) and np.count_nonzero(areas) == 3500
    r = w.handle({"type": "initialize", "generation": 2, "liquidParticleDiameter": 9.5, "gravity": G})
    assert r == {"type": "error", "generation": 2, "message": "liquidParticleDiameter must be between 10 and 21."}
    assert w.made[0].disposed                             # a new initialize disposes the old simulation


def test_advance_plans_the_backlog_and_oscillates_per_microstep(worker, pkg):
    w, wmod = worker
    cfl = _cfl(pkg)
    w.handle({"type": "initialize", "generation": 1, "liquidParticleDiameter": 20, "gravity": G})
    solver = w.made[-1]
    solver.calls.clear()
    t0 = w.simulated_seconds                               # the clock after the initial macrostep
    assert t0 == pytest.approx(1 / 240, rel=1e-15)
    bounds = cfl.derive_cfl_bounds_packed(solver.state.packed, solver.state.particle_mass_f32)
    amp = 25.0
    expect, _ = cfl.calculate_cfl_step_batch(
        bounds.velocityBound, 90.0, amp, 0.0, 0.0, 0.0, accumulatorSeconds=0.1, elapsedSeconds=0.0,
        simulationSpeed=1.0, fixedStepSeconds=1 / 240, cellSize=10.0, waveSpeedBound=bounds.waveSpeedBound,
        maxMicrosteps=4096, courantLimit=cfl.DEFAULT_MPM_COURANT_LIMIT, allowPartialBatch=True)
    r = w.handle({"type": "advance", "generation": 2, "accumulatorSeconds": 0.1, "maxMicrosteps": 4096,
                  "gravity": G, "oscillationAmplitude": amp})
    assert r["type"] == "advanced", r
    assert expect.macroStepCount == 24
    assert r["simulatedSeconds"] == expect.simulatedSeconds
    assert r["remainderSeconds"] == expect.remainderSeconds
    assert r["microStepCount"] == expect.microStepCount
    assert r["workerMilliseconds"] >= 0
    steps = solver.microsteps()
    assert len(steps) == expect.microStepCount
    micro = expect.microstepPlan.microDt
    k = 0
    for macro in range(expect.macroStepCount):
        for m in range(expect.microstepPlan.stepCount):
            dt, g, ref = steps[k]
            osc = cfl.mean_sinusoidal_acceleration(t0 + (macro * expect.macroDt + m * micro), micro, amp)
            assert dt == micro and ref == 1 / 240
            assert g == {"x": osc, "y": 90.0}
            k += 1
    assert w.simulated_seconds == t0 + expect.simulatedSeconds


def test_advance_respects_the_microstep_budget(worker, pkg):
    w, _ = worker
    w.handle({"type": "initialize", "generation": 1, "liquidParticleDiameter": 20, "gravity": G})
    solver = w.made[-1]
    per_macro = len(solver.microsteps())
    solver.calls.clear()
    r = w.handle({"type": "advance", "generation": 3, "accumulatorSeconds": 1.0, "maxMicrosteps": 3 * per_macro,
                  "gravity": G, "oscillationAmplitude": 0.0})
    assert r["type"] == "advanced"
    assert r["microStepCount"] <= 3 * per_macro
    assert r["simulatedSeconds"] + r["remainderSeconds"] == pytest.approx(1.0, rel=1e-12)
    assert len(solver.microsteps()) == r["microStepCount"]
    assert len(solver.calls) == 1                          # no oscillation: one batch of steps


def test_step_is_one_macrostep(worker):
    w, _ = worker
    w.handle({"type": "initialize", "generation": 1, "liquidParticleDiameter": 20, "gravity": G})
    r = w.handle({"type": "step", "generation": 4, "maxMicrosteps": 256, "gravity": G, "oscillationAmplitude": 3.0})
    assert r["type"] == "stepped" and r["generation"] == 4
    assert r["simulatedSeconds"] == pytest.approx(1 / 240) and r["remainderSeconds"] == 0.0
    r = w.handle({"type": "step", "generation": 5, "maxMicrosteps": 0, "gravity": G, "oscillationAmplitude": 0.0})
    assert r["type"] == "error" and r["generation"] == 5


def test_errors_are_answered_with_the_generation(worker):
    w, _ = worker
    r = w.handle({"type": "advance", "generation": 9, "accumulatorSeconds": 0.1, "maxMicrosteps": 10,
                  "gravity": G, "oscillationAmplitude": 0.0})
    assert r == {"type": "error", "generation": 9, "message": "MLS-MPM simulation is not initialized."}
    r = w.handle({"type": "resize", "generation": 10})
    assert r["type"] == "error" and r["generation"] == 10
    r = w.handle("advance")
    assert r["type"] == "error" and r["generation"] is None
