"""The simulation worker's message protocol (reference/cpu-simulation.worker.ts:
32-193, messages of reference/cpu-simulation.ts:22-63) with the MI355X solver
as its simulation.

The reference runs its CPU step in a web worker that answers three requests:

* ``initialize`` (generation, liquidParticleDiameter, gravity): a fresh Gate 4
  scene on the code grid, advanced by exactly one 1/240 s macrostep;
  answered ``ready`` with a render snapshot and the backlog remainder;
* ``advance`` (generation, accumulatorSeconds, maxMicrosteps, gravity,
  oscillationAmplitude): as many whole macrosteps of the backlog as fit the
  microstep budget (calculateMpmCflStepBatch with partial batches), answered
  ``advanced`` with the snapshot, simulated seconds, remainder, microstep
  count and the worker's milliseconds;
* ``step`` (the same fields without the backlog): exactly one macrostep,
  answered ``stepped``;

and ``error`` (generation, message) when a request fails.  Every batch plans
its CFL microsteps from the current particles' bounds, and each microstep
adds the mean of the horizontal oscillation over its interval to gravity.

Here the same protocol drives MpmMlsGpuSolver, so a host that talks to the
reference's worker can talk to the GPU instead, one request at a time.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import cfl as cflmod
from . import scenes
from .layout import MODEL_REMOVED, RangeError

FIXED_STEP_SECONDS = 1.0 / 240.0          # GATE4_FIXED_STEP_SECONDS, scenes/gate4.ts:11
MAX_CFL_MICROSTEPS_PER_MANUAL_STEP = 256  # MPM_CPU_MAX_CFL_MICROSTEPS_PER_MANUAL_STEP, cpu-simulation.ts:20
LIQUID_PARTICLE_DIAMETER_MIN, LIQUID_PARTICLE_DIAMETER_MAX = 10.0, 21.0  # scenes/gate4.ts:28-29


def gate4_liquid_area(diameter: float) -> float:
    """getGate4LiquidParticleArea (scenes/gate4.ts:169-182)."""
    if not isinstance(diameter, (int, float)) or not math.isfinite(diameter) or \
            diameter < LIQUID_PARTICLE_DIAMETER_MIN or diameter > LIQUID_PARTICLE_DIAMETER_MAX:
        raise RangeError(f"liquidParticleDiameter must be between {LIQUID_PARTICLE_DIAMETER_MIN:g} and "
                         f"{LIQUID_PARTICLE_DIAMETER_MAX:g}.")
    return float(diameter) ** 2


class MpmSimulationWorker:
    """One simulation and its clock; ``handle`` answers one request."""

    def __init__(self, make_solver=None, device: int = 0):
        if make_solver is None:
            from .solver import create_mpm_mls_gpu_solver

            def make_solver(particles, config):
                return create_mpm_mls_gpu_solver(particles, config, device=device)
        self._make_solver = make_solver
        self.solver = None
        self.config = None
        self.simulated_seconds = 0.0

    # ----------------------------------------------------------- protocol
    def handle(self, request: dict) -> dict:
        generation = request.get("generation") if isinstance(request, dict) else None
        try:
            if not isinstance(request, dict):
                raise TypeError("worker request must be an object")
            kind = request.get("type")
            if kind == "initialize":
                return self._initialize(request)
            if kind not in ("advance", "step"):
                raise RangeError(f"unknown worker request type {kind!r}")
            started = time.perf_counter()
            if kind == "advance":
                planned = self._calculate_batch(request["accumulatorSeconds"], 0.0, 1.0, request["maxMicrosteps"],
                                                request["gravity"], request["oscillationAmplitude"])
            else:
                planned = self._calculate_batch(0.0, FIXED_STEP_SECONDS, 1.0, request["maxMicrosteps"],
                                                request["gravity"], request["oscillationAmplitude"])
                if planned.macroStepCount != 1:
                    raise RangeError("A CFL-safe Gate 4 step could not be planned.")
            self._advance_batch(planned, request["gravity"], request["oscillationAmplitude"])
            worker_ms = (time.perf_counter() - started) * 1e3
            return {"type": "advanced" if kind == "advance" else "stepped", "generation": generation,
                    "snapshot": self.render_snapshot(), "simulatedSeconds": planned.simulatedSeconds,
                    "remainderSeconds": planned.remainderSeconds, "microStepCount": planned.microStepCount,
                    "workerMilliseconds": worker_ms}
        except Exception as error:  # noqa: BLE001 -- every failure is answered, as the reference's worker does
            text = str(error).strip() or "MLS-MPM worker failed with an unknown error."
            return {"type": "error", "generation": generation, "message": text}

    def _initialize(self, request: dict) -> dict:
        if self.solver is not None:
            self.solver.dispose()
            self.solver = None
        self.simulated_seconds = 0.0
        area = gate4_liquid_area(request.get("liquidParticleDiameter", 20.0))
        particles, config, _ = scenes.gate4_scene(mpm_grid=False, liquid_area=area)
        self.config = config
        self.solver = self._make_solver(particles, config)
        initial = self._calculate_batch(0.0, FIXED_STEP_SECONDS, 1.0, MAX_CFL_MICROSTEPS_PER_MANUAL_STEP,
                                        request["gravity"], 0.0)
        if initial.macroStepCount != 1 or initial.macroDt != FIXED_STEP_SECONDS:
            raise RangeError("The initial Gate 4 macrostep could not fit the CFL batch.")
        self._advance_batch(initial, request["gravity"], 0.0)
        return {"type": "ready", "generation": request.get("generation"), "snapshot": self.render_snapshot(),
                "remainderSeconds": initial.remainderSeconds}

    # ----------------------------------------------------------- batches
    def _require(self):
        if self.solver is None:
            raise RuntimeError("MLS-MPM simulation is not initialized.")
        return self.solver

    def _calculate_batch(self, accumulator, elapsed, speed, max_microsteps, gravity, amplitude):
        """calculateBatch (cpu-simulation.worker.ts:32-69): the CFL batch of
        the backlog from the current particles' bounds."""
        s = self._require()
        records, _ = s.ctx.read_particles()
        live = records[records[:, 13] != MODEL_REMOVED]
        bounds = cflmod.derive_cfl_bounds_packed(live, s.state.particle_mass_f32)
        if bounds is None:
            raise RangeError("The Gate 4 scene has no supported CFL wave-speed bound.")
        batch, _ = cflmod.calculate_cfl_step_batch(
            bounds.velocityBound, math.hypot(gravity["x"], gravity["y"]), amplitude, 0.0, 0.0, 0.0,
            accumulatorSeconds=accumulator, elapsedSeconds=elapsed, simulationSpeed=speed,
            fixedStepSeconds=FIXED_STEP_SECONDS, cellSize=self.config["grid"]["cellSize"],
            waveSpeedBound=bounds.waveSpeedBound, maxMicrosteps=max_microsteps,
            courantLimit=cflmod.DEFAULT_MPM_COURANT_LIMIT, allowPartialBatch=True)
        return batch

    def microstep_gravities(self, batch, gravity, amplitude, start_seconds=None):
        """The gravity of every microstep of a batch: the horizontal
        oscillation's mean over the microstep added to gravity.x
        (advanceBatch, cpu-simulation.worker.ts:71-101)."""
        start = self.simulated_seconds if start_seconds is None else start_seconds
        micro = batch.microstepPlan.microDt
        out = []
        for macro in range(batch.macroStepCount):
            for k in range(batch.microstepPlan.stepCount):
                offset = macro * batch.macroDt + k * micro
                osc = cflmod.mean_sinusoidal_acceleration(start + offset, micro, amplitude)
                out.append({"x": gravity["x"] + osc, "y": gravity["y"]})
        return out

    def _advance_batch(self, batch, gravity, amplitude):
        s = self._require()
        micro = batch.microstepPlan.microDt
        # Runs of microsteps with the same gravity (all of them when the
        # oscillation is off) go to the GPU as one batch of steps.
        run, count = None, 0
        for g in self.microstep_gravities(batch, gravity, amplitude):
            if g != run and count:
                s.step_batch(count, micro, run, bond_reference_dt=FIXED_STEP_SECONDS)
                count = 0
            run, count = g, count + 1
        if count:
            s.step_batch(count, micro, run, bond_reference_dt=FIXED_STEP_SECONDS)
        s.check_numerical_status()
        self.simulated_seconds += batch.simulatedSeconds

    # ----------------------------------------------------------- snapshot
    def render_snapshot(self) -> dict:
        """createRenderSnapshot (cpu-simulation.ts:144-164): positions (x, y
        per particle, float64), the current area V det F of every liquid
        particle (0 for others, float32) and each fracture bond's broken flag."""
        s = self._require()
        records, _ = s.ctx.read_particles()
        f = records.view(np.float32)
        positions = f[:, 0:2].astype(np.float64).ravel()
        liquid = (records[:, 13] == 2) | (records[:, 13] == 3)
        det = f[:, 8].astype(np.float64) * f[:, 11] - f[:, 9].astype(np.float64) * f[:, 10]
        volume = f[:, 17].astype(np.float64)
        areas = np.where(liquid, volume * det, 0.0).astype(np.float32)
        with np.errstate(invalid="ignore"):
            bad = liquid & ~((volume > 0) & (det > 0) & np.isfinite(areas) & (areas > 0))
        if np.any(bad):
            raise RangeError(f"Liquid particle {int(np.argmax(bad))} has an invalid current area.")
        bonds = s.read_fracture_bonds() if s.state.fracture_bonds is not None else []
        broken = np.array([1 if b["broken"] else 0 for b in bonds], np.uint8)
        return {"positions": positions, "currentLiquidAreas": areas, "brokenFractureBonds": broken}

    def dispose(self):
        if self.solver is not None:
            self.solver.dispose()
            self.solver = None
