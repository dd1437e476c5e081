"""``MpmMlsGpuSolver`` over the native MI355X step (libmpm355.so).

Mirrors the reference's solver host API (src/lib/mpm/kernels/solver.ts:106-144)
for the step/batch path: ``step``, ``step_batch``, ``step_cfl_batch``,
``step_cfl_batch_adaptive``, ``read_particles``,
``measure_maximum_particle_speed``, ``check_numerical_status``,
``append_particles``, drag interaction and ``dispose``.  Promises of the TS API
are synchronous calls here; validation raises ``RangeError``/``TypeError``
before any device work, numerical failures surface as
:class:`MpmMlsGpuNumericalError` after a whole-step rollback.
"""
from __future__ import annotations

import math
import struct
import time
from dataclasses import dataclass, replace

import numpy as np

from . import cfl as cflmod
from . import layout
from .layout import RangeError
from .native import (MPM_ECAPACITY, MPM_ESTATE, MPM_FLAG_KEEP_FIXED_GRID, MPM_P2G_FIXED_POINT, MPM_P2G_GATHER,
                     Context, MpmConfig, NativeError)

ERROR_DEFORMATION = 2
ERROR_DEFORMATION_SUBSTEP_LIMIT = 32
ERROR_CFL_BOUND = 16384
CFL_RECOVERY_BOUND_SAFETY_FACTOR = 1.25   # solver.ts:73
MAX_U32 = 0xFFFFFFFF

_FLAG_MESSAGES = [  # solver.ts:161-178
    (1, "invalid particle stencil or linked-list state"),
    (2, "non-positive trial deformation determinant"),
    (4, "P2G validation failed"),
    (8, "non-finite grid state"),
    (16, "non-finite G2P particle state"),
    (32, "deformation update exceeded the adaptive substep limit"),
    (64, "non-positive projected deformation determinant"),
    (128, "invalid P2G material parameters"),
    (256, "non-finite particle stress or momentum"),
    (1024, "non-finite P2G node accumulation"),
    (2048, "non-finite P2G node output"),
    (4096, "invalid triangle contact field or reaction"),
    (8192, "invalid bonded triangle state"),
    (16384, "particle exceeded the enforced CFL wave-speed or velocity bound"),
    (32768, "invalid persistent solid-fracture bond state"),
    (65536, "experimental fixed-point P2G contribution or accumulator overflow"),
]


def _u2f(u: int) -> float:
    return struct.unpack("<f", struct.pack("<I", int(u) & 0xFFFFFFFF))[0]


class MpmMlsGpuNumericalError(RuntimeError):
    """MpmMlsGpuNumericalError (solver.ts:146-206)."""

    def __init__(self, flags, particle_index=None, diagnostics=None, projected_deformation=None, cfl_details=None,
                 completed_microsteps=None, measured_maximum_speed=None):
        msg = "; ".join(d for f, d in _FLAG_MESSAGES if flags & f) or f"status 0x{flags:x}"
        ctx = "" if particle_index is None else f" at particle {particle_index}"
        super().__init__(f"GPU MLS-MPM rejected a step ({msg}{ctx}); all particles were rolled back to the last "
                         "valid state.")
        self.flags = flags
        self.particle_index = particle_index
        self.diagnostics = diagnostics
        self.projected_deformation = projected_deformation
        self.cfl_details = cfl_details
        self.completed_microsteps = completed_microsteps
        self.measured_maximum_speed = measured_maximum_speed


def decode_numerical_error(status, completed=None, max_speed=None):
    """decodeNumericalError (solver.ts:232-284)."""
    status = [int(s) for s in status]
    if status[0] == 0:
        return None
    idx = None if status[1] == MAX_U32 else status[1]
    vals = [_u2f(status[k]) for k in range(2, 6)]
    diag = None if status[6] != 1 else {"determinant": vals[0], "deformationIncrement": vals[1],
                                         "precedingDeterminant": vals[2]}
    proj = None if status[6] != 2 else {"m00": vals[0], "m01": vals[1], "m10": vals[2], "m11": vals[3]}
    cfl = None if status[6] != 3 else {"particleSpeed": vals[0], "waveSpeed": vals[1], "velocityLimit": vals[2],
                                        "waveSpeedLimit": vals[3]}
    return MpmMlsGpuNumericalError(status[0], idx, diag, proj, cfl, completed, max_speed)


def is_recoverable_cfl_error(error: MpmMlsGpuNumericalError) -> bool:
    """isRecoverableCflError (solver.ts:296-312)."""
    if error.flags == ERROR_CFL_BOUND:
        d = error.cfl_details
        return d is not None and all(math.isfinite(d[k]) for k in d)
    return error.flags in (ERROR_DEFORMATION, ERROR_DEFORMATION_SUBSTEP_LIMIT) and \
        error.measured_maximum_speed is not None and math.isfinite(error.measured_maximum_speed)


@dataclass(frozen=True)
class SpeedMeasurement:
    maximumSpeed: float
    sequence: int


@dataclass(frozen=True)
class CflBatchMeasurement:
    maximumSpeed: float
    sequence: int
    recommendedVelocityBound: float
    attemptedMicroStepCount: int
    retryCount: int
    retrySubmissionMilliseconds: float
    readbackMilliseconds: float
    simulatedDuration: float
    remainingDuration: float


MPM_MLS_GPU_DRAG_STIFFNESS = 20.0   # solver.ts:379-380
MPM_MLS_GPU_DRAG_DAMPING = 8.0
MAX_U32_OBJECT_ID = 0xFFFFFFFF


def _finite_interaction_float32(value, name: str) -> float:
    """finiteInteractionFloat32 (solver.ts:3271-3280)."""
    if isinstance(value, bool) or not isinstance(value, (int, float, np.floating, np.integer)) or \
            not math.isfinite(float(value)):
        raise TypeError(f"{name} must be a finite number.")
    with np.errstate(over="ignore"):
        result = float(np.float32(value))
    if not math.isfinite(result):
        raise RangeError(f"{name} is outside the supported float32 range.")
    return result


def validate_world_point(point, state, name: str) -> dict:
    """validateMpmWorldPoint (solver.ts:3228-3242): a float32 point inside the domain."""
    if not isinstance(point, dict):
        raise TypeError(f"{name} must be a vector.")
    x = _finite_interaction_float32(point.get("x"), f"{name}.x")
    y = _finite_interaction_float32(point.get("y"), f"{name}.y")
    if x < 0 or x > state.domain_width_f32 or y < 0 or y > state.domain_height_f32:
        raise RangeError(f"{name} must be inside the MLS-MPM domain.")
    return {"x": x, "y": y}


def validate_pick_radius(radius, state) -> float:
    """validateMpmPickRadius (solver.ts:3244-3250)."""
    value = _finite_interaction_float32(radius, "pick radius")
    if value < 0 or value > max(state.domain_width_f32, state.domain_height_f32):
        raise RangeError("pick radius must be non-negative and within the MLS-MPM domain extent.")
    return value


def validate_drag_radius(radius, state) -> float:
    """validateMpmDragRadius (solver.ts:3252-3256)."""
    value = validate_pick_radius(radius, state)
    if value <= 0:
        raise RangeError("drag radius must be greater than zero.")
    return value


def validate_clone_offset(offset) -> dict:
    """validateMpmCloneOffset (solver.ts:3258-3266)."""
    if not isinstance(offset, dict):
        raise TypeError("clone offset must be a vector.")
    return {"x": _finite_interaction_float32(offset.get("x"), "clone offset.x"),
            "y": _finite_interaction_float32(offset.get("y"), "clone offset.y")}


class MpmMlsGpuSolver:
    """Solver host over one native context (one HIP stream)."""

    def __init__(self, particles, config: dict, cfl_bounds: cflmod.CflBounds | None = None,
                 p2g_mode: int = MPM_P2G_GATHER, keep_fixed_grid: bool = False, device: int = -1,
                 bucket_slots: int = 0, tuning: dict | str | None = None):
        """tuning: None runs the default launch shape; a dict sets one
        (autotune.DEFAULT_TUNING's keys); "auto" takes the cached tuning of
        this scene's class, tuning and caching it first when there is none
        (autotune.tuning_for).  particles may also be a ValidatedInput (a
        restored snapshot, layout.validated_from_records), which is taken as
        it is."""
        if isinstance(particles, layout.ValidatedInput):
            if isinstance(tuning, str):
                raise RangeError("a validated particle set takes an explicit tuning or none")
            self.state = particles
        else:
            self.state = layout.validate_input(particles, config)
        derived = cfl_bounds
        if derived is None:
            derived = cflmod.derive_cfl_bounds_packed(self.state.packed, self.state.particle_mass_f32)
        self._validate_cfl_bounds(derived)
        self.cfl_bounds = derived
        cfg = MpmConfig(self.state.config["grid"]["cellsX"], self.state.config["grid"]["cellsY"],
                        self.state.cell_size_f32, self.state.particle_mass_f32, self.state.particle_volume_f32,
                        self.state.wall_friction_f32, self.state.particle_capacity, p2g_mode,
                        MPM_FLAG_KEEP_FIXED_GRID if keep_fixed_grid else 0, device, bucket_slots)
        self.ctx = Context(cfg, self.state.packed, self.state.object_ids)
        if self.state.fracture_bonds is not None:
            self.ctx.set_fracture_bonds(self.state.fracture_bonds)
        self.p2g_mode = p2g_mode
        if tuning is not None:
            from . import autotune
            if isinstance(tuning, str):
                if tuning != "auto":
                    raise RangeError('tuning must be None, a dict or "auto"')
                tuning = autotune.tuning_for(particles, config, p2g_mode, device=max(device, 0))
            autotune.apply(self, tuning)
        self._init_host_state()

    def _init_host_state(self):
        """Host bookkeeping of createGpuSolver (solver.ts:1242-1270): active
        count, object ranges and the next object id, removal and drag state."""
        self._disposed = False
        self._cfl_state_usable = True
        self._speed_measurement_submitted = False
        self._last_scheduled_velocity_bound = None
        self._latest_sequence = 0
        self._latest_microsteps = 0
        self._drag = None
        packed = self.state.packed
        self._active_count = int(np.count_nonzero(packed[:, 13] != layout.MODEL_REMOVED)) if len(packed) else 0
        self._has_removed = self._active_count != len(packed)
        self._object_ranges = {r["objectId"]: dict(r) for r in self.state.object_ranges}
        self._next_object_id = max((r["objectId"] + 1 for r in self.state.object_ranges), default=0)
        oid = self.state.object_ids
        self._body_count = self.state.bonded_body_count
        self._object_bodies = {}
        first_vertex = packed[:, 19] == 1 if len(packed) else np.zeros(0, bool)
        for r in self.state.object_ranges:
            k = int(np.count_nonzero(first_vertex[r["start"]:r["start"] + r["count"]]))
            if k:
                self._object_bodies[r["objectId"]] = k
        self._bonded_objects = set()
        if self.state.fracture_bonds is not None:
            for b in self.state.fracture_bonds:
                self._bonded_objects.add(int(oid[int(b[0])]))

    @classmethod
    def _over_context(cls, state, cfl_bounds, ctx, p2g_mode):
        """A solver over an existing native context (a scene of a batch)."""
        self = cls.__new__(cls)
        self.state = state
        self.cfl_bounds = cfl_bounds
        self.ctx = ctx
        self.p2g_mode = p2g_mode
        self._init_host_state()
        return self

    # ----------------------------------------------------------- properties
    @property
    def particle_count(self) -> int:
        """Active (non-removed) particles (activeParticleCount, solver.ts:1242)."""
        return self._active_count

    @property
    def allocated_particle_count(self) -> int:
        return self.ctx.counts()[0]

    @property
    def particle_capacity(self) -> int:
        return self.ctx.counts()[1]

    def _assert_active(self):
        if self._disposed:
            raise RuntimeError("MLS-MPM GPU solver has been disposed.")

    @staticmethod
    def _validate_cfl_bounds(bounds):
        """validateCflBounds (solver.ts:471-500)."""
        if bounds is None:
            return
        for name in ("waveSpeedBound", "velocityBound"):
            v = getattr(bounds, name)
            if not math.isfinite(v) or v < 0:
                raise RangeError(f"cflBounds.{name} must be finite and non-negative.")
        if not math.isfinite(bounds.courantLimit) or bounds.courantLimit <= 0:
            raise RangeError("cflBounds.courantLimit must be positive and finite.")

    # -------------------------------------------------------------- stepping
    def step(self, dt: float, gravity: dict):
        self.step_batch(1, dt, gravity)

    def step_batch(self, step_count: int, dt: float, gravity: dict, cfl_velocity_bound: float | None = None,
                   bond_reference_dt: float | None = None, measure_speed: bool = False):
        """stepBatch (solver.ts:1407-1599)."""
        self._assert_active()
        if not isinstance(step_count, int) or isinstance(step_count, bool) or step_count < 1:
            raise RangeError("stepCount must be a positive safe integer.")
        guard = None
        if cfl_velocity_bound is not None:
            if self.cfl_bounds is None:
                raise RangeError("this solver configuration has no safely derived CFL bounds.")
            if not math.isfinite(cfl_velocity_bound) or cfl_velocity_bound < 0:
                raise RangeError("velocityBound must be finite and non-negative.")
            guard = {"waveSpeedBound": self.cfl_bounds.waveSpeedBound, "velocityBound": cfl_velocity_bound}
        params = layout.pack_step_parameters(self.state, dt, gravity, self.allocated_particle_count,
                                             self._body_count, guard, bond_reference_dt)
        try:
            self.ctx.step_batch(params, step_count, measure_speed or guard is not None)
        except Exception as cause:
            self._cfl_state_usable = False
            self._speed_measurement_submitted = False
            raise RuntimeError(f"Failed to encode or submit an MLS-MPM batch of {step_count} step(s).") from cause
        self._speed_measurement_submitted = guard is not None
        if guard is None:
            self._cfl_state_usable = False

    def step_cfl_batch(self, macro_step_count: int, macro_dt: float, microstep_plan: cflmod.MicrostepPlan,
                       velocity_bound: float, max_microsteps: int, gravity: dict,
                       bond_reference_dt: float | None = None):
        """stepCflBatch (solver.ts:1601-1722): certify the plan, then one batch."""
        self._assert_active()
        if self.cfl_bounds is None:
            raise RangeError("CFL stepping requires a material-derived wave-speed bound; this solver configuration "
                             "has no safely derived CFL bounds.")
        if not self._cfl_state_usable:
            raise RuntimeError("a direct non-CFL GPU step or a failed CFL measurement invalidated the tracked "
                               "velocity bound; recreate the solver before CFL stepping.")
        if self._speed_measurement_submitted:
            raise RuntimeError("queue measureMaximumParticleSpeed() after the previous CFL batch before submitting "
                               "another one.")
        if self._last_scheduled_velocity_bound is not None and velocity_bound < self._last_scheduled_velocity_bound:
            raise RangeError("velocityBound must cover the last submitted GPU maximum-particle envelope.")
        if not isinstance(macro_step_count, int) or macro_step_count < 1:
            raise RangeError("macroStepCount must be a positive safe integer.")
        if not isinstance(max_microsteps, int) or max_microsteps < 1:
            raise RangeError("maxMicrosteps must be a positive safe integer.")
        cell = self.state.config["grid"]["cellSize"]
        b = self.cfl_bounds
        expected = cflmod.calculate_microsteps(macro_dt, cell, b.waveSpeedBound, velocity_bound, max_microsteps,
                                               courantLimit=b.courantLimit)
        certified = None
        if isinstance(microstep_plan, cflmod.MicrostepPlan) and microstep_plan.stepCount >= expected.stepCount:
            certified = cflmod.calculate_microsteps(macro_dt, cell, b.waveSpeedBound, velocity_bound,
                                                    max_microsteps, microstep_plan.stepCount, b.courantLimit)
        if certified is None or certified != microstep_plan:
            raise RangeError("microstepPlan does not match this solver configuration or is not a certified CFL "
                             "refinement.")
        micro_count = macro_step_count * certified.stepCount
        if micro_count > max_microsteps:
            raise RangeError(f"requested microstep count ({micro_count}) exceeds maxMicrosteps ({max_microsteps}).")
        represented = certified.microDt * certified.stepCount
        tol = max(macro_dt, 5e-324) * 2.220446049250313e-16 * 4
        if abs(represented - macro_dt) > tol:
            raise RangeError("microstepPlan does not preserve the requested macro dt.")
        gdt = cflmod.fround(certified.microDt)
        if self._drag is not None and (not math.isfinite(gdt) or gdt <= 0 or gdt * MPM_MLS_GPU_DRAG_DAMPING > 1):
            raise RangeError("CFL microDt is too large for the explicit drag damping update to remain dissipative.")
        limit = cflmod.calculate_velocity_limit(cell, b.waveSpeedBound, certified.microDt, b.courantLimit)
        if limit < velocity_bound:
            raise RangeError("float32 microDt does not preserve the planned CFL velocity bound; increase the "
                             "microstep count.")
        ref = bond_reference_dt if bond_reference_dt is not None else (
            self.state.bond_reference_dt_f32 if self.state.bond_reference_dt_f32 is not None else macro_dt)
        self.step_batch(micro_count, certified.microDt, gravity, limit, ref)
        self._latest_sequence += 1
        self._latest_microsteps = micro_count
        self._last_scheduled_velocity_bound = velocity_bound

    def step_cfl_batch_adaptive(self, macro_step_count, macro_dt, microstep_plan, velocity_bound, max_microsteps,
                                gravity, gravity_for_duration=None) -> CflBatchMeasurement:
        """stepCflBatchAdaptive (solver.ts:1724-2021): retry the failed
        interval with at least two finer microsteps and raised bounds."""
        if self.cfl_bounds is None:
            raise RangeError("CFL stepping requires a material-derived wave-speed bound; this solver configuration "
                             "has no safely derived CFL bounds.")
        if not isinstance(macro_step_count, int) or macro_step_count < 1:
            raise RangeError("macroStepCount must be a positive safe integer.")
        if not math.isfinite(macro_dt) or macro_dt <= 0:
            raise RangeError("macroDt must be positive and finite.")
        total = macro_step_count * macro_dt
        if not math.isfinite(total):
            raise RangeError("CFL batch duration is outside the supported range.")
        bond_ref = self.state.bond_reference_dt_f32 if self.state.bond_reference_dt_f32 is not None else macro_dt
        cell = self.state.config["grid"]["cellSize"]
        remaining = total
        a_count, a_dt, a_plan, a_bound = macro_step_count, macro_dt, microstep_plan, velocity_bound
        consumed = 0
        retries = 0
        submissions = 0
        retry_ms = 0.0
        readback_ms = 0.0
        tol = max(total, 5e-324) * 2.220446049250313e-16 * 4

        def result(measurement: SpeedMeasurement, recommended=None):
            return CflBatchMeasurement(measurement.maximumSpeed, measurement.sequence,
                                       measurement.maximumSpeed if recommended is None else recommended, consumed,
                                       retries, retry_ms, readback_ms, max(0.0, total - remaining), remaining)

        while True:
            a_micro = a_count * a_plan.stepCount
            a_duration = a_count * a_dt
            budget = max_microsteps - consumed
            if a_micro < 1 or not math.isfinite(a_duration) or a_duration <= 0 or a_micro > budget:
                raise RangeError(f"adaptive CFL retry requires {a_micro} microsteps with {budget} remaining.")
            t0 = time.perf_counter()
            g = gravity if gravity_for_duration is None else gravity_for_duration(total - remaining, a_duration)
            self.step_cfl_batch(a_count, a_dt, a_plan, a_bound, budget, g, bond_ref)
            if submissions > 0:
                retry_ms += (time.perf_counter() - t0) * 1e3
            submissions += 1
            t1 = time.perf_counter()
            try:
                measurement = self.measure_maximum_particle_speed()
            except MpmMlsGpuNumericalError as error:
                readback_ms += (time.perf_counter() - t1) * 1e3
                if not is_recoverable_cfl_error(error):
                    raise
                done = error.completed_microsteps
                if done is None or done < 0 or done >= a_micro:
                    raise RuntimeError(f"Adaptive CFL recovery could not determine completed microsteps: {error}") \
                        from error
                done_duration = done * a_plan.microDt
                if not math.isfinite(done_duration) or done_duration >= remaining:
                    raise RuntimeError(f"Adaptive CFL recovery received invalid simulation progress: {error}") \
                        from error
                remaining -= done_duration
                consumed += a_micro
                observed = max(error.measured_maximum_speed or 0.0,
                               (error.cfl_details or {}).get("particleSpeed", 0.0) or 0.0)
                retry_bound = max(a_bound, self._last_scheduled_velocity_bound or 0.0,
                                  observed * CFL_RECOVERY_BOUND_SAFETY_FACTOR)
                if not math.isfinite(retry_bound):
                    raise RuntimeError(f"Adaptive CFL recovery stopped after {retries} retry attempt(s) (non-finite "
                                       f"velocity bound); last GPU error: {error}") from error
                self._last_scheduled_velocity_bound = retry_bound
                if error.cfl_details is not None:
                    retry_wave = error.cfl_details["waveSpeed"] * CFL_RECOVERY_BOUND_SAFETY_FACTOR
                    if not math.isfinite(retry_wave):
                        raise RuntimeError("Adaptive CFL recovery stopped (non-finite wave-speed bound)") from error
                    self.cfl_bounds = replace(self.cfl_bounds,
                                              waveSpeedBound=max(self.cfl_bounds.waveSpeedBound, retry_wave))
                retry_budget = max_microsteps - consumed
                if retry_budget < 2:
                    self.ctx.clear_status()
                    self._cfl_state_usable = True
                    return result(SpeedMeasurement(observed, self._latest_sequence), retry_bound)
                failed = min(remaining, a_plan.microDt)
                try:
                    a_plan = cflmod.calculate_microsteps(failed, cell, self.cfl_bounds.waveSpeedBound, retry_bound,
                                                         retry_budget, 2, self.cfl_bounds.courantLimit)
                except RangeError as cause:
                    if not cflmod.is_microstep_capacity_error(cause):
                        raise RuntimeError(f"Adaptive CFL recovery could not schedule a finer step ({cause}): "
                                           f"{error}") from cause
                    self.ctx.clear_status()
                    self._cfl_state_usable = True
                    return result(SpeedMeasurement(observed, self._latest_sequence), retry_bound)
                self.ctx.clear_status()
                self._cfl_state_usable = True
                a_count, a_dt, a_bound = 1, failed, retry_bound
                retries += 1
                continue
            readback_ms += (time.perf_counter() - t1) * 1e3
            consumed += a_micro
            remaining = max(0.0, remaining - a_duration)
            if remaining <= tol:
                remaining = 0.0
                return result(measurement, max(a_bound, measurement.maximumSpeed) if retries > 0
                              else measurement.maximumSpeed)
            cont_budget = max_microsteps - consumed
            if cont_budget < 2:
                return result(measurement, max(a_bound, measurement.maximumSpeed))
            a_bound = max(a_bound, measurement.maximumSpeed)
            safe = cflmod.calculate_safe_dt(cell, self.cfl_bounds.waveSpeedBound, a_bound,
                                            self.cfl_bounds.courantLimit)
            max_step = min(a_plan.microDt, safe)
            if not math.isfinite(max_step) or max_step <= 0:
                raise RangeError(f"Adaptive CFL recovery produced an invalid continuation timestep ({max_step}).")
            required = math.ceil(remaining / max_step)
            cont_duration = remaining
            min_count = max(1, required)
            if required >= cont_budget:
                reserved = cont_budget - 1
                cont_duration = min(remaining, max_step * reserved * (1 - 2.220446049250313e-16 * 8))
                min_count = max(1, math.ceil(cont_duration / max_step))
            if not math.isfinite(cont_duration) or cont_duration <= 0:
                return result(measurement, max(a_bound, measurement.maximumSpeed))
            a_plan = cflmod.calculate_microsteps(cont_duration, cell, self.cfl_bounds.waveSpeedBound, a_bound,
                                                 cont_budget, min_count, self.cfl_bounds.courantLimit)
            a_count, a_dt = 1, cont_duration

    # ------------------------------------------------------------- readback
    def measure_maximum_particle_speed(self) -> SpeedMeasurement:
        """measureMaximumParticleSpeed (solver.ts:2388-2512): one 40-byte
        compact snapshot; raises the decoded numerical error if any."""
        self._assert_active()
        was_submitted = self._speed_measurement_submitted
        self._speed_measurement_submitted = False
        speed, done, status = self.ctx.read_compact()
        err = decode_numerical_error(status, done, speed)
        if err is not None:
            self._cfl_state_usable = False
            raise err
        if not was_submitted:
            # no guarded batch queued a measurement: measure the current state
            recs, _ = self.ctx.read_particles()
            v = recs.view(np.float32)[:, 2:4]
            live = recs[:, 13] != layout.MODEL_REMOVED
            sp = np.sqrt(v[live, 0] * v[live, 0] + v[live, 1] * v[live, 1])
            speed = float(np.max(np.where(np.isfinite(sp), sp, np.float32(3.4028234663852886e38)))) if sp.size else 0.0
        return SpeedMeasurement(float(speed), self._latest_sequence)

    def check_numerical_status(self):
        """checkNumericalStatus (solver.ts:2514-2580)."""
        self._assert_active()
        err = decode_numerical_error(self.ctx.read_status())
        if err is not None:
            raise err

    def clear_numerical_status(self):
        self._assert_active()
        self.ctx.clear_status()

    def read_particles(self, particle_indices=None, packed: bool = False):
        """readParticles (solver.ts:2594-2731)."""
        self._assert_active()
        if particle_indices is not None:
            idx = list(particle_indices)
            n = self.allocated_particle_count
            for i in idx:
                if not isinstance(i, (int, np.integer)) or i < 0 or i >= n:
                    raise RangeError(f"particle index {i} is outside the active particle range.")
        recs, ids = self.ctx.read_particles(particle_indices)
        err = decode_numerical_error(self.ctx.read_status())
        if err is not None:
            raise err
        if packed:
            return recs, ids
        return layout.unpack_particles(recs, ids)

    def append_particles(self, particles) -> dict:
        """appendParticles (solver.ts:2108-2278): validated, capacity-bounded."""
        self._assert_active()
        requested = len(particles)
        if requested == 0:
            return {"requested": 0, "accepted": 0, "refused": 0, "dropped": 0}
        cfg = dict(self.state.config)
        cfg["capacity"] = max(requested, 1)
        cfg.pop("fractureBonds", None)
        if self._next_object_id > MAX_U32_OBJECT_ID:
            return {"requested": requested, "accepted": 0, "refused": requested, "dropped": 0,
                    "reason": "no object ids remain for a new emitter group"}
        # every emitted group is one new object (solver.ts:2154-2165)
        object_id = self._next_object_id
        if isinstance(particles, layout.ParticleArrays):
            particles = replace(particles, object_id=np.full(len(particles), object_id, np.uint32))
        else:
            particles = [dict(p, objectId=object_id) for p in particles]
        add = layout.validate_input(particles, cfg)
        free = self.particle_capacity - self.allocated_particle_count
        accepted = min(free, requested)
        if accepted > 0:
            start = self.allocated_particle_count
            self.ctx.append(add.packed[:accepted], add.object_ids[:accepted])
            self._object_ranges[object_id] = {"objectId": object_id, "start": start, "count": accepted}
            self._active_count += accepted
            self._next_object_id += 1
            if self.cfl_bounds is not None:
                extra = cflmod.derive_cfl_bounds_packed(add.packed[:accepted], self.state.particle_mass_f32)
                if extra is None:
                    self.cfl_bounds = None
                else:
                    self.cfl_bounds = replace(self.cfl_bounds,
                                              waveSpeedBound=max(self.cfl_bounds.waveSpeedBound, extra.waveSpeedBound),
                                              velocityBound=max(self.cfl_bounds.velocityBound, extra.velocityBound))
        return {"requested": requested, "accepted": accepted, "refused": requested - accepted, "dropped": 0,
                "activeParticleCount": self.allocated_particle_count, "capacity": self.particle_capacity,
                **({} if accepted == requested else {"reason": "capacity"})}

    # ------------------------------------------------------------ interaction
    def _resolve_selection(self, selection) -> dict:
        """resolveSelection (solver.ts:1363-1387): the object range of a
        selection {particleIndex, objectId, anchor}."""
        if not isinstance(selection, dict):
            raise TypeError("selection must identify an MPM particle.")
        idx, oid = selection.get("particleIndex"), selection.get("objectId")
        if not isinstance(idx, (int, np.integer)) or isinstance(idx, bool) or idx < 0 or \
                idx >= self.allocated_particle_count or not isinstance(oid, (int, np.integer)) or \
                isinstance(oid, bool) or oid < 0 or oid > MAX_U32_OBJECT_ID:
            raise RangeError("selection must identify an active MPM particle.")
        validate_world_point(selection.get("anchor"), self.state, "selection anchor")
        r = self._object_ranges.get(int(oid))
        if r is None or idx < r["start"] or idx >= r["start"] + r["count"]:
            raise RangeError("selection objectId does not match the selected particle.")
        return r

    def _write_interaction(self, target: dict, radius: float):
        """writeInteractionParameters (solver.ts:1389-1401)."""
        self.ctx.set_interaction([target["x"], target["y"], radius, MPM_MLS_GPU_DRAG_STIFFNESS,
                                  MPM_MLS_GPU_DRAG_DAMPING, 0.0, 0.0, 0.0])

    def begin_drag(self, selection: dict, target: dict, radius: float):
        """beginDrag (solver.ts:2825-2836): a spring-damper on the grid
        nodes within `radius` of the target (finishGridVelocity)."""
        self._assert_active()
        self._resolve_selection(selection)
        t = validate_world_point(target, self.state, "drag target")
        r = validate_drag_radius(radius, self.state)
        self._write_interaction(t, r)
        self._drag = r

    def update_drag(self, target: dict):
        """updateDrag (solver.ts:2838-2843)."""
        self._assert_active()
        if self._drag is None:
            return
        self._write_interaction(validate_world_point(target, self.state, "drag target"), self._drag)

    def end_drag(self):
        """endDrag (solver.ts:2845-2850)."""
        if self._drag is None:
            return
        self._drag = None
        if self._disposed:
            return
        self._write_interaction({"x": 0.0, "y": 0.0}, 0.0)

    def gather_surface_occupancy(self) -> np.ndarray:
        """The surface field the reference's renderer gathers from its bins
        (gatherSurfaceOccupancy, shaders.ts:710-806): per grid node (gridIndex
        order) the water and oil display area 2 V det F, weighted by the
        quadratic B-spline and divided by the cell area.  (node_count, 2)."""
        self._assert_active()
        return self.ctx.gather_occupancy()

    def pick(self, point: dict, radius: float):
        """pick (solver.ts:2733-2823): the active particle nearest to the
        point within the radius, as a selection {particleIndex, objectId,
        anchor}, or None."""
        self._assert_active()
        q = validate_world_point(point, self.state, "pick point")
        r = validate_pick_radius(radius, self.state)
        if self._active_count == 0:
            return None
        out = self.ctx.pick(q["x"], q["y"], r)
        index = int(out[0])
        if index == MAX_U32:
            return None
        anchor = {"x": float(out[1:2].view(np.float32)[0]), "y": float(out[2:3].view(np.float32)[0])}
        oid = int(out[3])
        rng = self._object_ranges.get(oid)
        if rng is None or index < rng["start"] or index >= rng["start"] + rng["count"]:
            raise RuntimeError("GPU returned an object id inconsistent with the selected particle.")
        return {"particleIndex": index, "objectId": oid,
                "anchor": validate_world_point(anchor, self.state, "GPU selection anchor")}

    def remove_particles_in_outlet(self, bounds: dict) -> int:
        """removeParticlesInOutlet (solver.ts:2280-2386): water and oil
        particles inside the box become removed; returns how many."""
        self._assert_active()
        if not isinstance(bounds, dict):
            raise TypeError("outlet bounds must be an object.")
        lo = validate_world_point({"x": bounds.get("minX"), "y": bounds.get("minY")}, self.state, "outlet minimum")
        hi = validate_world_point({"x": bounds.get("maxX"), "y": bounds.get("maxY")}, self.state, "outlet maximum")
        if lo["x"] > hi["x"] or lo["y"] > hi["y"]:
            raise RangeError("outlet minimum coordinates must not exceed the maximum.")
        if self.allocated_particle_count == 0 or self._active_count == 0:
            return 0
        try:
            removed = self.ctx.remove_in_outlet([lo["x"], lo["y"], hi["x"], hi["y"]])
        except Exception as cause:
            self._cfl_state_usable = False
            raise RuntimeError("Failed to remove liquid particles from the GPU outlet region.") from cause
        if removed > self._active_count:
            raise RuntimeError("GPU reported an outlet removal count above the active particle count.")
        self._active_count -= removed
        self._has_removed = self._has_removed or removed > 0
        return removed

    def clone(self, selection: dict, offset: dict) -> bool:
        """clone (solver.ts:2852-2948): copy the selected object, translated
        by the offset and clamped to the domain, as a new object after the
        active slots.  False when the reference refuses (after an outlet
        removal, for an object with fracture bonds, or without capacity)."""
        self._assert_active()
        src = self._resolve_selection(selection)
        t = validate_clone_offset(offset)
        if self._has_removed or src["objectId"] in self._bonded_objects:
            return False
        if src["count"] > self.particle_capacity - self.allocated_particle_count:
            return False
        if self._next_object_id > MAX_U32_OBJECT_ID:
            raise RangeError("No unsigned 32-bit object ids remain for MLS-MPM cloning.")
        self.end_drag()
        start = self.allocated_particle_count
        object_id = self._next_object_id
        try:
            self.ctx.clone(src["start"], src["count"], t["x"], t["y"], object_id)
        except NativeError as err:
            if err.code in (MPM_ECAPACITY, MPM_ESTATE):
                return False
            raise RuntimeError("Failed to clone the selected MLS-MPM object on the GPU.") from err
        self._object_ranges[object_id] = {"objectId": object_id, "start": start, "count": src["count"]}
        bodies = self._object_bodies.get(src["objectId"], 0)
        if bodies:
            self._object_bodies[object_id] = bodies
            self._body_count += bodies
        self._active_count += src["count"]
        self._next_object_id += 1
        return True

    def read_fracture_bonds(self) -> list[dict]:
        """The committed state of the fracture bonds (damage, broken), as the
        reference's step result reports them (reference/solver.ts:815-819)."""
        self._assert_active()
        n = 0 if self.state.fracture_bonds is None else self.state.fracture_bonds.shape[0]
        return layout.unpack_fracture_bonds(self.ctx.read_fracture_bonds(n))

    def dispose(self):
        if not self._disposed:
            self._disposed = True
            self.ctx.close()

    # ------------------------------------------------------- parity helpers
    def set_diagnostics(self, enable: bool = True):
        """Turn the FP64 mass/momentum reductions of each batch on or off."""
        self._assert_active()
        self.ctx.set_diagnostics(enable)

    def read_diagnostics(self) -> dict:
        """MpmStepDiagnostics (reference/solver.ts:96-103) of the last batch:
        input = the state it started from, grid = its last P2G, output = the
        state it committed; FP64 sums on the GPU."""
        self._assert_active()
        d = self.ctx.read_diagnostics()
        return {"inputMass": float(d[0]), "gridMass": float(d[1]), "outputMass": float(d[2]),
                "inputMomentum": {"x": float(d[3]), "y": float(d[4])},
                "gridMomentum": {"x": float(d[5]), "y": float(d[6])},
                "outputMomentum": {"x": float(d[7]), "y": float(d[8])}}

    def read_fixed_grid(self) -> np.ndarray:
        return self.ctx.read_fixed_grid()

    def read_grid(self) -> np.ndarray:
        return self.ctx.read_grid()

    def read_bins(self):
        return self.ctx.read_bins()


class MpmMlsGpuBatch:
    """Independent scenes advanced together: one launch per microstep covers
    the tiles of every scene (mpm_batch_* in include/mpm355.h).  The
    reference has no counterpart (one solver per scene, solver.ts:390-406).
    scene(i) is an MpmMlsGpuSolver over scene i: reads, status, diagnostics
    and single-scene steps work on it as on any solver."""

    def __init__(self, scenes, p2g_mode: int = MPM_P2G_GATHER, device: int = -1, tuning: dict | None = None):
        if not scenes:
            raise RangeError("a batch needs at least one scene")
        from .native import Batch
        states, cfgs, bounds = [], [], []
        for particles, config in scenes:
            st = layout.validate_input(particles, config)
            states.append(st)
            bounds.append(cflmod.derive_cfl_bounds_packed(st.packed, st.particle_mass_f32))
            cfgs.append(MpmConfig(st.config["grid"]["cellsX"], st.config["grid"]["cellsY"], st.cell_size_f32,
                                  st.particle_mass_f32, st.particle_volume_f32, st.wall_friction_f32,
                                  st.particle_capacity, p2g_mode, 0, device, 0))
        self.native = Batch(cfgs, [st.packed for st in states])
        if tuning is not None:
            from .native import MpmTuning
            self.native.set_tuning(MpmTuning.from_dict(tuning))
        self.solvers = [MpmMlsGpuSolver._over_context(st, b, ctx, p2g_mode)
                        for st, b, ctx in zip(states, bounds, self.native.scenes)]

    def __len__(self) -> int:
        return len(self.solvers)

    def scene(self, i: int) -> "MpmMlsGpuSolver":
        return self.solvers[i]

    def step_batch(self, step_count: int, dt, gravity, measure_speed: bool = False):
        """Advance every scene `step_count` microsteps; dt and gravity are one
        value for all scenes or one per scene."""
        if not isinstance(step_count, int) or isinstance(step_count, bool) or step_count < 1:
            raise RangeError("stepCount must be a positive safe integer.")
        n = len(self.solvers)
        dts = list(dt) if isinstance(dt, (list, tuple, np.ndarray)) else [dt] * n
        gs = list(gravity) if isinstance(gravity, (list, tuple)) else [gravity] * n
        if len(dts) != n or len(gs) != n:
            raise RangeError("one dt and one gravity per scene")
        params = [layout.pack_step_parameters(s.state, d, g, s.allocated_particle_count, 0, None)
                  for s, d, g in zip(self.solvers, dts, gs)]
        self.native.step(params, step_count, measure_speed)
        for s in self.solvers:
            s._cfl_state_usable = False

    def dispose(self):
        for s in self.solvers:
            s._disposed = True
        self.native.close()


def create_mpm_mls_gpu_solver(particles, config: dict, cfl_bounds=None, **kw) -> MpmMlsGpuSolver:
    """createMpmMlsGpuSolver (solver.ts:390-397): float32 gather P2G."""
    return MpmMlsGpuSolver(particles, config, cfl_bounds, MPM_P2G_GATHER, **kw)


def create_experimental_mpm_mls_atomic_gpu_solver(particles, config: dict, cfl_bounds=None,
                                                  **kw) -> MpmMlsGpuSolver:
    """createExperimentalMpmMlsAtomicGpuSolver (solver.ts:399-406): int32 fixed-point P2G."""
    return MpmMlsGpuSolver(particles, config, cfl_bounds, MPM_P2G_FIXED_POINT, **kw)
