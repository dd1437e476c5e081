"""The adaptive CFL loop's failure paths (stepCflBatchAdaptive,
src/lib/mpm/kernels/solver.ts:1724-2021), driven from the host with a fake of
the native context that scripts the 40-byte compact readbacks, as the
reference's own spec does with its mocked GPU (kernels/solver.spec.ts:624-838).
CPU only: the fake replaces the C-ABI context, so no kernel runs."""
from __future__ import annotations

import struct

import numpy as np
import pytest

CONFIG = {"grid": {"cellsX": 16, "cellsY": 8, "cellSize": 1}, "particleMass": 0.1, "particleVolume": 0.01,
          "material": {"lambda": 20, "mu": 10}, "capacity": 4}           # solver.spec.ts:129-135
PARTICLE = {"position": {"x": 8, "y": 4}, "velocity": {"x": 0, "y": 0},
            "C": {"m00": 0, "m01": 0, "m10": 0, "m11": 0}, "F": {"m00": 1, "m01": 0, "m10": 0, "m11": 1}}


def f2u(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def ok(speed: float, completed: int):
    return speed, completed, np.array([0, 0xFFFFFFFF, 0, 0, 0, 0, 0, 0], np.uint32)


def cfl_failure(speed: float, completed: int, wave: float):
    """Status words of a CFL-bound failure at particle 0 (kind 3): particle
    speed 50, wave speed, velocity limit 40, wave limit."""
    return speed, completed, np.array([16384, 0, f2u(50.0), f2u(wave), f2u(40.0), f2u(wave), 3, 0], np.uint32)


class FakeContext:
    """The parts of native.Context the host loop uses; read_compact replays a
    script and step_batch records what was submitted."""

    def __init__(self, script, n=1, capacity=4):
        self.script = list(script)
        self.n, self.capacity = n, capacity
        self.submitted = []   # (microsteps, dt, wave bound, velocity bound, guard on)
        self.cleared = 0
        self.status = ok(0.0, 0)[2]

    def step_batch(self, params: bytes, steps: int, measure: bool):
        dt, = struct.unpack_from("<f", params, 44)
        wave, vel, on = struct.unpack_from("<3f", params, 96)
        self.submitted.append((steps, dt, wave, vel, on))

    def read_compact(self):
        speed, done, status = self.script.pop(0)
        self.status = status
        return speed, done, status

    def read_status(self):
        return self.status

    def clear_status(self):
        self.cleared += 1
        self.status = ok(0.0, 0)[2]

    def counts(self):
        return self.n, self.capacity, (16 + 3) * (8 + 3)

    def close(self):
        pass


def make_solver(pkg, script, bounds=None):
    state = pkg.layout.validate_input([PARTICLE], CONFIG)
    bounds = bounds or pkg.cfl.CflBounds(waveSpeedBound=25.0, velocityBound=1.0)
    ctx = FakeContext(script)
    return pkg.MpmMlsGpuSolver._over_context(state, bounds, ctx, pkg.native.MPM_P2G_GATHER), ctx


def plan(pkg, macro, wave, vel, maxm, courant=None):
    kw = {} if courant is None else {"courantLimit": courant}
    return pkg.cfl.calculate_microsteps(macro, 1.0, wave, vel, maxm, **kw)


G = {"x": 0.0, "y": 0.0}


def test_cfl_violation_surfaces_and_invalidates_the_bound(pkg):
    """solver.spec.ts:624-656: the compact status of a CFL failure raises the
    numerical error, and a further CFL batch is refused."""
    s, ctx = make_solver(pkg, [cfl_failure(50.0, 0, 25.0)])
    p = plan(pkg, 0.001, 25.0, 1.0, 1)
    s.step_cfl_batch(1, 0.001, p, 1.0, 1, G)
    with pytest.raises(pkg.MpmMlsGpuNumericalError, match="CFL wave-speed or velocity bound at particle 0"):
        s.measure_maximum_particle_speed()
    with pytest.raises(RuntimeError, match="failed CFL measurement invalidated the tracked velocity bound"):
        s.step_cfl_batch(1, 0.001, p, 1.0, 1, G)


def test_rejected_microstep_is_retried_finer_with_raised_bounds(pkg):
    """solver.spec.ts:659-719: three macrosteps of one microstep; the first
    batch fails after one committed microstep; the failed interval is retried
    as two microsteps with the velocity bound raised to 1.25 x the observed
    speed and the wave bound to 1.25 x the reported one; then the rest."""
    s, ctx = make_solver(pkg, [cfl_failure(50.0, 1, 25.0), ok(50.0, 2), ok(50.0, 1)])
    p = plan(pkg, 0.001, 25.0, 1.0, 8)
    assert p.stepCount == 1
    m = s.step_cfl_batch_adaptive(3, 0.001, p, 1.0, 8, G)
    assert (m.maximumSpeed, m.sequence, m.attemptedMicroStepCount, m.retryCount) == (50.0, 3, 7, 1)
    assert m.remainingDuration == 0 and m.simulatedDuration == pytest.approx(0.003, rel=1e-12)
    steps = [sub[0] for sub in ctx.submitted]
    assert steps == [3, 2, 2]
    assert ctx.cleared == 1
    # the retry: two microsteps over the failed 1 ms, bounds x1.25
    _, dt, wave, vel, on = ctx.submitted[1]
    assert on == 1.0 and dt == np.float32(0.0005)
    assert wave == np.float32(31.25) and s.cfl_bounds.waveSpeedBound == pytest.approx(31.25)
    assert vel >= 62.5


def test_recovery_continues_beyond_four_retries(pkg):
    """solver.spec.ts:721-759: five failed attempts in a row, each retried
    finer while the budget of 64 microsteps lasts, then success."""
    script = [cfl_failure(50.0, 0, 25.0)] * 5 + [ok(50.0, 0)] * 8
    s, ctx = make_solver(pkg, script)
    p = plan(pkg, 0.001, 25.0, 1.0, 64)
    m = s.step_cfl_batch_adaptive(1, 0.001, p, 1.0, 64, G)
    assert m.retryCount == 5 and m.remainingDuration == 0
    assert all(sub[0] >= 2 for sub in ctx.submitted[1:6])  # every retry at least two microsteps
    assert ctx.cleared == 5


def test_partial_progress_then_resume(pkg):
    """solver.spec.ts:761-838: with a budget of four microsteps the retry
    cannot fit after a failure at microstep 1 of 3, so the loop returns the
    valid progress (1 ms of 3) and the recommended bound 62.5; a second call
    with a plan at that bound finishes the remaining 2 ms."""
    script = [cfl_failure(50.0, 1, 25.0)] + [ok(0.5, 2)] * 4
    s, ctx = make_solver(pkg, script)
    p = plan(pkg, 0.001, 25.0, 1.0, 4)
    partial = s.step_cfl_batch_adaptive(3, 0.001, p, 1.0, 4, G)
    assert partial.maximumSpeed == 50.0 and partial.sequence == 1
    assert partial.recommendedVelocityBound == 62.5
    assert (partial.attemptedMicroStepCount, partial.retryCount) == (3, 0)
    assert partial.simulatedDuration == pytest.approx(0.001, rel=1e-12)
    assert partial.remainingDuration == pytest.approx(0.002, rel=1e-12)
    bound = partial.recommendedVelocityBound
    cont_plan = plan(pkg, 0.001, s.cfl_bounds.waveSpeedBound, bound, 8, s.cfl_bounds.courantLimit)
    cont = s.step_cfl_batch_adaptive(2, 0.001, cont_plan, bound, 8, G)
    assert cont.attemptedMicroStepCount == 2 and cont.remainingDuration == 0
    assert cont.simulatedDuration == pytest.approx(0.002, rel=1e-12)
    assert sum(sub[0] for sub in ctx.submitted) == 5  # 3 + 2 committed-microstep launches


def test_unrecoverable_failure_is_raised(pkg):
    """A failure that is not a CFL or deformation failure (here the P2G
    validation bit) is not retried (isRecoverableCflError, solver.ts:296-312)."""
    bad = (50.0, 0, np.array([4, 0, 0, 0, 0, 0, 0, 0], np.uint32))
    s, ctx = make_solver(pkg, [bad])
    p = plan(pkg, 0.001, 25.0, 1.0, 8)
    with pytest.raises(pkg.MpmMlsGpuNumericalError, match="P2G validation failed"):
        s.step_cfl_batch_adaptive(1, 0.001, p, 1.0, 8, G)
    assert len(ctx.submitted) == 1
