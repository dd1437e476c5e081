"""Host-side CFL planning: macrostep -> microstep scheduling and bounds.

Restates, for the host that drives the native step:

* src/lib/mpm/kernels/cfl.ts:121-315 — safe dt, the integer microstep plan
  whose float32 microDt never exceeds safe dt, and the float32 velocity
  allowance the GPU guard enforces;
* src/lib/mpm/kernels/cfl-bounds.ts:32-208 — material wave-speed bounds and
  the external-acceleration velocity envelope (with the 8-ulp round-up);
* src/lib/mpm/kernels/step-scheduler.ts:50-200 — macrostep batches with
  partial batches and backlog.
"""
from __future__ import annotations

import math
import re
import struct
from dataclasses import dataclass

import numpy as np

from .layout import RangeError

DEFAULT_MPM_COURANT_LIMIT = 0.3   # cfl.ts:1
MAX_FLOAT32 = 3.4028234663852886e38
F32_ROUNDING_ULPS = 8             # cfl-bounds.ts:19


@dataclass(frozen=True)
class MicrostepPlan:
    stepCount: int
    microDt: float
    safeDt: float


@dataclass(frozen=True)
class CflBounds:
    waveSpeedBound: float
    velocityBound: float
    courantLimit: float = DEFAULT_MPM_COURANT_LIMIT


def fround(x: float) -> float:
    with np.errstate(over="ignore"):
        return float(np.float32(x))


def _f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def _bits_f32(b: int) -> float:
    return struct.unpack("<f", struct.pack("<I", b & 0xFFFFFFFF))[0]


def _decompose_positive(value: float):
    """mantissa in [0.5, 1) and exponent with value = mantissa * 2**exponent."""
    m, e = math.frexp(value)
    return m, e


def _scale_pow2(mantissa: float, exponent: int) -> float:
    if exponent > 1024:
        return math.inf
    if exponent < -1074:
        return 0.0
    return math.ldexp(mantissa, exponent)


def _safe_dt(courant: float, cell: float, wave: float, vel: float) -> float:
    """calculateSafeDt (cfl.ts:52-95) with the overflow-safe fallback."""
    speed = wave + vel
    if speed == 0:
        return math.inf
    dist = courant * cell
    if math.isfinite(speed) and math.isfinite(dist) and dist > 0:
        safe = dist / speed
        if safe > 0:
            return safe
    scale = max(wave, vel)
    scaled_speed = wave / scale + vel / scale
    lm, le = _decompose_positive(courant)
    cm, ce = _decompose_positive(cell)
    sm, se = _decompose_positive(scale)
    qm, qe = _decompose_positive(scaled_speed)
    mantissa = (lm * cm) / (sm * qm)
    exponent = le + ce - se - qe
    while mantissa >= 1:
        mantissa *= 0.5
        exponent += 1
    while mantissa < 0.5:
        mantissa *= 2
        exponent -= 1
    return _scale_pow2(mantissa, exponent)


def calculate_safe_dt(cellSize, waveSpeedBound, velocityBound, courantLimit=DEFAULT_MPM_COURANT_LIMIT) -> float:
    """calculateMpmCflSafeDt (cfl.ts:121-149)."""
    for name, v, positive in (("cellSize", cellSize, True), ("waveSpeedBound", waveSpeedBound, False),
                              ("velocityBound", velocityBound, False), ("courantLimit", courantLimit, True)):
        if not isinstance(v, (int, float)) or not math.isfinite(v) or (v <= 0 if positive else v < 0):
            qual = "positive and finite" if positive else "finite and non-negative"
            raise RangeError(f"{name} must be {qual}.")
    safe = _safe_dt(courantLimit, cellSize, waveSpeedBound, velocityBound)
    if math.isnan(safe) or safe <= 0:
        raise RangeError("safeDt is outside the positive representable numeric range.")
    return safe


def calculate_microsteps(macroDt, cellSize, waveSpeedBound, velocityBound, maxMicrosteps, minimumStepCount=1,
                         courantLimit=DEFAULT_MPM_COURANT_LIMIT) -> MicrostepPlan:
    """calculateMpmCflMicrosteps (cfl.ts:151-232)."""
    if not isinstance(macroDt, (int, float)) or not math.isfinite(macroDt) or macroDt <= 0:
        raise RangeError("macroDt must be positive and finite.")
    if not isinstance(maxMicrosteps, int) or maxMicrosteps < 1:
        raise RangeError("maxMicrosteps must be a positive safe integer.")
    if not isinstance(minimumStepCount, int) or minimumStepCount < 1:
        raise RangeError("minimumStepCount must be a positive safe integer.")
    safe = calculate_safe_dt(cellSize, waveSpeedBound, velocityBound, courantLimit)
    ratio = macroDt / safe
    count = max(minimumStepCount, math.ceil(ratio)) if math.isfinite(ratio) else math.inf
    if count > maxMicrosteps:
        raise RangeError(f"required microstep count ({count}) exceeds maxMicrosteps ({maxMicrosteps}).")
    count = int(count)
    micro = macroDt / count
    gpu = fround(micro)
    if not math.isfinite(gpu):
        minimum = math.ceil(macroDt / MAX_FLOAT32)
        if minimum > maxMicrosteps:
            raise RangeError(f"required float32 microstep count ({minimum}) exceeds maxMicrosteps ({maxMicrosteps}).")
        count = max(count, minimum)
        micro = macroDt / count
        gpu = fround(micro)
    while micro > safe or gpu > safe:
        if gpu == 0:
            raise RangeError("microDt is too small to represent as a positive float32 value.")
        count += 1
        if count > maxMicrosteps:
            raise RangeError(f"required microstep count ({count}) exceeds maxMicrosteps ({maxMicrosteps}).")
        micro = macroDt / count
        gpu = fround(micro)
    if gpu <= 0:
        raise RangeError("microDt is too small to represent as a positive float32 value.")
    if not math.isfinite(gpu):
        raise RangeError("microDt cannot be represented as a finite float32 value.")
    return MicrostepPlan(count, micro, safe)


def _round_down_f32(value: float) -> float:
    """roundDownFloat32 (cfl.ts:275-295)."""
    if not math.isfinite(value) or value < 0:
        raise RangeError("CFL velocity limits must be finite and non-negative.")
    if value == 0:
        return 0.0
    r = fround(value)
    if not math.isfinite(r):
        raise RangeError("CFL velocity limit cannot be represented as a finite float32 value.")
    if r > value:
        r = _bits_f32(_f32_bits(r) - 1)
    return r


def _round_up_f32_once(value: float) -> float:
    """roundUpFloat32 of cfl.ts:297-315 (no extra ulps)."""
    if not math.isfinite(value) or value < 0:
        raise RangeError("CFL wave-speed bounds must be finite and non-negative.")
    if value == 0:
        return 0.0
    r = fround(value)
    if not math.isfinite(r):
        raise RangeError("CFL wave-speed bound cannot be represented as a finite float32 value.")
    if r < value:
        r = _bits_f32(_f32_bits(r) + 1)
    return r


def calculate_velocity_limit(cellSize, waveSpeedBound, microDt, courantLimit=DEFAULT_MPM_COURANT_LIMIT) -> float:
    """calculateMpmCflVelocityLimit (cfl.ts:234-273)."""
    if not math.isfinite(cellSize) or cellSize <= 0:
        raise RangeError("cellSize must be positive and finite.")
    if not math.isfinite(waveSpeedBound) or waveSpeedBound < 0:
        raise RangeError("waveSpeedBound must be finite and non-negative.")
    if not math.isfinite(microDt) or microDt <= 0:
        raise RangeError("microDt must be positive and finite.")
    if not math.isfinite(courantLimit) or courantLimit <= 0:
        raise RangeError("courantLimit must be positive and finite.")
    gcell = _round_down_f32(cellSize)
    gwave = _round_up_f32_once(waveSpeedBound)
    gdt = fround(microDt)
    if gcell <= 0 or not math.isfinite(gdt) or gdt <= 0:
        raise RangeError("CFL inputs must be representable as positive float32 values.")
    dist = courantLimit * gcell
    total = dist / gdt
    limit = total - gwave
    if not math.isfinite(dist) or dist <= 0 or not math.isfinite(total) or not math.isfinite(limit) or limit < 0:
        raise RangeError("microDt leaves no finite non-negative CFL velocity allowance.")
    return _round_down_f32(limit)


def round_up_float32(value: float) -> float:
    """roundUpFloat32 of cfl-bounds.ts:186-208: round up, then 8 more ulps."""
    if not math.isfinite(value) or value < 0:
        raise RangeError("CFL bounds must be finite and non-negative.")
    if value == 0:
        return 0.0
    bits = _f32_bits(fround(value))
    while _bits_f32(bits) < value:
        bits += 1
    bits += F32_ROUNDING_ULPS
    r = _bits_f32(bits)
    if not math.isfinite(r):
        raise RangeError("CFL bound cannot be represented as a finite float32 value.")
    return r


def particle_wave_speed_bound(material: dict, F: tuple, volume: float | None, particle_mass: float,
                              particle_volume: float) -> float | None:
    """deriveParticleWaveSpeedBound (cfl-bounds.ts:64-142)."""
    lam, mu = fround(material["lambda"]), fround(material["mu"])
    if not math.isfinite(lam) or not math.isfinite(mu):
        raise RangeError("CFL material parameters must be finite float32 values.")
    model = material.get("model", "neo-hookean")
    f00, f01, f10, f11 = (fround(v) for v in F)
    if model in ("water", "oil"):
        density = fround(material["density"])
        if not math.isfinite(density) or density <= 0 or lam <= 0:
            raise RangeError("liquid CFL parameters require positive finite lambda and density.")
        eos = material.get("equationOfState")
        if not eos or eos.get("model") != "tait":
            return math.sqrt(lam / density)
        n = fround(eos["exponent"])
        det = fround(f00 * f11 - f01 * f10)
        if not math.isfinite(n) or n <= 0 or not math.isfinite(det) or det <= 0:
            raise RangeError("Tait CFL bounds require a positive finite exponent and determinant.")
        ws2 = (n * lam * math.pow(1 / det, n - 1)) / density
        if not math.isfinite(ws2) or ws2 < 0:
            raise RangeError("Tait CFL wave-speed bound is outside the finite range.")
        return math.sqrt(ws2)
    if model != "neo-hookean":
        return None
    if lam < 0 or mu < 0:
        raise RangeError("Neo-Hookean CFL parameters must be non-negative.")
    vol = fround(particle_volume if volume is None else volume)
    mass = fround(particle_mass)
    rho = mass / vol
    det = f00 * f11 - f01 * f10
    if not (mass > 0 and vol > 0 and math.isfinite(rho) and rho > 0 and math.isfinite(det) and det > 0):
        raise RangeError("Neo-Hookean CFL bounds require positive density and determinant.")
    tr = f00 * f00 + f01 * f01 + f10 * f10 + f11 * f11
    disc = math.sqrt(max(0.0, tr * tr - 4 * det * det))
    smax = 0.5 * (tr + disc)
    vt = abs(lam + mu - lam * math.log(det))
    ws2 = (mu * smax + vt) / rho
    if not math.isfinite(ws2) or ws2 < 0:
        raise RangeError("Neo-Hookean CFL wave-speed bound is outside the finite range.")
    return math.sqrt(ws2)


def derive_cfl_bounds(particles, config: dict) -> CflBounds | None:
    """deriveMpmMlsGpuCflBounds (cfl-bounds.ts:32-62) over MpmParticle dicts."""
    velocity = 0.0
    wave = 0.0
    for i, p in enumerate(particles):
        vx, vy = fround(p["velocity"]["x"]), fround(p["velocity"]["y"])
        if not math.isfinite(vx) or not math.isfinite(vy):
            raise RangeError(f"particle {i} has a non-finite initial velocity.")
        velocity = max(velocity, math.hypot(vx, vy))
        mat = p.get("material") or config["material"]
        F = p["F"]
        w = particle_wave_speed_bound(mat, (F["m00"], F["m01"], F["m10"], F["m11"]), p.get("volume"),
                                      config["particleMass"], config["particleVolume"])
        if w is None:
            return None
        wave = max(wave, w)
    return CflBounds(round_up_float32(wave), round_up_float32(velocity))


def derive_cfl_bounds_packed(packed: np.ndarray, particle_mass: float) -> CflBounds | None:
    """The same bound over packed 20-word records (vectorised for large scenes)."""
    rec = np.asarray(packed, np.uint32).reshape(-1, 20)
    if rec.shape[0] == 0:
        return CflBounds(0.0, 0.0)
    fv = rec.view(np.float32).astype(np.float64)
    model = rec[:, 13]
    if np.any((model != 0) & (model != 2) & (model != 3)):
        return None
    vel = float(np.max(np.hypot(fv[:, 2], fv[:, 3])))
    lam, mu, dens, vol, param = fv[:, 14], fv[:, 15], fv[:, 16], fv[:, 17], fv[:, 18]
    liquid = (model == 2) | (model == 3)
    f00, f01, f10, f11 = fv[:, 8], fv[:, 9], fv[:, 10], fv[:, 11]
    ws = np.zeros(rec.shape[0])
    with np.errstate(divide="ignore", invalid="ignore"):
        tait = liquid & (param > 0)
        det32 = (f00 * f11 - f01 * f10).astype(np.float32).astype(np.float64)
        ws_l = np.where(tait, param * lam * np.power(1 / det32, param - 1) / dens, lam / dens)
        det = f00 * f11 - f01 * f10
        tr = f00 ** 2 + f01 ** 2 + f10 ** 2 + f11 ** 2
        smax = 0.5 * (tr + np.sqrt(np.maximum(0, tr * tr - 4 * det * det)))
        rho = np.float32(particle_mass) / vol
        ws_s = (mu * smax + np.abs(lam + mu - lam * np.log(det))) / rho
    ws = np.sqrt(np.where(liquid, ws_l, ws_s))
    if not np.all(np.isfinite(ws)):
        raise RangeError("CFL wave-speed bound is outside the finite range.")
    return CflBounds(round_up_float32(float(np.max(ws))), round_up_float32(vel))


def calculate_velocity_bound(currentVelocityBound, durationSeconds, gravityMagnitude, oscillationAmplitude=0.0,
                             dragRadius=0.0, dragStiffness=0.0, dragDamping=0.0) -> float:
    """calculateMpmCflVelocityBound (cfl-bounds.ts:153-184)."""
    for name, v in (("currentVelocityBound", currentVelocityBound), ("durationSeconds", durationSeconds),
                    ("gravityMagnitude", gravityMagnitude), ("oscillationAmplitude", oscillationAmplitude),
                    ("dragRadius", dragRadius), ("dragStiffness", dragStiffness), ("dragDamping", dragDamping)):
        if not math.isfinite(v) or v < 0:
            raise RangeError(f"{name} must be finite and non-negative.")
    acc = math.hypot(gravityMagnitude, oscillationAmplitude) + dragRadius * dragStiffness
    bound = currentVelocityBound + acc * durationSeconds
    if not math.isfinite(bound):
        raise RangeError("external-input velocity bound exceeds the finite numeric range.")
    return round_up_float32(bound)


_CAPACITY_ERROR = re.compile(r"exceeds maxMicrosteps|microDt is too small to represent")


def is_microstep_capacity_error(error: Exception) -> bool:
    return isinstance(error, RangeError) and bool(_CAPACITY_ERROR.search(str(error)))


@dataclass(frozen=True)
class StepBatch:
    macroStepCount: int
    macroDt: float
    microStepCount: int
    simulatedSeconds: float
    remainderSeconds: float
    microstepPlan: MicrostepPlan


def calculate_step_batch(accumulatorSeconds, elapsedSeconds, simulationSpeed, fixedStepSeconds, cellSize,
                         waveSpeedBound, velocityBound, maxMicrosteps, courantLimit=DEFAULT_MPM_COURANT_LIMIT,
                         allowPartialBatch=False) -> StepBatch:
    """calculateMpmStepBatch (step-scheduler.ts:50-158)."""
    if not math.isfinite(accumulatorSeconds) or accumulatorSeconds < 0:
        raise RangeError("accumulatorSeconds must be finite and non-negative.")
    if not math.isfinite(elapsedSeconds) or elapsedSeconds < 0:
        raise RangeError("elapsedSeconds must be finite and non-negative.")
    if not math.isfinite(simulationSpeed) or simulationSpeed < 0:
        raise RangeError("simulationSpeed must be finite and non-negative.")
    if not math.isfinite(fixedStepSeconds) or fixedStepSeconds <= 0:
        raise RangeError("fixedStepSeconds must be positive and finite.")
    accumulated = accumulatorSeconds + elapsedSeconds * simulationSpeed
    if not math.isfinite(accumulated):
        raise RangeError("accumulated simulation time is outside the supported range.")
    macro = fixedStepSeconds
    try:
        plan = calculate_microsteps(macro, cellSize, waveSpeedBound, velocityBound, maxMicrosteps,
                                    courantLimit=courantLimit)
    except RangeError as err:
        if not allowPartialBatch or not is_microstep_capacity_error(err):
            raise
        safe = calculate_safe_dt(cellSize, waveSpeedBound, velocityBound, courantLimit)
        reserved = max(1, maxMicrosteps - (1 if maxMicrosteps > 1 else 0))
        macro = min(fixedStepSeconds, safe * reserved * (1 - 1e-6))
        if not math.isfinite(macro) or macro <= 0:
            raise RangeError(f"No positive CFL-safe macrostep duration fits the {maxMicrosteps}-microstep batch cap.")
        plan = calculate_microsteps(macro, cellSize, waveSpeedBound, velocityBound, maxMicrosteps,
                                    courantLimit=courantLimit)
    requested = math.floor(accumulated / fixedStepSeconds)
    if macro < fixedStepSeconds:
        count = 1 if requested > 0 else 0
    elif allowPartialBatch:
        count = min(requested, maxMicrosteps // plan.stepCount)
    else:
        count = requested
    micro_count = count * plan.stepCount
    if micro_count > maxMicrosteps:
        raise RangeError(f"requested microstep count ({micro_count}) exceeds maxMicrosteps ({maxMicrosteps}).")
    simulated = count * macro
    return StepBatch(count, macro, micro_count, simulated, max(0.0, accumulated - simulated), plan)


def calculate_cfl_step_batch(currentVelocityBound, gravityMagnitude, oscillationAmplitude, dragRadius,
                             dragStiffness, dragDamping, **batch_input):
    """calculateMpmCflStepBatch (step-scheduler.ts:160-200) -> (batch, velocityBound)."""
    preview = calculate_step_batch(velocityBound=currentVelocityBound, **batch_input)
    bound = calculate_velocity_bound(currentVelocityBound, preview.simulatedSeconds, gravityMagnitude,
                                     oscillationAmplitude, dragRadius, dragStiffness, dragDamping)
    batch = calculate_step_batch(velocityBound=bound, **batch_input)
    gpu_dt = fround(batch.microstepPlan.microDt)
    gpu_damp = fround(dragDamping)
    if batch.macroStepCount > 0 and dragRadius > 0 and (
            not math.isfinite(gpu_dt) or gpu_dt <= 0 or not math.isfinite(gpu_damp) or gpu_dt * gpu_damp > 1):
        raise RangeError("CFL microDt is too large for the explicit drag damping update to remain dissipative.")
    return batch, bound


HORIZONTAL_OSCILLATION_FREQUENCY_HZ = 0.16  # kernels/oscillation.ts:5


def mean_sinusoidal_acceleration(startSeconds, durationSeconds, amplitude,
                                 frequencyHz=HORIZONTAL_OSCILLATION_FREQUENCY_HZ) -> float:
    """meanSinusoidalAcceleration (kernels/oscillation.ts:7-32): the mean of
    amplitude * sin(2 pi f t) over [start, start + duration]."""
    for name, v in (("startSeconds", startSeconds), ("durationSeconds", durationSeconds),
                    ("amplitude", amplitude), ("frequencyHz", frequencyHz)):
        if not isinstance(v, (int, float)) or not math.isfinite(v):
            raise RangeError("Oscillation inputs must be finite, with non-negative duration and frequency.")
    if durationSeconds < 0 or frequencyHz < 0:
        raise RangeError("Oscillation inputs must be finite, with non-negative duration and frequency.")
    if amplitude == 0 or frequencyHz == 0:
        return 0.0
    w = 2 * math.pi * frequencyHz
    half = (w * durationSeconds) / 2
    sinc = 1.0 if half == 0 else math.sin(half) / half
    return amplitude * math.sin(w * (startSeconds + durationSeconds / 2)) * sinc
