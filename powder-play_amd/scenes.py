"""Synthetic scenes of the step path's workloads.

* ``lcg_uniform`` — the Gate 4 scene generator's LCG
  (src/lib/mpm/scenes/gate4.ts:134-144): state = state*1664525 + 1013904223
  mod 2^32, value = state / 2^32; evaluated in bulk by jump-ahead.
* ``tait_tank`` — SI Tait water tank: 40x40 particles at 2.5 mm filling the
  lower-left 100x100 mm of a 200x200 mm box, 80x80 cells, rho0 1000, n 7,
  c 20 m/s (Tait material mapping: reference/tait-eos.ts:26-86; SI box:
  si-domain.ts:1-12; wall friction 1 as scenes/si-fluid.ts:21).
* ``water_oil_column`` — 1,000,000 particles, 60% water / 40% oil of the Gate 4
  liquids (scenes/gate4.ts:82-93), one per cell of a 1000x1000 grid, jittered,
  velocities U(+-2).
* ``five_model_block`` — the 45-particle five-material fixture of
  experiments/transfer-regression.ts:34-53.
* ``batched_tanks`` — independent Tait water tanks of log-uniform size for
  batched stepping: the SI tank's material and cell size, a dam-break block
  at one particle per cell in a corner, gravity in a random direction.
* ``gate4_scene`` — the multi-material Gate 4 scene: water, oil, 50 bonded
  triangle pebbles and a fracture solid (scenes/gate4.ts:184-357).
* ``create_mpm_emitter_particles`` — the deterministic emitter burst of
  scenes/emitter.ts:49-123, as MpmParticle dicts for append_particles.
* ``elastic_block`` — the differentiable fixed-corotated block: 64x64
  particles (lambda 200, mu 100, density 1) launched into the floor of a
  256x256 grid for 200 steps of 1e-3, with a seeded centre-of-mass target.
"""
from __future__ import annotations

import math

import numpy as np

from .layout import ParticleArrays, RangeError

LCG_A = 1664525
LCG_C = 1013904223
MASK32 = 0xFFFFFFFF
GATE4_SCENE_SEED = 0x4D504D34          # gate4.ts:10
GATE4_WATER = {"model": "water", "lambda": 160000.0, "mu": 0.0, "density": 0.02}   # gate4.ts:82-87
GATE4_OIL = {"model": "oil", "lambda": 120000.0, "mu": 0.0, "density": 0.016}      # gate4.ts:88-93


def lcg_uniform(seed: int, n: int) -> np.ndarray:
    """n successive draws of the Gate 4 LCG starting from `seed`, as float64 in [0, 1)."""
    if n <= 0:
        return np.zeros(0)
    A = np.zeros(n + 1, np.uint64)
    C = np.zeros(n + 1, np.uint64)
    A[1], C[1] = LCG_A, LCG_C
    length = 1
    while length < n:
        m = min(length, n - length)
        j = np.arange(1, m + 1)
        A[length + j] = (A[j] * A[length]) & MASK32
        C[length + j] = (A[j] * C[length] + C[j]) & MASK32
        length += m
    states = (A[1:] * np.uint64(seed & MASK32) + C[1:]) & MASK32
    return states.astype(np.float64) / 4294967296.0


def tait_material(reference_density=1000.0, exponent=7.0, sound_speed=20.0, extrusion=0.02, model="water") -> dict:
    """createMpmTaitLiquidMaterial + deriveTaitPressureCoefficientPa
    (tait-eos.ts:26-44, :73-86): B = rho0 c^2 / n mapped to areal units."""
    b = reference_density * sound_speed * sound_speed / exponent
    return {"model": model, "lambda": b * extrusion, "mu": 0.0, "density": reference_density * extrusion,
            "equationOfState": {"model": "tait", "exponent": exponent}}


def tait_tank():
    """SI Tait tank: returns (particles as dicts, config, macro dt, macro count)."""
    spacing = 0.0025
    material = tait_material()
    area = spacing * spacing
    particles = []
    for j in range(40):
        for i in range(40):
            particles.append({
                "position": {"x": (i + 0.5) * spacing, "y": (j + 0.5) * spacing},
                "velocity": {"x": 0.0, "y": 0.0},
                "C": {"m00": 0.0, "m01": 0.0, "m10": 0.0, "m11": 0.0},
                "F": {"m00": 1.0, "m01": 0.0, "m10": 0.0, "m11": 1.0},
                "volume": area,
            })
    config = {"grid": {"cellsX": 80, "cellsY": 80, "cellSize": spacing},
              "particleMass": material["density"] * area, "particleVolume": area, "material": material,
              "wallFriction": 1.0}
    return particles, config, 1.0 / 120.0, 240


TAIT_TANK_GRAVITY = {"x": 0.0, "y": -9.81}


def water_oil_column(cells: int = 1000, seed: int = GATE4_SCENE_SEED, water_fraction: float = 0.6,
                     speed: float = 2.0):
    """One particle per cell of a cells x cells grid (cell size 1), jittered
    uniformly inside its cell; water for the first water_fraction of an LCG
    ordering, oil for the rest; velocities U(-speed, speed)^2.
    Returns (ParticleArrays, config)."""
    n = cells * cells
    u = lcg_uniform(seed, 5 * n)
    jx, jy, vx, vy, order_key = (u[k::5] for k in range(5))
    cx = np.tile(np.arange(cells, dtype=np.float64), cells)
    cy = np.repeat(np.arange(cells, dtype=np.float64), cells)
    x = cx + 0.5 + (jx - 0.5)
    y = cy + 0.5 + (jy - 0.5)
    order = np.argsort(order_key, kind="stable")
    is_water = np.zeros(n, bool)
    is_water[order[: int(round(water_fraction * n))]] = True
    model = np.where(is_water, 2, 3).astype(np.uint32)
    arr = ParticleArrays(
        x=x, y=y, vx=(2 * vx - 1) * speed, vy=(2 * vy - 1) * speed, model=model,
        lambda_=np.where(is_water, GATE4_WATER["lambda"], GATE4_OIL["lambda"]),
        mu=np.zeros(n), density=np.where(is_water, GATE4_WATER["density"], GATE4_OIL["density"]),
        volume=np.ones(n), param=np.zeros(n), object_id=np.zeros(n, np.uint32))
    config = {"grid": {"cellsX": cells, "cellsY": cells, "cellSize": 1.0}, "particleMass": 0.02,
              "particleVolume": 1.0, "material": dict(GATE4_WATER), "capacity": n}
    return arr, config


WATER_OIL_GRAVITY = {"x": 0.0, "y": 90.0}
WATER_OIL_MACRO_DT = 1.0 / 240.0

FIVE_MODEL_MATERIALS = [  # transfer-regression.ts:34-40
    {"model": "neo-hookean", "lambda": 20.0, "mu": 10.0},
    {"model": "snow", "lambda": 20.0, "mu": 10.0},
    {"model": "water", "lambda": 8.0, "mu": 0.0, "density": 1.0},
    {"model": "oil", "lambda": 6.0, "mu": 0.0, "density": 0.8},
    {"model": "powder", "lambda": 20.0, "mu": 10.0, "frictionCoefficient": 0.55},
]


def five_model_block(models=None):
    """9 particles per material on a 3x3 patch (transfer-regression.ts:41-53)."""
    mats = FIVE_MODEL_MATERIALS if models is None else [m for m in FIVE_MODEL_MATERIALS if m["model"] in models]
    particles = []
    for group, material in enumerate(FIVE_MODEL_MATERIALS):
        if material not in mats:
            continue
        for local in range(9):
            particles.append({
                "position": {"x": 4 + group * 8 + local % 3, "y": 4 + local // 3},
                "velocity": {"x": 0.03 * group, "y": -0.02},
                "C": {"m00": 0.0, "m01": 0.001, "m10": 0.0, "m11": 0.0},
                "F": {"m00": 0.98, "m01": 0.01, "m10": 0.0, "m11": 0.99},
                "material": material, "objectId": group + 1, "plasticVolume": 1.0,
            })
    config = {"grid": {"cellsX": 48, "cellsY": 16, "cellSize": 1.0}, "particleMass": 0.1,
              "particleVolume": 0.05, "material": FIVE_MODEL_MATERIALS[0], "capacity": 54}
    return particles, config


BATCH_SEED = 0x62617463
BATCH_DT = (1.0 / 120.0) / 224.0  # the SI tank's CFL microstep (first macrostep plan)


def batched_tanks(count: int = 256, seed: int = BATCH_SEED, min_particles: int = 500,
                  max_particles: int = 50_000):
    """`count` independent SI-material Tait tanks.  Scene i has
    N_i = round(exp(U(ln min, ln max))) particles at one per cell (cell size
    2.5 mm, as tait_tank) in a block from cell (2, 2), on a square grid of
    side round(64 * 8^((ln N_i - ln min) / ln(max / min))) in [64, 512]
    cells, with gravity 9.81 m/s^2 at an angle U(0, 2 pi).  Draws come from
    the Gate 4 LCG.  Returns a list of (ParticleArrays, config, gravity)."""
    u = lcg_uniform(seed, 2 * count)
    spacing = 0.0025
    material = tait_material()
    area = spacing * spacing
    lo, hi = np.log(min_particles), np.log(max_particles)
    scenes = []
    for i in range(count):
        n = int(round(np.exp(lo + u[2 * i] * (hi - lo))))
        side = int(round(64 * 8 ** ((np.log(n) - lo) / (hi - lo))))
        w = int(np.ceil(np.sqrt(n)))
        k = np.arange(n)
        x = (2 + k % w + 0.5) * spacing
        y = (2 + k // w + 0.5) * spacing
        theta = 2 * np.pi * u[2 * i + 1]
        arr = ParticleArrays(
            x=x, y=y, vx=np.zeros(n), vy=np.zeros(n), model=np.full(n, 2, np.uint32),
            lambda_=np.full(n, material["lambda"]), mu=np.zeros(n), density=np.full(n, material["density"]),
            volume=np.full(n, area), param=np.full(n, material["equationOfState"]["exponent"]),
            object_id=np.zeros(n, np.uint32))
        config = {"grid": {"cellsX": side, "cellsY": side, "cellSize": spacing},
                  "particleMass": material["density"] * area, "particleVolume": area, "material": material,
                  "wallFriction": 1.0, "capacity": n}
        gravity = {"x": 9.81 * float(np.cos(theta)), "y": 9.81 * float(np.sin(theta))}
        scenes.append((arr, config, gravity))
    return scenes


ELASTIC_SEED = 0x61646A74


def elastic_block(seed: int = ELASTIC_SEED, side: int = 64, spacing: float = 0.5, cells: int = 256,
                  origin=(96.0, 6.0), lam: float = 200.0, mu: float = 100.0, density: float = 1.0,
                  steps: int = 200, dt: float = 1e-3, gravity=(0.0, -9.81)):
    """Differentiable elastic block: a side x side lattice of fixed-corotated
    particles at `spacing` (4 per cell at the defaults) whose lower-left
    corner is `origin`, in a cells x cells grid of cell size 1, launched
    toward the floor (wall y = 0).  Every particle gets the block velocity
    (U(-2, 2), -U(8, 12)) plus its own U(-0.5, 0.5) jitter per axis; the
    target of the centre-of-mass loss is the initial centre of mass plus
    (U(-4, 4), U(-2, 2)).  lambda, mu and density are per particle (uniform
    here); mass = density * spacing^2.  Draws come from the Gate 4 LCG.

    Returns (state0 (n, 12) float64 rows x, y, vx, vy, C00..C11, F00..F11 with
    C = 0 and F = I, lambda (n,), mu (n,), config dict, target (2,))."""
    n = side * side
    u = lcg_uniform(seed, 4 + 2 * n)
    k = np.arange(n)
    state = np.zeros((n, 12))
    state[:, 0] = origin[0] + (k % side + 0.5) * spacing
    state[:, 1] = origin[1] + (k // side + 0.5) * spacing
    state[:, 2] = -2.0 + 4.0 * u[0] + (u[4 + 2 * k] - 0.5)
    state[:, 3] = -(8.0 + 4.0 * u[1]) + (u[5 + 2 * k] - 0.5)
    state[:, 8] = 1.0
    state[:, 11] = 1.0
    area = spacing * spacing
    config = {"cellsX": cells, "cellsY": cells, "cellSize": 1.0, "dt": dt, "steps": steps,
              "particleMass": density * area, "particleVolume": area,
              "gravity": (float(gravity[0]), float(gravity[1])), "wallFriction": 0.0}
    com = state[:, :2].mean(axis=0)
    target = com + np.array([-4.0 + 8.0 * u[2], -2.0 + 4.0 * u[3]])
    return state, np.full(n, lam), np.full(n, mu), config, target


# ---------------------------------------------------------------- emitter
MAX_MPM_EMITTER_PARTICLES = 100_000                  # emitter.ts:9
GATE4_LIQUID_BULK_MODULUS_SCALE = 4                  # gate4.ts:35
_GOLDEN_ANGLE = math.pi * (3 - math.sqrt(5))
EMITTER_MATERIALS = {                                # emitter.ts:28-46
    "water": {"lambda": 40_000 * GATE4_LIQUID_BULK_MODULUS_SCALE, "mu": 0, "model": "water", "density": 0.02},
    "oil": {"lambda": 30_000 * GATE4_LIQUID_BULK_MODULUS_SCALE, "mu": 0, "model": "oil", "density": 0.016},
    "neo-hookean": {"lambda": 20, "mu": 10, "model": "neo-hookean"},
}


def _positive_finite(value, name):
    if not isinstance(value, (int, float)) or isinstance(value, bool) or not math.isfinite(value) or value <= 0:
        raise RangeError(f"{name} must be a positive finite number")
    return float(value)


def create_mpm_emitter_particles(config: dict) -> list[dict]:
    """createMpmEmitterParticles (scenes/emitter.ts:49-123): `count`
    particles of one material around `center`, on a golden-angle spiral
    (circle) or a jittered lattice (rectangle), all with `velocity`, C = 0,
    F = I, plastic volume 1 and `particleVolume` (default 25).  Draws come
    from the seeded LCG state = state * 1664525 + 1013904223 mod 2^32
    (emitter.ts:164-170); the first draw rotates the spiral, then one draw
    per particle (circle) or two (rectangle) jitter it by +-6% of the
    spacing.  Returns MpmParticle-shaped dicts for append_particles."""
    count = config.get("count")
    if not isinstance(count, int) or isinstance(count, bool) or count < 1:
        raise RangeError("count must be a positive safe integer")
    if count > MAX_MPM_EMITTER_PARTICLES:
        raise RangeError(f"count must not exceed {MAX_MPM_EMITTER_PARTICLES}")
    seed = config.get("seed")
    if not isinstance(seed, int) or isinstance(seed, bool) or seed < 0 or seed > 0xFFFFFFFF:
        raise RangeError("seed must be an unsigned 32-bit integer")
    shape = config.get("shape")
    if shape not in ("rectangle", "circle"):
        raise RangeError("shape must be rectangle or circle")
    material = config.get("material")
    if material not in EMITTER_MATERIALS:
        raise RangeError("material must be water, oil, or neo-hookean")
    center, velocity = config.get("center"), config.get("velocity")
    for vec, name in ((center, "center"), (velocity, "velocity")):
        if not isinstance(vec, dict):
            raise TypeError(f"{name} must be a vector")
        if not all(isinstance(vec.get(k), (int, float)) and math.isfinite(vec.get(k)) for k in ("x", "y")):
            raise RangeError(f"{name} components must be finite")
    grid = config.get("grid")
    if not isinstance(grid, dict):
        raise TypeError("grid must be an object")
    cells_x, cells_y = grid.get("cellsX"), grid.get("cellsY")
    if not all(isinstance(v, int) and not isinstance(v, bool) and v >= 1 for v in (cells_x, cells_y)):
        raise RangeError("grid cell counts must be positive safe integers")
    cell_size = _positive_finite(grid.get("cellSize"), "grid.cellSize")
    spacing = _positive_finite(config.get("spacing"), "spacing")
    volume = _positive_finite(config.get("particleVolume", 25), "particleVolume")
    width, height = cells_x * cell_size, cells_y * cell_size
    cx, cy = float(center["x"]), float(center["y"])
    if cx < 0 or cx > width or cy < 0 or cy > height:
        raise RangeError("center must be inside the simulation domain")

    state = seed

    def random() -> float:
        nonlocal state
        state = (state * LCG_A + LCG_C) & MASK32
        return state / 4294967296.0

    columns = math.ceil(math.sqrt(count))
    rows = math.ceil(count / columns)
    radius = spacing * math.sqrt(count / math.pi)
    angular_offset = random() * math.pi * 2
    mat = EMITTER_MATERIALS[material]
    out = []
    for index in range(count):
        if shape == "circle":
            radial = math.sqrt((index + 0.5) / count)
            angle = angular_offset + index * _GOLDEN_ANGLE
            jitter = (random() - 0.5) * spacing * 0.12
            dist = max(0.0, radial * radius + jitter)
            x, y = cx + math.cos(angle) * dist, cy + math.sin(angle) * dist
        else:
            column, row = index % columns, index // columns
            x = cx + (column - (columns - 1) * 0.5) * spacing + (random() - 0.5) * spacing * 0.12
            y = cy + (row - (rows - 1) * 0.5) * spacing + (random() - 0.5) * spacing * 0.12
        if x < 0 or x > width or y < 0 or y > height:
            raise RangeError(f"emitted particle {index} would be outside the simulation domain")
        out.append({"position": {"x": x, "y": y}, "velocity": {"x": float(velocity["x"]), "y": float(velocity["y"])},
                    "C": {"m00": 0.0, "m01": 0.0, "m10": 0.0, "m11": 0.0},
                    "F": {"m00": 1.0, "m01": 0.0, "m10": 0.0, "m11": 1.0},
                    "plasticVolume": 1.0, "volume": volume, "material": dict(mat)})
    return out


# ---------------------------------------------------------------- Gate 4
GATE4_FIXED_STEP_SECONDS = 1.0 / 240.0               # gate4.ts:11
GATE4_DEFAULT_GRAVITY = {"x": 0.0, "y": 90.0}        # gate4.ts:12
GATE4_PEBBLE_EDGE_LENGTH = 220 * 10 * 0.005 * 12     # gate4.ts:63-67 (132)
GATE4_PEBBLE_MATERIAL = {"lambda": 20.0, "mu": 10.0, "model": "neo-hookean"}  # gate4.ts:72-76


def gate4_scene(seed: int = GATE4_SCENE_SEED, mpm_grid: bool = True, liquid_area: float | None = None):
    """createMpmGate4Scene + createMpmGate4StepConfig (scenes/gate4.ts:184-357):
    2,000 water and 1,500 oil particles on 18-unit lattices, 50 three-particle
    neo-Hookean pebbles (bonded triangles of edge 132, each its own object)
    and a 6x4 fracture solid with 53 bonds (failure strain 0.12), velocities
    from the scene LCG (+-64 liquids, +-240 pebble vertices, one +-80 vector
    for the fracture solid).  mpm_grid=True is the MPM page's configuration
    (kernels/README.md:116-127): a 220 x 600 grid of cell size 10, liquid
    reference area 1,600 and the whole scene shifted 4,500 units in +y;
    False is the 220 x 150 code grid with liquid area 400.
    Returns (particles as MpmParticle dicts, solver config, gravity)."""
    area = liquid_area if liquid_area is not None else (1600.0 if mpm_grid else 400.0)
    shift = 4500.0 if mpm_grid else 0.0
    cells_y = 600 if mpm_grid else 150
    state = seed & MASK32

    def random() -> float:
        nonlocal state
        state = (state * LCG_A + LCG_C) & MASK32
        return state / 4294967296.0

    def velocity(vmax):
        return {"x": (random() * 2 - 1) * vmax, "y": (random() * 2 - 1) * vmax}

    zero = {"m00": 0.0, "m01": 0.0, "m10": 0.0, "m11": 0.0}
    ident = {"m00": 1.0, "m01": 0.0, "m10": 0.0, "m11": 1.0}
    particles = []
    for oid, start_y, rows, mat in ((1, 640.0, 40, GATE4_WATER), (2, 100.0, 30, GATE4_OIL)):  # gate4.ts:95-118
        for row in range(rows):
            for col in range(50):
                particles.append({"position": {"x": 100.0 + col * 18.0, "y": start_y + row * 18.0 + shift},
                                  "velocity": velocity(64.0), "C": dict(zero), "F": dict(ident), "objectId": oid,
                                  "plasticVolume": 1.0, "volume": area, "material": dict(mat)})
    columns, count = 8, 50
    rows = math.ceil(count / columns)
    edge = GATE4_PEBBLE_EDGE_LENGTH
    for k in range(count):
        row = k // columns
        in_row = min(columns, count - row * columns)
        col = k - row * columns + (columns - in_row) * 0.5
        cx = 700.0 + (col - (columns - 1) / 2) * 150.0
        cy = 569.0 + (row - (rows - 1) / 2) * 150.0 + shift
        verts = [(cx - edge * 0.5, cy - edge / (2 * math.sqrt(3))), (cx + edge * 0.5, cy - edge / (2 * math.sqrt(3))),
                 (cx, cy + edge / math.sqrt(3))]
        for v, (x, y) in enumerate(verts):
            particles.append({"position": {"x": x, "y": y}, "velocity": velocity(240.0), "C": dict(zero),
                              "F": dict(ident), "objectId": 3 + k, "bondedTriangleVertex": v, "plasticVolume": 1.0,
                              "volume": 100.0, "material": dict(GATE4_PEBBLE_MATERIAL)})
    fracture_start = len(particles)
    fvel = velocity(80.0)
    for row in range(4):
        for col in range(6):
            particles.append({"position": {"x": 1520.0 + (col - 2.5) * 20.0, "y": 1380.0 + (row - 1.5) * 20.0 + shift},
                              "velocity": dict(fvel), "C": dict(zero), "F": dict(ident), "objectId": 3 + count,
                              "plasticVolume": 1.0, "volume": 100.0, "material": dict(GATE4_PEBBLE_MATERIAL)})
    bonds = []
    for row in range(4):
        for col in range(6):
            a = fracture_start + row * 6 + col
            if col + 1 < 6:
                bonds.append({"firstParticle": a, "secondParticle": a + 1, "restLength": 20.0, "failureStrain": 0.12})
            if row + 1 < 4:
                bonds.append({"firstParticle": a, "secondParticle": a + 6, "restLength": 20.0, "failureStrain": 0.12})
            if row + 1 < 4 and col + 1 < 6:
                bonds.append({"firstParticle": a, "secondParticle": a + 7, "restLength": 20.0 * math.sqrt(2),
                              "failureStrain": 0.12})
    assert len(particles) == 3674 and len(bonds) == 53
    config = {"grid": {"cellsX": 220, "cellsY": cells_y, "cellSize": 10.0}, "particleMass": 4.0,
              "particleVolume": 100.0, "bondRestLength": edge, "bondReferenceDt": GATE4_FIXED_STEP_SECONDS,
              "material": dict(GATE4_PEBBLE_MATERIAL), "fractureBonds": bonds, "capacity": len(particles)}
    return particles, config, dict(GATE4_DEFAULT_GRAVITY)
