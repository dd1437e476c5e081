"""Boundary wire format: validation, float32 packing and the step-parameter block.

Restates src/lib/mpm/kernels/layout.ts for the host side of the C-ABI:

* ``validate_input`` — validateMpmMlsGpuInput (layout.ts:122-607): checks the
  configuration and every particle, rounds to float32 with the same range
  rules (finiteFloat32, layout.ts:1126-1138) and packs the 20-word records
  (layout.ts:385-437).
* ``pack_step_parameters`` — packMpmMlsGpuStepParameters (layout.ts:896-997):
  the 144-byte parameter block the native step reads.

Particles are accepted either as a list of ``MpmParticle``-shaped dicts (the
reference's object model, reference/solver.ts:29-40) or, for large scenes, as
a :class:`ParticleArrays` of numpy columns validated in bulk.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

PARTICLE_FLOATS = 20
PARTICLE_STRIDE_BYTES = 80
PREPARED_AFFINE_STRIDE_BYTES = 16
SIMULATION_PARAMS_BYTES = 144
COMPUTE_WORKGROUP_SIZE = 64
NUMERICAL_STATUS_WORDS = 8
MAX_FLOAT32 = 3.4028234663852886e38
UINT32_MAX = 0xFFFFFFFF
INT32_MAX = 0x7FFFFFFF
DEFAULT_BONDED_TRIANGLE_STIFFNESS = 0.35  # reference/solver.ts:277

MODEL_CODES = {"neo-hookean": 0, "snow": 1, "water": 2, "oil": 3, "powder": 4}
MODEL_NAMES = {v: k for k, v in MODEL_CODES.items()}
MODEL_REMOVED = 0xFFFFFFFF
SNOW_MIN_PLASTIC_VOLUME = 0.6
SNOW_MAX_PLASTIC_VOLUME = 20.0


class RangeError(ValueError):
    """The reference's RangeError: an argument outside its valid range."""


def f32(x) -> float:
    return float(np.float32(x))


def finite_float32(value, name: str, positive: bool = False) -> float:
    """finiteFloat32 (layout.ts:1126-1138)."""
    if not isinstance(value, (int, float, np.floating, np.integer)) or not math.isfinite(float(value)):
        raise RangeError(f"{name} is outside the supported float32 range")
    value = float(value)
    with np.errstate(over="ignore"):
        result = float(np.float32(value))
    if not math.isfinite(result) or abs(value) > MAX_FLOAT32:
        raise RangeError(f"{name} is outside the supported float32 range")
    if value != 0 and result == 0:
        raise RangeError(f"{name} underflows the supported float32 range")
    if positive and result <= 0:
        raise RangeError(f"{name} must remain positive in the supported float32 range")
    return result


def finite_number(value, name: str) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float, np.floating, np.integer)):
        raise TypeError(f"{name} must be a finite number")
    if not math.isfinite(float(value)):
        raise RangeError(f"{name} must be a finite number")
    return float(value)


def validate_material(value, name: str) -> dict:
    """validateMaterial / normalizeMpmConstitutiveModel (materials.ts:18-77)."""
    if not isinstance(value, dict):
        raise TypeError(f"{name} must be an object")
    model = value.get("model", "neo-hookean")
    if model not in MODEL_CODES:
        raise RangeError(f"{name}.model must be 'neo-hookean', 'snow', 'water', 'oil', or 'powder'")
    out = {"model": model, "lambda": finite_number(value.get("lambda"), f"{name}.lambda"),
           "mu": finite_number(value.get("mu"), f"{name}.mu")}
    if model in ("water", "oil"):
        out["density"] = finite_number(value.get("density"), f"{name}.density")
        if out["density"] <= 0:
            raise RangeError(f"{name}.density must be greater than zero")
        eos = value.get("equationOfState")
        if eos is not None:
            if not isinstance(eos, dict) or eos.get("model") != "tait":
                raise RangeError(f"{name}.equationOfState.model must be 'tait'.")
            exponent = eos.get("exponent")
            if not isinstance(exponent, (int, float)) or not math.isfinite(exponent) or exponent <= 0:
                raise RangeError(f"{name}.equationOfState.exponent must be positive and finite.")
            out["equationOfState"] = {"model": "tait", "exponent": float(exponent)}
    if model == "powder":
        out["frictionCoefficient"] = finite_number(value.get("frictionCoefficient"), f"{name}.frictionCoefficient")
    return out


def material_param(material: dict, damage: float = 0.0) -> float:
    """Word 18 of the record: Tait exponent / powder friction / NH damage."""
    model = material["model"]
    if model == "powder":
        return material["frictionCoefficient"]
    if model == "neo-hookean":
        return damage
    if model in ("water", "oil") and material.get("equationOfState", {}).get("model") == "tait":
        return material["equationOfState"]["exponent"]
    return 0.0


@dataclass
class ParticleArrays:
    """Column form of a particle set (float64 inputs; packed to float32)."""

    x: np.ndarray
    y: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    model: np.ndarray           # uint32 model codes
    lambda_: np.ndarray
    mu: np.ndarray
    density: np.ndarray
    volume: np.ndarray
    param: np.ndarray
    C: np.ndarray | None = None  # (n, 4) row-major m00 m01 m10 m11
    F: np.ndarray | None = None
    plastic_volume: np.ndarray | None = None
    object_id: np.ndarray | None = None
    bonded_vertex: np.ndarray | None = None  # bondedTriangleVertex (0, 1, 2) or -1

    def __len__(self) -> int:
        return int(np.asarray(self.x).shape[0])


@dataclass
class ValidatedInput:
    config: dict
    particle_count: int
    particle_capacity: int
    packed: np.ndarray            # (n, 20) uint32
    object_ids: np.ndarray        # (n,) uint32
    domain_width: float
    domain_height: float
    domain_width_f32: float
    domain_height_f32: float
    cell_size_f32: float
    particle_mass_f32: float
    particle_volume_f32: float
    bond_rest_length_f32: float
    bond_stiffness_f32: float
    bond_reference_dt_f32: float | None
    lambda_f32: float
    mu_f32: float
    wall_friction_f32: float
    gradient_scale: float
    gradient_scale_f32: float
    object_ranges: list = field(default_factory=list)
    bonded_body_count: int = 0
    fracture_bonds: np.ndarray | None = None     # (bonds, 6) uint32: the 24-byte records
    fracture_incidence_count: int = 0


def _validate_config(config: dict):
    if not isinstance(config, dict):
        raise TypeError("config must be an object")
    grid = config.get("grid")
    if not isinstance(grid, dict):
        raise TypeError("grid must be an object")
    material = validate_material(config.get("material"), "material")
    cells_x = finite_number(grid.get("cellsX"), "grid.cellsX")
    cells_y = finite_number(grid.get("cellsY"), "grid.cellsY")
    cell_size = finite_number(grid.get("cellSize"), "grid.cellSize")
    particle_mass = finite_number(config.get("particleMass"), "particleMass")
    particle_volume = finite_number(config.get("particleVolume"), "particleVolume")
    bond_rest = config.get("bondRestLength")
    bond_rest = None if bond_rest is None else finite_number(bond_rest, "bondRestLength")
    bond_stiffness = finite_number(config.get("bondStiffness", DEFAULT_BONDED_TRIANGLE_STIFFNESS), "bondStiffness")
    bond_ref = config.get("bondReferenceDt")
    bond_ref = None if bond_ref is None else finite_number(bond_ref, "bondReferenceDt")
    wall_friction = finite_number(config.get("wallFriction", 0), "wallFriction")
    if not float(cells_x).is_integer() or cells_x < 1 or cells_x + 3 > UINT32_MAX:
        raise RangeError("grid.cellsX must be a positive safe integer that fits the GPU layout")
    if not float(cells_y).is_integer() or cells_y < 1 or cells_y + 3 > UINT32_MAX:
        raise RangeError("grid.cellsY must be a positive safe integer that fits the GPU layout")
    if cell_size <= 0:
        raise RangeError("grid.cellSize must be greater than zero")
    if particle_mass <= 0:
        raise RangeError("particleMass must be greater than zero")
    if particle_volume <= 0:
        raise RangeError("particleVolume must be greater than zero")
    if bond_rest is not None and bond_rest <= 0:
        raise RangeError("bondRestLength must be greater than zero")
    if bond_stiffness <= 0 or bond_stiffness > 1:
        raise RangeError("bondStiffness must be greater than zero and at most one")
    if bond_ref is not None and bond_ref <= 0:
        raise RangeError("bondReferenceDt must be greater than zero")
    if wall_friction < 0 or wall_friction > 1:
        raise RangeError("wallFriction must be between zero and one")
    if material["lambda"] < 0:
        raise RangeError("material.lambda must be non-negative")
    if material["mu"] <= 0 and material["model"] not in ("water", "oil"):
        raise RangeError("material.mu must be greater than zero")
    cells_x, cells_y = int(cells_x), int(cells_y)
    return (material, cells_x, cells_y, cell_size, particle_mass, particle_volume, bond_rest, bond_stiffness,
            bond_ref, wall_friction)


def particle_arrays_from_dicts(particles, default_material: dict, particle_volume_f32: float) -> ParticleArrays:
    """Per-particle checks of validateMpmMlsGpuInput (layout.ts:300-437) on
    MpmParticle-shaped dicts, gathered into columns."""
    if not isinstance(particles, (list, tuple)):
        raise TypeError("particles must be an array")
    n = len(particles)
    cols = {k: np.zeros(n) for k in ("x", "y", "vx", "vy", "lambda_", "mu", "density", "volume", "param", "pv")}
    C = np.zeros((n, 4))
    F = np.zeros((n, 4))
    model = np.zeros(n, np.uint32)
    oid = np.zeros(n, np.uint32)
    vertex = np.full(n, -1, np.int64)
    for i, p in enumerate(particles):
        if not isinstance(p, dict):
            raise TypeError(f"particles[{i}] must be an object")
        name = f"particles[{i}]"
        mat = default_material if p.get("material") is None else validate_material(p["material"], f"{name}.material")
        bv = p.get("bondedTriangleVertex")
        if bv is not None and (isinstance(bv, bool) or bv not in (0, 1, 2)):
            raise RangeError(f"{name}.bondedTriangleVertex must be 0, 1, or 2")
        vertex[i] = -1 if bv is None else int(bv)
        pos, vel = p.get("position"), p.get("velocity")
        for vec, nm in ((pos, "position"), (vel, "velocity")):
            if not isinstance(vec, dict):
                raise TypeError(f"{name}.{nm} must be a vector")
        for mat_name in ("C", "F"):
            if not isinstance(p.get(mat_name), dict):
                raise TypeError(f"{name}.{mat_name} must be a matrix")
        oid_v = p.get("objectId", 0)
        if not float(oid_v).is_integer() or oid_v < 0 or oid_v > UINT32_MAX:
            raise RangeError(f"{name}.objectId must be an unsigned 32-bit integer")
        cols["x"][i] = finite_number(pos.get("x"), f"{name}.position.x")
        cols["y"][i] = finite_number(pos.get("y"), f"{name}.position.y")
        cols["vx"][i] = finite_number(vel.get("x"), f"{name}.velocity.x")
        cols["vy"][i] = finite_number(vel.get("y"), f"{name}.velocity.y")
        for k, key in enumerate(("m00", "m01", "m10", "m11")):
            C[i, k] = finite_number(p["C"].get(key), f"{name}.C.{key}")
            F[i, k] = finite_number(p["F"].get(key), f"{name}.F.{key}")
        damage = p.get("damage")
        if damage is not None:
            damage = finite_number(damage, f"{name}.damage")
            if damage < 0 or damage > 1:
                raise RangeError(f"{name}.damage must be between zero and one")
            if mat["model"] != "neo-hookean":
                raise RangeError(f"{name}.damage is only valid for neo-hookean solids")
        pv = p.get("plasticVolume")
        cols["pv"][i] = 1.0 if pv is None else finite_number(pv, f"{name}.plasticVolume")
        vol = p.get("volume")
        cols["volume"][i] = particle_volume_f32 if vol is None else finite_number(vol, f"{name}.volume")
        if cols["volume"][i] <= 0:
            raise RangeError(f"{name}.volume must be greater than zero")
        model[i] = MODEL_CODES[mat["model"]]
        cols["lambda_"][i] = mat["lambda"]
        cols["mu"][i] = mat["mu"]
        cols["density"][i] = mat.get("density", 0.0)
        cols["param"][i] = material_param(mat, 0.0 if damage is None else damage)
        oid[i] = int(oid_v)
    return ParticleArrays(cols["x"], cols["y"], cols["vx"], cols["vy"], model, cols["lambda_"], cols["mu"],
                          cols["density"], cols["volume"], cols["param"], C, F, cols["pv"], oid,
                          vertex if np.any(vertex >= 0) else None)


def _f32_checked(values: np.ndarray, name: str, positive=False) -> np.ndarray:
    v = np.asarray(values, np.float64)
    if not np.all(np.isfinite(v)) or np.any(np.abs(v) > MAX_FLOAT32):
        raise RangeError(f"{name} is outside the supported float32 range")
    r = v.astype(np.float32)
    if np.any((v != 0) & (r == 0)):
        raise RangeError(f"{name} underflows the supported float32 range")
    if positive and np.any(r <= 0):
        raise RangeError(f"{name} must remain positive in the supported float32 range")
    return r


def pack_arrays(arr: ParticleArrays, domain_w: float, domain_h: float, domain_w32: float, domain_h32: float,
                particle_mass_f32: float) -> np.ndarray:
    """Vectorised packing to the 20-word records (layout.ts:385-437)."""
    n = len(arr)
    out = np.zeros((n, 20), np.uint32)
    fv = out.view(np.float32)
    C = np.zeros((n, 4)) if arr.C is None else np.asarray(arr.C, np.float64).reshape(n, 4)
    F = np.tile([1.0, 0.0, 0.0, 1.0], (n, 1)) if arr.F is None else np.asarray(arr.F, np.float64).reshape(n, 4)
    pv = np.ones(n) if arr.plastic_volume is None else np.asarray(arr.plastic_volume, np.float64)
    model = np.asarray(arr.model, np.uint32)
    liquid = (model == 2) | (model == 3)
    x, y = np.asarray(arr.x, np.float64), np.asarray(arr.y, np.float64)
    if np.any(model > 4):
        raise RangeError("particle material model code must be 0..4")
    if np.any((pv < SNOW_MIN_PLASTIC_VOLUME) | (pv > SNOW_MAX_PLASTIC_VOLUME)):
        raise RangeError(f"particles plasticVolume must be between {SNOW_MIN_PLASTIC_VOLUME} and {SNOW_MAX_PLASTIC_VOLUME}")
    if np.any((x < 0) | (x > domain_w) | (y < 0) | (y > domain_h)):
        raise RangeError("particles position must be inside the grid domain")
    det = F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]
    if not np.all(np.isfinite(det)) or np.any(det <= 0):
        raise RangeError("particles F must have a finite positive determinant")
    fv[:, 0] = _f32_checked(x, "position.x")
    fv[:, 1] = _f32_checked(y, "position.y")
    fv[:, 2] = _f32_checked(arr.vx, "velocity.x")
    fv[:, 3] = _f32_checked(arr.vy, "velocity.y")
    fv[:, 4:8] = _f32_checked(C, "C").reshape(n, 4)
    fv[:, 8:12] = _f32_checked(F, "F").reshape(n, 4)
    fv[:, 12] = _f32_checked(pv, "plasticVolume", positive=True)
    out[:, 13] = model
    fv[:, 14] = _f32_checked(arr.lambda_, "material.lambda")
    mu = np.asarray(arr.mu, np.float64)
    if np.any(~liquid & (mu <= 0)):
        raise RangeError("material.mu must be greater than zero")
    fv[:, 15] = _f32_checked(mu, "material.mu")
    dens = np.where(liquid, np.asarray(arr.density, np.float64), 0.0)
    fv[:, 16] = _f32_checked(dens, "material.density")
    if np.any(liquid & (fv[:, 16] <= 0)):
        raise RangeError("material.density must remain positive in the supported float32 range")
    fv[:, 17] = _f32_checked(arr.volume, "volume", positive=True)
    fv[:, 18] = _f32_checked(arr.param, "material parameter")
    if arr.bonded_vertex is not None:
        bv = np.asarray(arr.bonded_vertex, np.int64)
        out[:, 19] = np.where(bv >= 0, bv + 1, 0).astype(np.uint32)
    packed_mass = np.where(liquid, fv[:, 16] * fv[:, 17], np.float32(particle_mass_f32))
    if not np.all(np.isfinite(packed_mass)) or np.any(packed_mass <= 0):
        raise RangeError("particle mass is outside the supported float32 range")
    with np.errstate(over="ignore", invalid="ignore"):
        pdet = (fv[:, 8] * fv[:, 11]) - (fv[:, 9] * fv[:, 10])
    if not np.all(np.isfinite(pdet)) or np.any(pdet <= 0):
        raise RangeError("particles F must retain a finite positive determinant in float32")
    if np.any((fv[:, 0] < 0) | (fv[:, 0] > np.float32(domain_w32)) | (fv[:, 1] < 0) | (fv[:, 1] > np.float32(domain_h32))):
        raise RangeError("particles position is outside the float32 GPU grid domain")
    return out


def validate_input(particles, config: dict) -> ValidatedInput:
    """validateMpmMlsGpuInput (layout.ts:122-607)."""
    (material, cells_x, cells_y, cell_size, particle_mass, particle_volume, bond_rest, bond_stiffness, bond_ref,
     wall_friction) = _validate_config(config)
    domain_w = cells_x * cell_size
    domain_h = cells_y * cell_size
    gradient_scale = 4 / (cell_size * cell_size)
    if not math.isfinite(domain_w) or not math.isfinite(domain_h):
        raise RangeError("grid dimensions produce a non-finite domain")
    if not math.isfinite(gradient_scale) or cell_size * cell_size <= 0:
        raise RangeError("grid.cellSize is outside the supported numeric range")
    if not math.isfinite(particle_mass / particle_volume) or not math.isfinite(material["lambda"] + 2 * material["mu"]):
        raise RangeError("mass and material parameters are outside the supported numeric range")
    cell_size_f32 = finite_float32(cell_size, "grid.cellSize", True)
    mass_f32 = finite_float32(particle_mass, "particleMass", True)
    volume_f32 = finite_float32(particle_volume, "particleVolume", True)
    bond_rest_f32 = 0.0 if bond_rest is None else finite_float32(bond_rest, "bondRestLength", True)
    bond_stiffness_f32 = finite_float32(bond_stiffness, "bondStiffness", True)
    bond_ref_f32 = None if bond_ref is None else finite_float32(bond_ref, "bondReferenceDt", True)
    lambda_f32 = finite_float32(material["lambda"], "material.lambda")
    mu_f32 = finite_float32(material["mu"], "material.mu", material["model"] not in ("water", "oil"))
    wall_f32 = finite_float32(wall_friction, "wallFriction")
    grad_f32 = finite_float32(gradient_scale, "grid gradient scale", True)
    dw32 = finite_float32(f32(cells_x * np.float32(cell_size_f32)), "GPU grid width", True)
    dh32 = finite_float32(f32(cells_y * np.float32(cell_size_f32)), "GPU grid height", True)

    if isinstance(particles, ParticleArrays):
        arr = particles
    else:
        arr = particle_arrays_from_dicts(particles, material, volume_f32)
    n = len(arr)
    capacity = config.get("capacity", n)
    if isinstance(capacity, bool) or not float(capacity).is_integer() or capacity < 1 and n > 0:
        raise RangeError("capacity must be a positive safe integer")
    capacity = int(capacity)
    if capacity < n or capacity > UINT32_MAX:
        raise RangeError("capacity must be at least the initial particle count and fit the GPU layout")
    if capacity * 9 > INT32_MAX:
        raise RangeError("particle stencil entries exceed the linked-list index range")
    packed = pack_arrays(arr, domain_w, domain_h, dw32, dh32, mass_f32)
    oid = np.zeros(n, np.uint32) if arr.object_id is None else np.asarray(arr.object_id, np.uint32)
    ranges = []
    if n:
        change = np.flatnonzero(np.diff(oid.astype(np.int64)) != 0) + 1
        starts = np.concatenate(([0], change))
        ends = np.concatenate((change, [n]))
        seen = set()
        for s_, e_ in zip(starts, ends):
            o = int(oid[s_])
            if o in seen:
                raise RangeError(f"particles with objectId {o} must be contiguous for GPU cloning")
            seen.add(o)
            ranges.append({"objectId": o, "start": int(s_), "count": int(e_ - s_)})
    # bonded triangle bodies per object range (layout.ts:456-494)
    vertex = packed[:, 19].astype(np.int64) - 1
    bodies = 0
    for r in ranges:
        v = vertex[r["start"]:r["start"] + r["count"]]
        if not np.any(v >= 0):
            continue
        if bond_rest is None:
            raise RangeError(f"object {r['objectId']} uses bonded triangles but bondRestLength is missing")
        if r["count"] % 3 != 0:
            raise RangeError(f"bonded object {r['objectId']} must contain complete three-particle triangles")
        if not np.array_equal(v, np.arange(r["count"]) % 3):
            raise RangeError(f"bonded object {r['objectId']} must store contiguous vertices 0, 1, and 2")
        m = packed[r["start"]:r["start"] + r["count"], 13]
        if np.any((m == 2) | (m == 3)):
            raise RangeError("bonded triangle vertices must use a solid material")
        bodies += r["count"] // 3
    if bodies > 0 and bond_rest_f32 <= 0:
        raise RangeError("bondRestLength must remain positive in float32")
    fracture = config.get("fractureBonds")
    bonds_packed, incidence = None, 0
    if fracture is not None:
        bonds_packed = pack_fracture_bonds(fracture, packed, oid)
        incidence = 2 * bonds_packed.shape[0]
    cfg = {"grid": {"cellsX": cells_x, "cellsY": cells_y, "cellSize": cell_size}, "particleMass": particle_mass,
           "particleVolume": particle_volume, "bondStiffness": bond_stiffness, "material": material,
           "wallFriction": wall_friction, "capacity": capacity}
    if fracture is not None:
        cfg["fractureBonds"] = fracture
    if bond_rest is not None:
        cfg["bondRestLength"] = bond_rest
    if bond_ref is not None:
        cfg["bondReferenceDt"] = bond_ref
    return ValidatedInput(cfg, n, capacity, packed, oid, domain_w, domain_h, dw32, dh32, cell_size_f32, mass_f32,
                          volume_f32, bond_rest_f32, bond_stiffness_f32, bond_ref_f32, lambda_f32, mu_f32, wall_f32,
                          gradient_scale, grad_f32, ranges, bodies, bonds_packed, incidence)


def validated_from_records(records: np.ndarray, object_ids: np.ndarray, config: dict) -> ValidatedInput:
    """A ValidatedInput holding exactly the given 20-word records (a restored
    snapshot): the records go through validate_input as particle columns,
    and must pack back to themselves word for word.  Removed tombstones
    (word 13 = 0xFFFFFFFF, left by outlets) are validated as the liquids they
    were and keep their place and their tombstone word."""
    rec = np.ascontiguousarray(records, np.uint32).reshape(-1, 20)
    n = rec.shape[0]
    fv = rec.view(np.float32).astype(np.float64)
    removed = rec[:, 13] == MODEL_REMOVED
    model = np.where(removed, MODEL_CODES["water"], rec[:, 13]).astype(np.uint32)
    vertex = rec[:, 19].astype(np.int64) - 1
    arr = ParticleArrays(x=fv[:, 0], y=fv[:, 1], vx=fv[:, 2], vy=fv[:, 3], model=model, lambda_=fv[:, 14],
                         mu=fv[:, 15], density=fv[:, 16], volume=fv[:, 17], param=fv[:, 18], C=fv[:, 4:8],
                         F=fv[:, 8:12], plastic_volume=fv[:, 12], object_id=np.asarray(object_ids, np.uint32),
                         bonded_vertex=vertex if np.any(vertex >= 0) else None)
    state = validate_input(arr, config)
    state.packed[removed, 13] = MODEL_REMOVED
    if state.packed.shape != rec.shape or not np.array_equal(state.packed, rec):
        bad = int(np.flatnonzero(np.any(state.packed != rec, axis=1))[0]) if state.packed.shape == rec.shape else -1
        raise RangeError(f"snapshot record {bad} does not pack back to itself")
    state.particle_count = n
    return state


def pack_fracture_bonds(bonds, packed: np.ndarray, oid: np.ndarray) -> np.ndarray:
    """validateMpmSolidBondGraph (reference/solid-bonds.ts:11-66) and the
    24-byte bond records (layout.ts:497-548): u32 first, u32 second,
    f32 restLength, f32 failureStrain, f32 damage (1 when broken), u32 broken."""
    if not isinstance(bonds, (list, tuple)):
        raise TypeError("fractureBonds must be an array")
    n = packed.shape[0]
    out = np.zeros((len(bonds), 6), np.uint32)
    fv = out.view(np.float32)
    seen = set()
    for k, b in enumerate(bonds):
        if not isinstance(b, dict):
            raise TypeError(f"fractureBonds[{k}] must be an object")
        ends = []
        for key in ("firstParticle", "secondParticle"):
            v = b.get(key)
            if isinstance(v, bool) or not isinstance(v, int) or v < 0 or v >= n:
                raise RangeError(f"fractureBonds[{k}].{key} must identify an existing particle")
            ends.append(v)
        a, c = ends
        if a == c:
            raise RangeError(f"fractureBonds[{k}] must connect two distinct particles")
        rest = b.get("restLength")
        if not isinstance(rest, (int, float)) or not math.isfinite(rest) or rest <= 0:
            raise RangeError(f"fractureBonds[{k}].restLength must be positive and finite")
        strain = b.get("failureStrain")
        if not isinstance(strain, (int, float)) or not math.isfinite(strain) or strain <= 0:
            raise RangeError(f"fractureBonds[{k}].failureStrain must be positive and finite")
        damage = b.get("damage", 0.0)
        if not isinstance(damage, (int, float)) or not math.isfinite(damage) or damage < 0 or damage > 1:
            raise RangeError(f"fractureBonds[{k}].damage must be between zero and one")
        broken = b.get("broken")
        if broken is not None and not isinstance(broken, bool):
            raise TypeError(f"fractureBonds[{k}].broken must be a boolean")
        if oid[a] != oid[c]:
            raise RangeError(f"fractureBonds[{k}] must connect particles of one object")
        if packed[a, 19] != 0 or packed[c, 19] != 0:
            raise RangeError(f"fractureBonds[{k}] cannot reuse bonded triangle vertices")
        if packed[a, 13] != 0 or packed[c, 13] != 0:
            raise RangeError(f"fractureBonds[{k}] endpoints must be neo-hookean solids")
        pair = (min(a, c), max(a, c))
        if pair in seen:
            raise RangeError(f"fractureBonds[{k}] duplicates an existing particle pair")
        seen.add(pair)
        out[k, 0], out[k, 1] = a, c
        fv[k, 2] = finite_float32(rest, f"fractureBonds[{k}].restLength", True)
        fv[k, 3] = finite_float32(strain, f"fractureBonds[{k}].failureStrain", True)
        fv[k, 4] = finite_float32(1.0 if broken else damage, f"fractureBonds[{k}].damage")
        out[k, 5] = 1 if broken else 0
    return out


def unpack_fracture_bonds(records: np.ndarray) -> list[dict]:
    """24-byte bond records back to MpmSolidBond dicts."""
    rec = np.asarray(records, np.uint32).reshape(-1, 6)
    fv = rec.view(np.float32)
    return [{"firstParticle": int(r[0]), "secondParticle": int(r[1]), "restLength": float(f[2]),
             "failureStrain": float(f[3]), "damage": float(f[4]), "broken": bool(r[5])} for r, f in zip(rec, fv)]


def pack_step_parameters(state: ValidatedInput, dt: float, gravity: dict, particle_count: int | None = None,
                         bonded_body_count: int | None = None, cfl_guard: dict | None = None,
                         bond_reference_dt: float | None = None) -> bytes:
    """packMpmMlsGpuStepParameters (layout.ts:896-997): 144 bytes."""
    particle_count = state.particle_count if particle_count is None else particle_count
    if bonded_body_count is None:
        bonded_body_count = state.bonded_body_count
    if not isinstance(particle_count, int) or particle_count < 0 or particle_count > state.particle_capacity:
        raise RangeError("particleCount must be an integer within the allocated GPU capacity")
    if particle_count * 9 > INT32_MAX:
        raise RangeError("active particle stencil entries exceed the linked-list index range")
    step_dt = finite_number(dt, "dt")
    if step_dt <= 0:
        raise RangeError("dt must be greater than zero")
    dt32 = finite_float32(step_dt, "dt", True)
    if not isinstance(gravity, dict):
        raise TypeError("gravity must be a vector")
    gx = finite_float32(finite_number(gravity.get("x"), "gravity.x"), "gravity.x")
    gy = finite_float32(finite_number(gravity.get("y"), "gravity.y"), "gravity.y")
    wave = 0.0 if cfl_guard is None else finite_float32(cfl_guard["waveSpeedBound"], "cflGuard.waveSpeedBound")
    vel = 0.0 if cfl_guard is None else finite_float32(cfl_guard["velocityBound"], "cflGuard.velocityBound")
    ref = bond_reference_dt if bond_reference_dt is not None else (
        state.bond_reference_dt_f32 if state.bond_reference_dt_f32 is not None else dt32)
    ref32 = finite_float32(ref, "bondReferenceDt", True)
    stress_scale = -step_dt * state.config["particleVolume"] * state.gradient_scale
    if not math.isfinite(stress_scale):
        raise RangeError("dt, particleVolume, and grid.cellSize produce a non-finite stress scale")
    ss32 = finite_float32(stress_scale, "stress scale")
    cells_x = state.config["grid"]["cellsX"]
    cells_y = state.config["grid"]["cellsY"]
    nw, nh = cells_x + 3, cells_y + 3
    ints = struct.pack("<8I", cells_x, cells_y, nw, nh, particle_count, nw * nh, particle_count * 9,
                       bonded_body_count)
    cs = np.float32(state.cell_size_f32)
    floats = np.array([
        state.cell_size_f32, state.particle_mass_f32, state.particle_volume_f32, dt32,
        state.lambda_f32, state.mu_f32, ss32, state.gradient_scale_f32,
        gx, gy, state.domain_width_f32, state.domain_height_f32,
        state.bond_rest_length_f32, state.bond_stiffness_f32, float(cs * np.float32(2)),
        math.sqrt(state.particle_volume_f32 / math.pi) + float(cs) * 0.75,
        wave, vel, 0.0 if cfl_guard is None else 1.0, ref32,
    ], np.float32).tobytes()
    nbonds = 0 if state.fracture_bonds is None else state.fracture_bonds.shape[0]
    fracture = struct.pack("<4I", state.particle_count if nbonds else 0, nbonds, state.fracture_incidence_count,
                           1 if 0 < bonded_body_count <= 512 else 0)
    wall = np.array([state.wall_friction_f32, 0, 0, 0], np.float32).tobytes()
    out = ints + floats + fracture + wall
    assert len(out) == SIMULATION_PARAMS_BYTES
    return out


def unpack_particles(records: np.ndarray, object_ids: np.ndarray | None = None) -> list[dict]:
    """Decode 20-word records back to MpmParticle-shaped dicts (unpackParticles,
    solver.ts:3148-3226): removed tombstones are skipped."""
    rec = np.asarray(records, np.uint32).reshape(-1, 20)
    fv = rec.view(np.float32).astype(np.float64)
    out = []
    for i in range(rec.shape[0]):
        model = int(rec[i, 13])
        if model == MODEL_REMOVED:
            continue
        name = MODEL_NAMES.get(model, "removed")
        mat = {"lambda": fv[i, 14], "mu": fv[i, 15], "model": name}
        p = {
            "position": {"x": fv[i, 0], "y": fv[i, 1]},
            "velocity": {"x": fv[i, 2], "y": fv[i, 3]},
            "C": {"m00": fv[i, 4], "m01": fv[i, 5], "m10": fv[i, 6], "m11": fv[i, 7]},
            "F": {"m00": fv[i, 8], "m01": fv[i, 9], "m10": fv[i, 10], "m11": fv[i, 11]},
            "plasticVolume": fv[i, 12],
            "volume": fv[i, 17],
            "objectId": 0 if object_ids is None else int(object_ids[i]),
            "material": mat,
        }
        if name in ("water", "oil"):
            mat["density"] = fv[i, 16]
            if fv[i, 18] > 0:
                mat["equationOfState"] = {"model": "tait", "exponent": fv[i, 18]}
        elif name == "powder":
            mat["frictionCoefficient"] = fv[i, 18]
        elif name == "neo-hookean":
            p["damage"] = fv[i, 18]
        if rec[i, 19]:
            p["bondedTriangleVertex"] = int(rec[i, 19]) - 1
        out.append(p)
    return out
