"""Autotuner of the step kernels' launch shape and bucket capacity.

The reference has no counterpart: its WGSL kernels have fixed workgroup sizes
(src/lib/mpm/kernels/shaders.ts).  Here the step kernels are compiled for
several launch shapes (``mpm_tuning`` in include/mpm355.h):

* P2G tile side (28, 20 or 12 grid nodes): larger tiles re-read fewer region
  buckets, smaller ones keep more blocks in flight per CU;
* G2P tile side (8 or 4 buckets) and threads per G2P block, which together set
  the particles each wavefront processes;
* bucket capacity (slots per 4x4-cell bucket, the "bin size").

A step's results do not depend on any of them, so the choice is pure
performance, and it depends on the GPU and on the scene (grid size, particle
count and density, phases present, P2G mode).  ``autotune`` times every
candidate on a scratch copy of the caller's scene (a context of its own
built from the same particles and configuration: the caller's solver is never
stepped), with the library's per-kernel HIP-event timing:

1. every P2G tile with the default G2P shape, ranked by the P2G kernel's mean
   time;
2. every G2P shape with the chosen P2G tile, ranked by the G2P kernel's mean
   time (the two kernels share no knob);
3. bucket capacities around the derived one, ranked by the whole step.

``autotune_batch`` does the same for a batch of scenes (one launch covers
every scene), ranking each shape by the batch's microstep time.

Each candidate is timed twice, interleaved, and keeps its faster pass.  The
result is cached as JSON keyed by (GPU architecture and compute units, scene
class); ``tuning_for`` returns a cached entry or tunes and stores one.
"""
from __future__ import annotations

import datetime as _dt
import json
import math
import os
import pathlib
import time

import numpy as np

from . import cfl as cflmod
from . import layout
from .native import MPM_P2G_FIXED_POINT, MPM_P2G_GATHER, MpmTuning, device_arch

CACHE_VERSION = 1
# the library's default shape on a grid with at least one 8x8-bucket G2P tile
# per compute unit (smaller grids default to 4x4-bucket tiles)
DEFAULT_TUNING = {"p2g_tile": 20, "g2p_tile": 8, "g2p_threads": 256, "bucket_capacity": 0}
BODIES_TUNING = {"p2g_tile": 28, "g2p_tile": 8, "g2p_threads": 256, "bucket_capacity": 0}
BODIES_G2P_TILES = (8, 4)  # scenes with bodies or bonds: 28-node P2G, 256-thread G2P of either tile
P2G_TILES = (28, 20, 12)
G2P_SHAPES = ((8, 256), (8, 512), (8, 128), (4, 256), (4, 128))
# particles one float-gather chunk stages for each P2G tile side: a bucket
# must fit one chunk, so a tile whose chunk is below the bucket capacity
# runs as the 28-node tile
P2G_CHUNK = {28: 1408, 20: 960, 12: 448}
MIN_BUCKET_CAPACITY, MAX_BUCKET_CAPACITY = 32, 1024
CAPACITY_FACTORS = (1.0, 0.75, 1.5)  # of the capacity the library derived
MACRO_DT = 1.0 / 240.0


def shape_supported(p2g_tile: int, g2p_tile: int, g2p_threads: int) -> bool:
    """The launch shapes libmpm355 compiles (shape_supported, mpm_internal.h)."""
    return p2g_tile in P2G_TILES and (g2p_tile, g2p_threads) in G2P_SHAPES


def default_cache_path() -> pathlib.Path:
    env = os.environ.get("MPM355_TUNING_CACHE")
    if env:
        return pathlib.Path(env)
    return pathlib.Path.home() / ".cache" / "mpm355" / "tuning.json"


def default_tuning(config: dict, device: int = 0, bodies: bool = False) -> dict:
    """The shape the library runs a scene with when it is given none
    (include/mpm355.h, "kernel tuning"): DEFAULT_TUNING, with 4x4-bucket G2P
    tiles when the grid has fewer 8x8-bucket tiles than the device has
    compute units.  Scenes with bonded bodies or fracture bonds (``bodies``)
    run BODIES_TUNING's 28-node P2G tile and 256-thread G2P blocks, with the
    G2P tile side from the same grid-size rule."""
    bw = int(config["grid"]["cellsX"]) // 4 + 1
    bh = int(config["grid"]["cellsY"]) // 4 + 1
    tiles8 = -(-bw // 8) * -(-bh // 8)
    _, cus = device_arch(max(device, 0))
    return dict(BODIES_TUNING if bodies else DEFAULT_TUNING, g2p_tile=8 if tiles8 >= cus else 4)


def device_key(device: int = 0) -> str:
    arch, cus = device_arch(max(device, 0))
    return f"{arch}/{cus}cu"


def _log2_class(x: float) -> int:
    return int(round(math.log2(x))) if x > 0 else -1


def scene_class(state: layout.ValidatedInput, p2g_mode: int) -> str:
    """Scene class of a validated scene: what the kernels' speed depends on.

    P2G mode; grid nodes and live particles (nearest power of two); the phases
    present (solid / water / oil), which pick the kernel variant and the
    channels it moves; whether bonded bodies or fracture bonds are present
    (such scenes run the default shape); and the mean particles per occupied
    4x4-cell bucket (nearest power of two), which sets the work per tile."""
    cfg = state.config["grid"]
    cells_x, cells_y = int(cfg["cellsX"]), int(cfg["cellsY"])
    nodes = (cells_x + 3) * (cells_y + 3)
    packed = state.packed
    model = packed[:, 13] if len(packed) else np.zeros(0, np.uint32)
    live = model != layout.MODEL_REMOVED
    models = set(int(m) for m in np.unique(model[live]))
    phases = []
    if models & {0, 1, 4}:
        phases.append("solid")
    if 2 in models:
        phases.append("water")
    if 3 in models:
        phases.append("oil")
    n = int(live.sum())
    density = 0.0
    if n:
        cell = float(state.cell_size_f32)
        x = packed[live, 0].view(np.float32).astype(np.float64) / cell
        y = packed[live, 1].view(np.float32).astype(np.float64) / cell
        bx = (np.floor(x - 0.5).astype(np.int64) + 1) >> 2
        by = (np.floor(y - 0.5).astype(np.int64) + 1) >> 2
        occupied = np.unique(by * (cells_x // 4 + 2) + bx).size
        density = n / max(occupied, 1)
    bodies = state.bonded_body_count > 0 or state.fracture_bonds is not None
    mode = "fixed" if p2g_mode == MPM_P2G_FIXED_POINT else "gather"
    return (f"{mode}|nodes=2^{_log2_class(nodes)}|particles=2^{_log2_class(n)}|phases={'+'.join(phases) or 'none'}"
            f"|bodies={int(bodies)}|per_bucket=2^{_log2_class(density)}")


# ------------------------------------------------------------------ cache
def load_cache(path: pathlib.Path | str | None = None) -> dict:
    """The cache file's entries ({} when absent, unreadable or of another
    version)."""
    p = pathlib.Path(path) if path is not None else default_cache_path()
    try:
        data = json.loads(p.read_text())
    except (OSError, ValueError):
        return {}
    if not isinstance(data, dict) or data.get("version") != CACHE_VERSION or not isinstance(data.get("entries"), dict):
        return {}
    return data["entries"]


def store_entry(device: str, klass: str, entry: dict, path: pathlib.Path | str | None = None) -> None:
    """Merge one entry into the cache file (written atomically)."""
    p = pathlib.Path(path) if path is not None else default_cache_path()
    entries = load_cache(p)
    entries.setdefault(device, {})[klass] = entry
    p.parent.mkdir(parents=True, exist_ok=True)
    tmp = p.with_name(p.name + f".{os.getpid()}.tmp")
    tmp.write_text(json.dumps({"version": CACHE_VERSION, "entries": entries}, indent=1, sort_keys=True))
    os.replace(tmp, p)


def lookup(device: str, klass: str, path: pathlib.Path | str | None = None) -> dict | None:
    entry = load_cache(path).get(device, {}).get(klass)
    if not isinstance(entry, dict) or not isinstance(entry.get("tuning"), dict):
        return None
    t = entry["tuning"]
    if not shape_supported(int(t.get("p2g_tile", 0)), int(t.get("g2p_tile", 0)), int(t.get("g2p_threads", 0))):
        return None
    return entry


# ----------------------------------------------------------------- tuning
def _tuning(p2g_tile, g2p_tile, g2p_threads, bucket_capacity=0) -> MpmTuning:
    return MpmTuning(p2g_tile, g2p_tile, g2p_threads, bucket_capacity)


class _Trial:
    """A scratch context of the caller's scene and the step parameters the
    candidates are timed with."""

    def __init__(self, particles, config, p2g_mode, dt, gravity, device):
        from .solver import MpmMlsGpuSolver
        self.solver = MpmMlsGpuSolver(particles, config, None, p2g_mode, device=device)
        self.ctx = self.solver.ctx
        if dt is None:
            dt = _default_dt(self.solver, gravity)
        self.params = layout.pack_step_parameters(self.solver.state, dt, gravity,
                                                  self.solver.allocated_particle_count, None, None)
        self.p2g_kernel = "p2g_grid_fixed" if p2g_mode == MPM_P2G_FIXED_POINT else "p2g_grid"

    def run(self, tuning: MpmTuning, warmup: int, steps: int) -> dict:
        ctx = self.ctx
        ctx.set_tuning(tuning)
        ctx.step_batch(self.params, warmup, False)
        ctx.read_compact()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        ctx.step_batch(self.params, steps, False)
        _, _, status = ctx.read_compact()
        wall = time.perf_counter() - t0
        p2g, _ = ctx.read_timing(self.p2g_kernel)
        g2p, _ = ctx.read_timing("g2p")
        ctx.set_timing(False)
        if status[0] != 0:
            raise RuntimeError(f"autotune step failed with status {list(status)}")
        eff = ctx.get_tuning()
        return {"p2g_us": p2g, "g2p_us": g2p, "step_us": 1e6 * wall / steps, "effective": eff.as_dict()}

    def close(self):
        self.solver.dispose()


def _default_dt(solver, gravity) -> float:
    """The CFL microstep of one 1/240 s macrostep of the scene."""
    b = solver.cfl_bounds
    cell = float(solver.state.cell_size_f32)
    if b is None:
        return 1e-4
    g = math.hypot(float(gravity["x"]), float(gravity["y"]))
    bound = cflmod.calculate_velocity_bound(b.velocityBound, MACRO_DT, g)
    return cflmod.calculate_microsteps(MACRO_DT, cell, b.waveSpeedBound, bound, 1024).microDt


def _best(results: dict, key: str):
    return min(results, key=lambda k: results[k][key])


def autotune(particles, config: dict, p2g_mode: int = MPM_P2G_GATHER, *, dt: float | None = None,
             gravity: dict | None = None, device: int = 0, warmup: int = 3, steps: int = 12,
             p2g_tiles=P2G_TILES, g2p_shapes=G2P_SHAPES, capacity_factors=CAPACITY_FACTORS) -> dict:
    """Time the candidates on a scratch copy of the scene; returns
    {"tuning": {...}, "class": ..., "device": ..., "candidates": {...}}."""
    gravity = gravity if gravity is not None else {"x": 0.0, "y": -9.81}
    trial = _Trial(particles, config, p2g_mode, dt, gravity, device)
    try:
        state = trial.solver.state
        klass = scene_class(state, p2g_mode)
        dev = device_key(device)
        derived_cap = trial.ctx.buckets()[2]
        log = {"p2g": {}, "g2p": {}, "capacity": {}}

        def timed(table, key, tuning, rank):
            # each candidate runs once per pass over the list; the faster pass stays
            r = trial.run(tuning, warmup, steps)
            old = table.get(key)
            if old is None or r[rank] < old[rank]:
                table[key] = r

        if state.bonded_body_count > 0 or state.fracture_bonds is not None:
            # bodies and bonds run 28-node P2G tiles and 256-thread G2P blocks
            for _ in range(2):
                for tb in BODIES_G2P_TILES:
                    timed(log["g2p"], f"{tb}x256", _tuning(28, tb, 256), "g2p_us")
            best = dict(BODIES_TUNING, g2p_tile=int(_best(log["g2p"], "g2p_us").split("x")[0]))
        else:
            for _ in range(2):
                for t in p2g_tiles:
                    if P2G_CHUNK[t] < derived_cap:
                        continue
                    timed(log["p2g"], str(t), _tuning(t, 8, 256), "p2g_us")
            p2g_tile = int(_best(log["p2g"], "p2g_us")) if log["p2g"] else 28  # only 28 holds every bucket
            for _ in range(2):
                for tb, th in g2p_shapes:
                    timed(log["g2p"], f"{tb}x{th}", _tuning(p2g_tile, tb, th), "g2p_us")
            g2p_key = _best(log["g2p"], "g2p_us")
            g2p_tile, g2p_threads = (int(v) for v in g2p_key.split("x"))
            caps = sorted({min(MAX_BUCKET_CAPACITY, max(MIN_BUCKET_CAPACITY, int(16 * round(derived_cap * f / 16))))
                           for f in capacity_factors})
            caps = [c for c in caps if P2G_CHUNK[p2g_tile] >= c]
            for _ in range(2):
                for cap in caps:
                    timed(log["capacity"], str(cap), _tuning(p2g_tile, g2p_tile, g2p_threads, cap), "step_us")
            cap_key = _best(log["capacity"], "step_us") if log["capacity"] else "0"
            best = {"p2g_tile": p2g_tile, "g2p_tile": g2p_tile, "g2p_threads": g2p_threads,
                    "bucket_capacity": int(cap_key)}
        return {"tuning": best, "class": klass, "device": dev, "derived_capacity": int(derived_cap),
                "candidates": log, "when": _dt.datetime.now(_dt.timezone.utc).isoformat(timespec="seconds")}
    finally:
        trial.close()


def autotune_batch(scenes, dt, gravity, p2g_mode: int = MPM_P2G_GATHER, *, device: int = 0, warmup: int = 4,
                   steps: int = 48, p2g_tiles=P2G_TILES, g2p_shapes=G2P_SHAPES) -> dict:
    """Launch shape of a batch of scenes (MpmMlsGpuBatch): every compiled
    shape is timed on a scratch batch built from the same scenes, by the wall
    time of `steps` microsteps of all scenes (a batch launch covers them all,
    so the two kernels are ranked together), each shape twice, interleaved,
    keeping its faster pass.  A step_batch call also pays host work per scene
    (256 tanks: about a quarter of a millisecond a microstep over 8 steps),
    so `steps` is long enough for the kernels to decide.  The scenes keep
    their bucket capacities.
    Returns {"tuning": {...}, "candidates": {shape: microstep microseconds}}."""
    from .solver import MpmMlsGpuBatch
    batch = MpmMlsGpuBatch(scenes, p2g_mode=p2g_mode, device=device)
    try:
        times: dict[str, float] = {}
        effective: dict[str, dict] = {}
        for _ in range(2):
            for t in p2g_tiles:
                for tb, th in g2p_shapes:
                    key = f"{t}/{tb}x{th}"
                    batch.native.set_tuning(_tuning(t, tb, th, 0))
                    batch.step_batch(warmup, dt, gravity)
                    batch.scene(0).ctx.read_compact()  # the scenes share one stream
                    t0 = time.perf_counter()
                    batch.step_batch(steps, dt, gravity)
                    batch.scene(0).ctx.read_compact()
                    us = 1e6 * (time.perf_counter() - t0) / steps
                    times[key] = min(times.get(key, us), us)
                    effective[key] = batch.native.get_tuning().as_dict()
        for i in range(len(batch)):
            batch.scene(i).check_numerical_status()
        best = min(times, key=times.get)
        return {"tuning": effective[best], "candidates": {k: round(v, 2) for k, v in times.items()}}
    finally:
        batch.dispose()


def tuning_for(particles, config: dict, p2g_mode: int = MPM_P2G_GATHER, *, cache: pathlib.Path | str | None = None,
               tune_if_missing: bool = True, device: int = 0, **kw) -> dict:
    """The cached tuning of this scene's class on this device; when there is
    none, tune (and store) it, or return the default shape."""
    state = layout.validate_input(particles, config)
    klass = scene_class(state, p2g_mode)
    dev = device_key(device)
    hit = lookup(dev, klass, cache)
    if hit is not None:
        return dict(hit["tuning"])
    if not tune_if_missing:
        return default_tuning(config, device, bodies=state.bonded_body_count > 0 or state.fracture_bonds is not None)
    entry = autotune(particles, config, p2g_mode, device=device, **kw)
    store_entry(dev, klass, entry, cache)
    return dict(entry["tuning"])


def apply(solver, tuning: dict) -> dict:
    """Set a tuning on a solver's context; returns the tuning it runs with."""
    solver.ctx.set_tuning(MpmTuning.from_dict(tuning))
    return solver.ctx.get_tuning().as_dict()
