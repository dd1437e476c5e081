"""Particle-state snapshots: the boundary's wire format as a file.

A snapshot holds a solver's committed state in the reference's own
formats, so that a run can stop and resume, and states can be exchanged:
the 20-word particle records (layout.ts:385-437, as readParticles returns
them, solver.ts:2594-2731; removed tombstones keep their slots), the object
ids, the fracture bonds in their 24-byte records (layout.ts:527-548) with
their current damage, and the 8 status words (shaders.ts:157-172), next to
a JSON header with the configuration and the host's bookkeeping (object
ranges, next object id, CFL bounds and the tracked velocity bound).

File layout (little-endian):

    magic     8 bytes  b"MPM355S1"
    header    u32 byte length, then that many bytes of UTF-8 JSON
    records   particleCount x 20 u32
    objectIds particleCount x u32
    bonds     bondCount x 6 u32
    status    8 u32
    crc32     u32 of every byte before it

``load`` creates a solver from the file.  It continues bit for bit like the
solver the snapshot was taken from: the step's results do not depend on
where particles sit in the buckets.  The numerical status is kept in the file
for inspection; the restored solver starts with a clear one.
"""
from __future__ import annotations

import json
import pathlib
import struct
import zlib

import numpy as np

from . import cfl as cflmod
from . import layout
from .layout import RangeError

MAGIC = b"MPM355S1"
VERSION = 1


def encode(solver) -> bytes:
    """The snapshot bytes of a solver's committed state."""
    solver._assert_active()
    ctx = solver.ctx
    records, ids = ctx.read_particles()
    status = np.asarray(ctx.read_status(), np.uint32)
    nb = 0 if solver.state.fracture_bonds is None else solver.state.fracture_bonds.shape[0]
    bonds = ctx.read_fracture_bonds(nb) if nb else np.zeros((0, 6), np.uint32)
    config = dict(solver.state.config)
    if nb:
        config["fractureBonds"] = layout.unpack_fracture_bonds(bonds)
    b = solver.cfl_bounds
    header = {
        "version": VERSION,
        "config": config,
        "p2gMode": int(solver.p2g_mode),
        "particleCount": int(records.shape[0]),
        "bondCount": int(nb),
        "activeParticleCount": int(solver._active_count),
        "nextObjectId": int(solver._next_object_id),
        "objectRanges": sorted(solver._object_ranges.values(), key=lambda r: r["start"]),
        "bondedBodyCount": int(solver._body_count),
        "cflBounds": None if b is None else {"waveSpeedBound": b.waveSpeedBound, "velocityBound": b.velocityBound,
                                             "courantLimit": b.courantLimit},
        "lastScheduledVelocityBound": solver._last_scheduled_velocity_bound,
        "latestSequence": int(solver._latest_sequence),
        "cflStateUsable": bool(solver._cfl_state_usable),
    }
    head = json.dumps(header, sort_keys=True).encode()
    body = b"".join([MAGIC, struct.pack("<I", len(head)), head,
                     np.ascontiguousarray(records, "<u4").tobytes(), np.ascontiguousarray(ids, "<u4").tobytes(),
                     np.ascontiguousarray(bonds, "<u4").tobytes(), status.astype("<u4").tobytes()])
    return body + struct.pack("<I", zlib.crc32(body) & 0xFFFFFFFF)


def decode(data: bytes) -> dict:
    """Parse and check snapshot bytes: the header, records, object ids, bonds
    and status words.  RangeError for a damaged or foreign file."""
    if len(data) < len(MAGIC) + 8 or data[:len(MAGIC)] != MAGIC:
        raise RangeError("not an mpm355 particle snapshot")
    (crc,) = struct.unpack("<I", data[-4:])
    if zlib.crc32(data[:-4]) & 0xFFFFFFFF != crc:
        raise RangeError("snapshot checksum mismatch: the file is damaged")
    (hlen,) = struct.unpack_from("<I", data, len(MAGIC))
    off = len(MAGIC) + 4
    header = json.loads(data[off:off + hlen].decode())
    if header.get("version") != VERSION:
        raise RangeError(f"unsupported snapshot version {header.get('version')}")
    off += hlen
    n, nb = int(header["particleCount"]), int(header["bondCount"])
    need = off + 4 * (20 * n + n + 6 * nb + 8) + 4
    if len(data) != need:
        raise RangeError("snapshot length does not match its header")

    def take(words):
        nonlocal off
        a = np.frombuffer(data, "<u4", words, off).astype(np.uint32)
        off += 4 * words
        return a

    records = take(20 * n).reshape(n, 20)
    ids = take(n)
    bonds = take(6 * nb).reshape(nb, 6)
    status = take(8)
    return {"header": header, "records": records, "objectIds": ids, "bonds": bonds, "status": status}


def save(solver, path) -> pathlib.Path:
    path = pathlib.Path(path)
    tmp = path.with_name(path.name + ".tmp")
    tmp.write_bytes(encode(solver))
    tmp.replace(path)
    return path


def read(path) -> dict:
    return decode(pathlib.Path(path).read_bytes())


def restore(snap: dict, device: int = -1, tuning: dict | None = None, keep_fixed_grid: bool = False):
    """A solver continuing from a decoded snapshot."""
    from .solver import MpmMlsGpuSolver
    h = snap["header"]
    state = layout.validated_from_records(snap["records"], snap["objectIds"], h["config"])
    b = h["cflBounds"]
    bounds = None if b is None else cflmod.CflBounds(b["waveSpeedBound"], b["velocityBound"], b["courantLimit"])
    s = MpmMlsGpuSolver(state, h["config"], bounds, int(h["p2gMode"]), keep_fixed_grid=keep_fixed_grid,
                        device=device, tuning=tuning)
    s.cfl_bounds = bounds
    s._active_count = int(h["activeParticleCount"])
    s._has_removed = s._active_count != state.particle_count
    s._next_object_id = int(h["nextObjectId"])
    s._object_ranges = {int(r["objectId"]): dict(r) for r in h["objectRanges"]}
    s._body_count = int(h["bondedBodyCount"])
    s._last_scheduled_velocity_bound = h["lastScheduledVelocityBound"]
    s._latest_sequence = int(h["latestSequence"])
    s._cfl_state_usable = bool(h["cflStateUsable"])
    return s


def load(path, device: int = -1, tuning: dict | None = None, keep_fixed_grid: bool = False):
    return restore(read(path), device=device, tuning=tuning, keep_fixed_grid=keep_fixed_grid)
