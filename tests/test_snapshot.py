"""The snapshot file (the boundary's wire formats as a file): records pack
back to themselves (tombstones included), the file round-trips and rejects
damage.  CPU only; tests/test_gpu_snapshot.py restores solvers from it."""
from __future__ import annotations

import numpy as np
import pytest


def scene_records(pkg):
    parts, config, _ = pkg.scenes.gate4_scene()
    state = pkg.layout.validate_input(parts, config)
    return state, config


def test_records_pack_back_to_themselves(pkg):
    state, config = scene_records(pkg)
    rec = state.packed.copy()
    liquid = np.flatnonzero((rec[:, 13] == 2) | (rec[:, 13] == 3))
    rec[liquid[[3, 50, 700]], 13] = pkg.layout.MODEL_REMOVED  # outlet tombstones
    again = pkg.layout.validated_from_records(rec, state.object_ids, config)
    assert np.array_equal(again.packed, rec)
    assert again.bonded_body_count == state.bonded_body_count
    assert again.object_ranges == state.object_ranges


def test_records_that_do_not_repack_are_refused(pkg):
    state, config = scene_records(pkg)
    rec = state.packed.copy()
    rec[10, 16] = np.float32(0.5).view(np.uint32)  # a density on a solid record: the packer writes 0
    solid = np.flatnonzero(rec[:, 13] == 0)[0]
    rec[solid, 16] = np.float32(0.5).view(np.uint32)
    with pytest.raises(pkg.RangeError, match="does not pack back"):
        pkg.layout.validated_from_records(rec, state.object_ids, config)


def fake_snapshot(pkg):
    state, config = scene_records(pkg)
    header = {"version": 1, "config": config, "p2gMode": 0, "particleCount": int(state.packed.shape[0]),
              "bondCount": 2, "activeParticleCount": 7, "nextObjectId": 3, "objectRanges": [],
              "bondedBodyCount": 0, "cflBounds": None, "lastScheduledVelocityBound": None, "latestSequence": 0,
              "cflStateUsable": True}
    import json
    import struct
    import zlib
    head = json.dumps(header).encode()
    n = state.packed.shape[0]
    body = (pkg.snapshot.MAGIC + struct.pack("<I", len(head)) + head + state.packed.astype("<u4").tobytes() +
            state.object_ids.astype("<u4").tobytes() + np.arange(12, dtype="<u4").tobytes() +
            np.array([0, 0xFFFFFFFF, 0, 0, 0, 0, 0, 0], "<u4").tobytes())
    return body + struct.pack("<I", zlib.crc32(body)), state, n


def test_file_round_trip_and_damage(pkg, tmp_path):
    data, state, n = fake_snapshot(pkg)
    snap = pkg.snapshot.decode(data)
    assert np.array_equal(snap["records"], state.packed)
    assert np.array_equal(snap["objectIds"], state.object_ids)
    assert np.array_equal(snap["bonds"].ravel(), np.arange(12))
    assert snap["status"][1] == 0xFFFFFFFF and snap["header"]["nextObjectId"] == 3
    (tmp_path / "s.mpm").write_bytes(data)
    assert np.array_equal(pkg.snapshot.read(tmp_path / "s.mpm")["records"], state.packed)
    bad = bytearray(data)
    bad[200] ^= 1
    with pytest.raises(pkg.RangeError, match="checksum"):
        pkg.snapshot.decode(bytes(bad))
    with pytest.raises(pkg.RangeError, match="not an mpm355"):
        pkg.snapshot.decode(b"PK" + data[2:])
    with pytest.raises(pkg.RangeError):
        pkg.snapshot.decode(data[:-8] + data[-4:])
