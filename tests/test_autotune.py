"""Host side of the autotuner: scene classes and the tuning cache (no GPU)."""
from __future__ import annotations

import json

import numpy as np


def test_scene_class_tracks_what_the_kernels_depend_on(pkg):
    at = pkg.autotune
    arr, cfg = pkg.scenes.water_oil_column(cells=64)
    st = pkg.layout.validate_input(arr, cfg)
    k = at.scene_class(st, pkg.native.MPM_P2G_GATHER)
    assert k == "gather|nodes=2^12|particles=2^12|phases=water+oil|bodies=0|per_bucket=2^4"
    # the P2G mode is part of the class
    assert at.scene_class(st, pkg.native.MPM_P2G_FIXED_POINT) == k.replace("gather", "fixed", 1)
    # another seed of the same scene: same class
    arr2, cfg2 = pkg.scenes.water_oil_column(cells=64, seed=12345)
    assert at.scene_class(pkg.layout.validate_input(arr2, cfg2), 0) == k
    # four particles per cell: four times the bucket occupancy
    n = len(arr)
    dense = pkg.ParticleArrays(
        x=np.repeat(arr.x, 4) + np.tile([-0.2, 0.2, -0.2, 0.2], n), y=np.repeat(arr.y, 4) + np.tile([-0.2, -0.2, 0.2, 0.2], n),
        vx=np.repeat(arr.vx, 4), vy=np.repeat(arr.vy, 4), model=np.repeat(arr.model, 4),
        lambda_=np.repeat(arr.lambda_, 4), mu=np.repeat(arr.mu, 4), density=np.repeat(arr.density, 4),
        volume=np.repeat(arr.volume, 4) / 4, param=np.repeat(arr.param, 4), object_id=np.repeat(arr.object_id, 4))
    dense.x = np.clip(dense.x, 0.51, 63.49)
    dense.y = np.clip(dense.y, 0.51, 63.49)
    kd = at.scene_class(pkg.layout.validate_input(dense, dict(cfg, capacity=4 * n)), 0)
    assert "per_bucket=2^6" in kd and "particles=2^14" in kd
    # solids change the phases
    solid = pkg.ParticleArrays(x=arr.x, y=arr.y, vx=arr.vx, vy=arr.vy, model=np.ones(n, np.uint32),
                               lambda_=np.full(n, 40.0), mu=np.full(n, 20.0), density=np.ones(n),
                               volume=np.ones(n), param=np.zeros(n), object_id=np.zeros(n, np.uint32))
    ks = at.scene_class(pkg.layout.validate_input(solid, cfg), 0)
    assert "phases=solid|" in ks
    # bonded bodies and fracture bonds are flagged (they run 28-node P2G tiles and 256-thread G2P)
    particles, gcfg, _ = pkg.scenes.gate4_scene()
    kg = at.scene_class(pkg.layout.validate_input(particles, gcfg), 0)
    assert "bodies=1" in kg and "phases=solid+water+oil" in kg


def test_supported_shapes(pkg):
    at = pkg.autotune
    assert at.shape_supported(28, 8, 256)
    for t in at.P2G_TILES:
        for tb, th in at.G2P_SHAPES:
            assert at.shape_supported(t, tb, th)
    assert not at.shape_supported(24, 8, 256)
    assert not at.shape_supported(28, 4, 512)
    assert not at.shape_supported(28, 8, 64)
    # a chunk of every tile holds the default derived capacities of 1 and 4 particles per cell
    assert min(at.P2G_CHUNK.values()) >= 2 * 4 * 16


def test_cache_round_trip_and_guards(pkg, tmp_path):
    at = pkg.autotune
    path = tmp_path / "sub" / "tuning.json"
    assert at.load_cache(path) == {}
    good = {"tuning": {"p2g_tile": 20, "g2p_tile": 4, "g2p_threads": 128, "bucket_capacity": 48}, "class": "k1"}
    at.store_entry("gfx950:sramecc+:xnack-/256cu", "k1", good, path)
    at.store_entry("gfx950:sramecc+:xnack-/256cu", "k2", dict(good, tuning=dict(good["tuning"], p2g_tile=12)), path)
    at.store_entry("gfx942/304cu", "k1", good, path)
    assert at.lookup("gfx950:sramecc+:xnack-/256cu", "k1", path)["tuning"]["p2g_tile"] == 20
    assert at.lookup("gfx950:sramecc+:xnack-/256cu", "k2", path)["tuning"]["p2g_tile"] == 12
    assert at.lookup("gfx942/304cu", "k1", path) is not None
    assert at.lookup("gfx950:sramecc+:xnack-/256cu", "missing", path) is None
    # an entry naming a shape this library does not compile is ignored
    at.store_entry("dev", "bad", {"tuning": {"p2g_tile": 24, "g2p_tile": 8, "g2p_threads": 256}}, path)
    assert at.lookup("dev", "bad", path) is None
    # another cache version, or a damaged file, reads as empty
    data = json.loads(path.read_text())
    path.write_text(json.dumps(dict(data, version=at.CACHE_VERSION + 1)))
    assert at.load_cache(path) == {}
    path.write_text("{not json")
    assert at.load_cache(path) == {}
    # no temporary file is left behind
    assert sorted(p.name for p in path.parent.iterdir()) == ["tuning.json"]


def test_cache_path_from_environment(pkg, tmp_path, monkeypatch):
    monkeypatch.setenv("MPM355_TUNING_CACHE", str(tmp_path / "x.json"))
    assert pkg.autotune.default_cache_path() == tmp_path / "x.json"
