"""GPU: launch-shape tuning and the autotuner.

A step's results must not depend on the tuning (include/mpm355.h, "kernel
tuning"): fixed-point P2G sums integers, the float gather sums every node in
a tile-independent order, G2P is per particle, and the bucket capacity only
sizes the slot arrays.  So every compiled shape is checked bit for bit
against the default shape, on a liquid scene and on a scene with solids
(which runs the stress-cache kernel variants), in both P2G modes, and so is a
bucket-capacity change in the middle of a run and a batch.  Then the
autotuner end to end: it picks a compiled shape, caches it per (device,
scene class), and the cache answers the second call.
"""
from __future__ import annotations

import pathlib
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRAVITY = {"x": 0.0, "y": -9.81}


def mixed_scene(pkg, per_side=2, seed=7):
    """A snow block, a neo-Hookean block and a water/oil pool on a 64x64 grid,
    per_side^2 particles per cell, with jittered positions and velocities."""
    rng = np.random.default_rng(seed)
    xs, ys, models, lam, mu, dens = [], [], [], [], [], []
    blocks = [((6, 20), (30, 44), 1, 40.0, 20.0, 1.0),    # snow
              ((36, 52), (30, 44), 0, 40.0, 20.0, 1.0),   # neo-Hookean
              ((4, 60), (2, 14), 2, 60.0, 0.0, 1.0),      # water
              ((4, 60), (14, 20), 3, 40.0, 0.0, 0.8)]     # oil
    h = 1.0 / per_side
    for (x0, x1), (y0, y1), m, la, mu_, de in blocks:
        gx, gy = np.meshgrid(np.arange(x0, x1, h) + h / 2, np.arange(y0, y1, h) + h / 2)
        n = gx.size
        xs.append(gx.ravel() + rng.uniform(-0.2, 0.2, n) * h)
        ys.append(gy.ravel() + rng.uniform(-0.2, 0.2, n) * h)
        models.append(np.full(n, m, np.uint32))
        lam.append(np.full(n, la))
        mu.append(np.full(n, mu_))
        dens.append(np.full(n, de))
    n = sum(len(a) for a in xs)
    arr = pkg.ParticleArrays(
        x=np.concatenate(xs), y=np.concatenate(ys), vx=rng.uniform(-1, 1, n), vy=rng.uniform(-1, 1, n),
        model=np.concatenate(models), lambda_=np.concatenate(lam), mu=np.concatenate(mu),
        density=np.concatenate(dens), volume=np.full(n, h * h), param=np.zeros(n), object_id=np.zeros(n, np.uint32))
    config = {"grid": {"cellsX": 64, "cellsY": 64, "cellSize": 1.0}, "particleMass": h * h, "particleVolume": h * h,
              "material": {"model": "water", "lambda": 60.0, "mu": 0.0, "density": 1.0}, "capacity": n}
    return arr, config


def scenes(pkg):
    return {"liquid": pkg.scenes.water_oil_column(cells=72), "mixed": mixed_scene(pkg)}


def factory_of(pkg, mode):
    return pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver


def run(pkg, mode, particles, config, tuning=None, steps=12, dt=None, capacity_change_at=None):
    s = factory_of(pkg, mode)(particles, config, device=0)
    try:
        if tuning is not None:
            pkg.autotune.apply(s, tuning)
        if dt is None:
            dt = pkg.autotune._default_dt(s, GRAVITY)
        if capacity_change_at is None:
            s.step_batch(steps, dt, GRAVITY)
        else:
            s.step_batch(capacity_change_at, dt, GRAVITY)
            t = s.ctx.get_tuning().as_dict()
            t["bucket_capacity"] = 2 * t["bucket_capacity"]
            pkg.autotune.apply(s, t)
            s.step_batch(steps - capacity_change_at, dt, GRAVITY)
        s.check_numerical_status()
        recs, _ = s.read_particles(packed=True)
        return recs, dt, s.ctx.get_tuning().as_dict()
    finally:
        s.dispose()


def default_shape(pkg, configs):
    """The library's default (include/mpm355.h, "kernel tuning"): 20-node P2G
    tiles, and 8x8-bucket G2P tiles of 256 threads unless the grids have
    fewer of those than the device has compute units, then 4x4-bucket ones."""
    tiles8 = 0
    for c in configs:
        bw, bh = c["grid"]["cellsX"] // 4 + 1, c["grid"]["cellsY"] // 4 + 1
        tiles8 += -(-bw // 8) * -(-bh // 8)
    return (20, 8 if tiles8 >= pkg.native.device_arch(0)[1] else 4, 256)


@pytest.mark.parametrize("mode", ["gather", "fixed"])
@pytest.mark.parametrize("scene", ["liquid", "mixed"])
def test_every_shape_is_bit_identical(pkg, mode, scene):
    particles, config = scenes(pkg)[scene]
    ref, dt, eff = run(pkg, mode, particles, config)
    assert (eff["p2g_tile"], eff["g2p_tile"], eff["g2p_threads"]) == default_shape(pkg, [config])
    for tile in pkg.autotune.P2G_TILES:
        for tb, th in pkg.autotune.G2P_SHAPES:
            want = {"p2g_tile": tile, "g2p_tile": tb, "g2p_threads": th, "bucket_capacity": 0}
            print(f"[tuning] {scene} {mode} {want}", flush=True)
            got, _, eff = run(pkg, mode, particles, config, want, dt=dt)
            assert (eff["p2g_tile"], eff["g2p_tile"], eff["g2p_threads"]) == (tile, tb, th)
            assert np.array_equal(got, ref), f"{scene} {mode} {want}: {np.count_nonzero(got != ref)} words differ"


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_bucket_capacity_change_mid_run(pkg, mode):
    particles, config = mixed_scene(pkg, per_side=3)
    ref, dt, eff = run(pkg, mode, particles, config, steps=10)
    got, _, eff2 = run(pkg, mode, particles, config, steps=10, dt=dt, capacity_change_at=4)
    assert eff2["bucket_capacity"] == 2 * eff["bucket_capacity"]
    assert np.array_equal(got, ref)


def test_default_shape_follows_the_grid_size(pkg):
    particles, config = pkg.scenes.water_oil_column(cells=32)
    big = dict(config, grid=dict(config["grid"], cellsX=1024, cellsY=1024))
    for cfg, want_g2p in ((config, 4), (big, 8)):
        assert default_shape(pkg, [cfg])[1] == want_g2p
        s = pkg.create_mpm_mls_gpu_solver(particles, cfg, device=0)
        try:
            eff = s.ctx.get_tuning().as_dict()
            assert (eff["p2g_tile"], eff["g2p_tile"], eff["g2p_threads"]) == (20, want_g2p, 256)
            s.step_batch(4, 1e-4, GRAVITY)
            s.check_numerical_status()
        finally:
            s.dispose()
    # the autotuner's fallback when it may not tune follows the same rule
    for cfg, want_g2p in ((config, 4), (big, 8)):
        assert pkg.autotune.default_tuning(cfg, 0)["g2p_tile"] == want_g2p
    # and for a scene with bodies reports the shape the library runs it with
    gp, gc, _ = pkg.scenes.gate4_scene()
    s = pkg.create_mpm_mls_gpu_solver(gp, gc, device=0)
    try:
        eff = s.ctx.get_tuning().as_dict()
    finally:
        s.dispose()
    fallback = pkg.autotune.tuning_for(gp, gc, tune_if_missing=False, device=0,
                                       cache=pathlib.Path(tempfile.mkdtemp()) / "none.json")
    assert {k: fallback[k] for k in ("p2g_tile", "g2p_tile", "g2p_threads")} == \
        {k: eff[k] for k in ("p2g_tile", "g2p_tile", "g2p_threads")}
    # a batch counts the tiles of all its grids
    b = pkg.MpmMlsGpuBatch([(particles, config)] * 3, device=0)
    try:
        eff = b.native.get_tuning().as_dict()
        assert (eff["p2g_tile"], eff["g2p_tile"], eff["g2p_threads"]) == default_shape(pkg, [config] * 3)
    finally:
        b.dispose()


def test_invalid_tuning_is_rejected_without_state_change(pkg):
    particles, config = pkg.scenes.water_oil_column(cells=32)
    s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0)
    try:
        before = s.ctx.get_tuning().as_dict()
        bad = [{"p2g_tile": 24, "g2p_tile": 8, "g2p_threads": 256},
               {"p2g_tile": 28, "g2p_tile": 4, "g2p_threads": 512},
               {"p2g_tile": 28, "g2p_tile": 16, "g2p_threads": 256},
               {"p2g_tile": 28, "g2p_tile": 8, "g2p_threads": 256, "bucket_capacity": 40},
               {"p2g_tile": 28, "g2p_tile": 8, "g2p_threads": 256, "bucket_capacity": 2048}]
        for t in bad:
            with pytest.raises(pkg.native.NativeError) as e:
                pkg.autotune.apply(s, t)
            assert e.value.code == pkg.native.MPM_EINVAL
            assert s.ctx.get_tuning().as_dict() == before
    finally:
        s.dispose()


def test_small_tile_falls_back_when_a_bucket_exceeds_its_chunk(pkg):
    particles, config = pkg.scenes.water_oil_column(cells=32)
    s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0)
    try:
        eff = pkg.autotune.apply(s, {"p2g_tile": 12, "g2p_tile": 4, "g2p_threads": 128, "bucket_capacity": 512})
        assert eff == {"p2g_tile": 28, "g2p_tile": 4, "g2p_threads": 128, "bucket_capacity": 512, "tile_queue": 0}
        eff = pkg.autotune.apply(s, {"p2g_tile": 12, "g2p_tile": 4, "g2p_threads": 128, "bucket_capacity": 64})
        assert eff["p2g_tile"] == 12 and eff["bucket_capacity"] == 64
        s.step_batch(3, 1e-4, GRAVITY)
        s.check_numerical_status()
    finally:
        s.dispose()


def test_scene_with_bodies_runs_28_node_p2g_and_both_g2p_tiles(pkg):
    particles, config, gravity = pkg.scenes.gate4_scene(mpm_grid=False, liquid_area=100.0)
    out = {}
    for tb in (8, 4):
        s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0)
        try:
            eff = pkg.autotune.apply(s, {"p2g_tile": 20, "g2p_tile": tb, "g2p_threads": 128, "bucket_capacity": 0})
            assert (eff["p2g_tile"], eff["g2p_tile"], eff["g2p_threads"]) == (28, tb, 256)
            print(f"[tuning] gate4 g2p tile {tb}", flush=True)
            s.step_batch(6, 1.0 / 240.0, gravity)
            s.check_numerical_status()
            out[tb] = s.read_particles(packed=True)[0]
        finally:
            s.dispose()
    assert np.array_equal(out[8], out[4])


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_batch_shape_is_bit_identical(pkg, mode):
    p_mode = pkg.native.MPM_P2G_GATHER if mode == "gather" else pkg.native.MPM_P2G_FIXED_POINT
    members = [pkg.scenes.water_oil_column(cells=40), mixed_scene(pkg, seed=3), pkg.scenes.water_oil_column(cells=24)]

    def go(tuning):
        b = pkg.MpmMlsGpuBatch(members, p2g_mode=p_mode, device=0, tuning=tuning)
        try:
            b.step_batch(8, 2e-4, GRAVITY)
            out = []
            for i in range(len(b)):
                b.scene(i).check_numerical_status()
                out.append(b.scene(i).read_particles(packed=True)[0])
            if tuning is not None:
                # the scenes of a batch share its shape
                other = dict(tuning, p2g_tile=28 if tuning["p2g_tile"] != 28 else 20)
                with pytest.raises(pkg.native.NativeError):
                    pkg.autotune.apply(b.scene(0), other)
            return out, b.native.get_tuning().as_dict()
        finally:
            b.dispose()

    ref, eff = go(None)
    assert eff["p2g_tile"] == 20
    for t in ({"p2g_tile": 12, "g2p_tile": 4, "g2p_threads": 128, "bucket_capacity": 0},
              {"p2g_tile": 28, "g2p_tile": 8, "g2p_threads": 512, "bucket_capacity": 0}):
        got, eff = go(t)
        assert (eff["p2g_tile"], eff["g2p_tile"], eff["g2p_threads"]) == (t["p2g_tile"], t["g2p_tile"], t["g2p_threads"])
        for a, r in zip(got, ref):
            assert np.array_equal(a, r)


def test_autotuner_picks_a_compiled_shape_and_caches_it(pkg, tmp_path, monkeypatch):
    particles, config = pkg.scenes.water_oil_column(cells=96)
    cache = tmp_path / "tuning.json"
    arch, cus = pkg.native.device_arch(0)
    assert arch.startswith("gfx") and cus > 0
    t = pkg.autotune.tuning_for(particles, config, cache=cache, steps=4, warmup=1)
    assert pkg.autotune.shape_supported(t["p2g_tile"], t["g2p_tile"], t["g2p_threads"])
    entries = pkg.autotune.load_cache(cache)
    dev = pkg.autotune.device_key(0)
    klass = pkg.autotune.scene_class(pkg.layout.validate_input(particles, config), 0)
    entry = entries[dev][klass]
    assert entry["tuning"] == t
    assert set(entry["candidates"]["p2g"]) == {"28", "20", "12"}
    assert len(entry["candidates"]["g2p"]) == len(pkg.autotune.G2P_SHAPES)
    for r in entry["candidates"]["p2g"].values():
        assert r["p2g_us"] > 0 and r["g2p_us"] > 0
    # the second call is answered by the cache
    monkeypatch.setattr(pkg.autotune, "autotune", lambda *a, **k: pytest.fail("tuned again"))
    assert pkg.autotune.tuning_for(particles, config, cache=cache) == t
    # and a solver takes it
    s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0, tuning=t)
    try:
        eff = s.ctx.get_tuning().as_dict()
        assert (eff["p2g_tile"], eff["g2p_tile"], eff["g2p_threads"]) == (t["p2g_tile"], t["g2p_tile"],
                                                                          t["g2p_threads"])
    finally:
        s.dispose()


def test_batch_autotuner_picks_a_compiled_shape(pkg):
    members = [pkg.scenes.water_oil_column(cells=40), mixed_scene(pkg, seed=5), pkg.scenes.water_oil_column(cells=24)]
    r = pkg.autotune.autotune_batch(members, 2e-4, GRAVITY, device=0, steps=2, warmup=1)
    t = r["tuning"]
    assert pkg.autotune.shape_supported(t["p2g_tile"], t["g2p_tile"], t["g2p_threads"])
    assert len(r["candidates"]) == len(pkg.autotune.P2G_TILES) * len(pkg.autotune.G2P_SHAPES)
    assert all(v > 0 for v in r["candidates"].values())
