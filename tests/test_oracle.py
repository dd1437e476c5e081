"""Pins the CPU oracle (oracle/mpm_oracle.c) against the reference's own spec
cases (src/lib/mpm/reference/solver.spec.ts) and against what the mathematics
fixes (partition of unity, conservation, exact clamps, hand-computed
fixed-point words, brute-force binning).  CPU only."""
from __future__ import annotations

import math
import pathlib

import numpy as np
import pytest


def particle(orc, x=8.0, y=8.0, vx=0.0, vy=0.0, C=(0, 0, 0, 0), F=(1, 0, 0, 1), model=0, lam=20.0, mu=10.0,
             density=0.0, volume=0.01, param=0.0, pv=1.0):
    p = np.zeros(1, orc.PARTICLE_DTYPE)
    p["x"], p["y"], p["vx"], p["vy"] = x, y, vx, vy
    for k, name in enumerate(("c00", "c01", "c10", "c11")):
        p[name] = C[k]
    for k, name in enumerate(("f00", "f01", "f10", "f11")):
        p[name] = F[k]
    p["plastic_volume"], p["lambda_"], p["mu"], p["density"], p["volume"], p["param"] = pv, lam, mu, density, volume, param
    p["model"] = model
    return p


def cfg(orc, dt=0.001, cells=16, cell=1.0, mass=0.1, volume=0.01, g=(0.0, 0.0), friction=0.0, cells_y=None):
    return orc.make_config(dt, cells, cells if cells_y is None else cells_y, cell, mass, volume, g, friction)


def det(p):
    return float(p["f00"] * p["f11"] - p["f01"] * p["f10"])


def test_empty_step_has_zero_diagnostics(orc):  # solver.spec.ts:63-75
    out, d = orc.step_f64(np.zeros(0, orc.PARTICLE_DTYPE), cfg(orc))
    assert len(out) == 0
    assert (d.input_mass, d.grid_mass, d.output_mass, d.grid_px, d.output_py) == (0, 0, 0, 0, 0)


def test_rest_state_is_a_fixed_point(orc):
    p = particle(orc)
    out, _ = orc.step_f64(p, cfg(orc))
    for name in ("x", "y", "vx", "vy", "c00", "c01", "c10", "c11", "f00", "f01", "f10", "f11"):
        assert out[name][0] == p[name][0], name


@pytest.mark.parametrize("x,y", [(8.0, 8.0), (7.3, 9.81), (3.05, 12.5), (0.2, 15.9)])
def test_partition_of_unity_and_zero_first_moment(orc, x, y):
    """Quadratic B-spline weights sum to one and have zero first moment, so a
    uniformly moving particle deposits exactly m and m*v and comes back with
    C = 0 (no walls, no gravity)."""
    p = particle(orc, x=x, y=y, vx=0.3, vy=-0.7)
    c = cfg(orc, mass=0.25)
    g = orc.grid_f64(p, c)
    m = g[:, orc.GRID_FIELDS.index("m")].sum()
    qx = g[:, orc.GRID_FIELDS.index("qx")].sum()
    assert m == pytest.approx(0.25, abs=1e-15)
    assert qx == pytest.approx(0.25 * 0.3, abs=1e-15)
    if 2 <= x <= 14 and 2 <= y <= 14:  # away from the wall cushion
        out, _ = orc.step_f64(p, c)
        assert out["vx"][0] == pytest.approx(0.3, abs=1e-14)
        assert out["x"][0] == pytest.approx(x + 0.001 * 0.3, abs=1e-14)
        for name in ("c00", "c01", "c10", "c11"):
            assert abs(out[name][0]) < 1e-12


def test_wall_collisions_are_exact(orc):  # solver.spec.ts:213-252
    c = cfg(orc, dt=0.1, cells=8)
    cases = [((0, 4), (-10, 0)), ((8, 4), (10, 0)), ((4, 0), (0, -10)), ((4, 8), (0, 10))]
    for (x, y), (vx, vy) in cases:
        out, _ = orc.step_f64(particle(orc, x=x, y=y, vx=vx, vy=vy), c)
        assert (out["x"][0], out["y"][0]) == (x, y)
        assert (out["vx"][0], out["vy"][0]) == (0.0, 0.0)


def test_wall_friction_damps_tangential_velocity(orc):  # solver.spec.ts:254-270
    p = particle(orc, x=0, y=4, vx=0, vy=10)
    free, _ = orc.step_f64(p, cfg(orc, dt=0.1, cells=8))
    noslip, _ = orc.step_f64(p, cfg(orc, dt=0.1, cells=8, friction=1.0))
    assert free["vy"][0] > 0
    assert noslip["vy"][0] < free["vy"][0]


def test_tait_restores_more_than_linear(orc):  # solver.spec.ts:272-302
    c = cfg(orc, dt=1e-5, cells=8, mass=0.8, volume=0.04)
    kw = dict(x=4, y=4, F=(0.975, 0, 0, 0.975), model=2, lam=1000.0, mu=0.0, density=20.0, volume=0.04)
    tait, _ = orc.step_f64(particle(orc, param=7.15, **kw), c)
    lin, _ = orc.step_f64(particle(orc, param=0.0, **kw), c)
    assert det(tait[0]) > det(lin[0])


def test_mass_and_momentum_conservation(orc):  # solver.spec.ts:354-390
    ps = np.concatenate([particle(orc, x=24.2, y=21.7, vx=0.5, vy=-0.2),
                         particle(orc, x=25.1, y=21.9, vx=-0.1, vy=0.4),
                         particle(orc, x=24.6, y=22.8, vx=0.2, vy=0.1)])
    _, d = orc.step_f64(ps, cfg(orc, cells=64))
    px, py = 0.1 * (0.5 - 0.1 + 0.2), 0.1 * (-0.2 + 0.4 + 0.1)
    assert abs(d.grid_mass - d.input_mass) <= 1e-12 and abs(d.output_mass - d.input_mass) <= 1e-12
    for v, want in ((d.grid_px, px), (d.grid_py, py), (d.output_px, px), (d.output_py, py)):
        assert abs(v - want) <= 1e-10


def test_snow_projection_clamps_to_yield(orc):  # solver.spec.ts:447-476
    ps = np.concatenate([particle(orc, x=7.5, model=0), particle(orc, x=8.5, F=(0.9, 0, 0, 0.9), model=1)])
    out, _ = orc.step_f64(ps, cfg(orc, dt=1e-9))
    assert out["f00"][0] == pytest.approx(1, abs=1e-5)
    assert out["plastic_volume"][0] == 1
    assert out["f00"][1] == pytest.approx(0.975, abs=1e-6)
    assert out["f11"][1] == pytest.approx(0.975, abs=1e-6)
    assert out["plastic_volume"][1] == pytest.approx(0.81 / 0.975 ** 2, abs=1e-6)


def test_rotated_snow_keeps_principal_rotation(orc):  # solver.spec.ts:1175-1200
    a = 0.37
    cs, sn = math.cos(a), math.sin(a)
    out, _ = orc.step_f64(particle(orc, F=(cs * 0.9, -sn * 1.1, sn * 0.9, cs * 1.1), model=1), cfg(orc, dt=1e-9))
    assert out["f00"][0] == pytest.approx(cs * 0.975, abs=1e-6)
    assert out["f01"][0] == pytest.approx(-sn * 1.0075, abs=1e-6)
    assert out["f10"][0] == pytest.approx(sn * 0.975, abs=1e-6)
    assert out["f11"][0] == pytest.approx(cs * 1.0075, abs=1e-6)


def test_neo_hookean_stretch_clamp(orc):  # solver.spec.ts:478-508
    out, _ = orc.step_f64(particle(orc, F=(1, 0, 0, 1e-8)), cfg(orc, dt=1e-8))
    assert det(out[0]) >= 0.05 * 0.05 - 1e-10
    assert out["f11"][0] == pytest.approx(0.05, abs=1e-6)
    out, _ = orc.step_f64(particle(orc, F=(1000, 0, 0, 1e-8)), cfg(orc, dt=1e-12))
    assert out["f00"][0] == pytest.approx(20, abs=1e-5)
    assert out["f11"][0] == pytest.approx(0.05, abs=1e-5)


def test_liquid_shear_projects_to_rotation(orc):  # solver.spec.ts:510-551, :1156-1173
    for model in (2, 3):
        out, _ = orc.step_f64(particle(orc, F=(1, 1, 0, 1), model=model, lam=20.0, mu=0.0, density=0.04), cfg(orc))
        f = out[0]
        assert f["f00"] == pytest.approx(f["f11"], abs=1e-6)
        assert f["f01"] == pytest.approx(-f["f10"], abs=1e-6)
        assert det(f) == pytest.approx(1, abs=1e-6)


def test_liquid_determinant_bounded(orc):  # solver.spec.ts:1062-1082
    p = particle(orc, F=(3, 0, 0, 3), model=2, lam=8.0, mu=0.0, density=1.0)
    for _ in range(20):
        p, _ = orc.step_f64(p, cfg(orc))
        assert 0.5 - 1e-10 <= det(p[0]) <= 2 + 1e-10


def test_powder_determinant_floor(orc):  # solver.spec.ts:1084-1102
    out, _ = orc.step_f64(particle(orc, F=(0.1, 0, 0, 1), model=4, param=0.55), cfg(orc, dt=1e-12))
    assert det(out[0]) == pytest.approx(0.5, abs=1e-10)


def test_unyielded_powder_is_untouched(orc):  # solver.spec.ts:1141-1154
    out, _ = orc.step_f64(particle(orc, model=4, param=0.55), cfg(orc))
    assert (out["f00"][0], out["f01"][0], out["f10"][0], out["f11"][0]) == (1, 0, 0, 1)


@pytest.mark.parametrize("dt,speed", [(1e-4, 0.5), (1e-3, 0.5), (1e-2, 0.5), (1e-3, 64.0)])
def test_water_oil_contact_conserves_momentum(orc, dt, speed):  # solver.spec.ts:652-733
    ps = np.concatenate([
        particle(orc, x=7.5, vx=speed, model=2, lam=8.0, mu=0.0, density=1.0, volume=0.25),
        particle(orc, x=8.5, vx=-speed, model=3, lam=8.0, mu=0.0, density=1.0, volume=0.25)])
    out, d = orc.step_f64(ps, cfg(orc, dt=dt, volume=0.25))
    assert out["vx"][0] < speed and out["vx"][1] > -speed
    assert 0 <= out["vx"][1] - out["vx"][0] <= max(2 * speed, 5) + 1e-8
    assert out["x"][0] <= out["x"][1]
    assert d.input_mass == 0.5 and d.output_mass == 0.5
    assert d.output_px == pytest.approx(d.input_px, abs=1e-10)
    if speed > 1:
        ke = 0.5 * 0.25 * float(np.sum(out["vx"] ** 2 + out["vy"] ** 2))
        assert ke < 0.25 * speed ** 2


def test_bonded_particles_are_reported_unsupported(orc):
    p = particle(orc)
    p["bonded_vertex"] = 1
    with pytest.raises(orc.OracleError) as e:
        orc.step_f64(p, cfg(orc))
    assert e.value.code == -5


# ----------------------------------------------------------- integer path

def packed_one(x, y, vx=0.0, vy=0.0, model=2, lam=8.0, density=1.0, volume=1.0, F=(1, 0, 0, 1), C=(0, 0, 0, 0),
               param=0.0, mu=0.0):
    r = np.zeros(20, np.float32)
    r[0:4] = (x, y, vx, vy)
    r[4:8] = C
    r[8:12] = F
    r[12] = 1.0
    r[14], r[15], r[16], r[17], r[18] = lam, mu, density, volume, param
    u = r.view(np.uint32)
    u[13] = model
    return u


def test_fixed_point_words_by_hand(orc):
    """One water particle at rest at a cell centre (F = I, linear EOS so the
    stress is zero): each word is round(w_x w_y * m / m_ref * 2^24)."""
    x, y = np.float32(5.5), np.float32(6.25)
    rec = packed_one(x, y, density=0.02, volume=1.0)
    prm = orc.params_f32(16, 16, 1.0, 0.02, 1.0, 1e-3, 1)
    grid, flags = orc.fixed_p2g_f32(rec, prm)
    assert flags == 0
    bx, by = 5, 5  # floor(5.5-0.5), floor(6.25-0.5)
    fx, fy = 5.5 - 5, 6.25 - 5
    wx = [0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1) ** 2, 0.5 * (fx - 0.5) ** 2]
    wy = [0.5 * (1.5 - fy) ** 2, 0.75 - (fy - 1) ** 2, 0.5 * (fy - 0.5) ** 2]
    W = 19
    total = 0
    for oy in range(3):
        for ox in range(3):
            node = (by + oy + 1) * W + (bx + ox + 1)
            mass_ratio = np.float32(np.float32(0.02) * np.float32(1.0))  # density*volume / particle mass
            w = wx[ox] * wy[oy]
            expect = int(np.rint(np.float32(np.float32(w) * mass_ratio) / np.float32(0.02) * np.float32(2 ** 24)))
            assert grid[node, 3] == expect, (ox, oy)
            assert grid[node, 6] == int(np.rint(np.float32(w) * np.float32(2 ** 24)))  # volume word
            assert grid[node, 4] == 0 and grid[node, 5] == 0 and grid[node, 0] == 0
            total += grid[node, 3]
    assert abs(total - 2 ** 24) <= 9  # sum of weights is one, up to per-word rounding
    nonzero_weights = sum(1 for a in wx for b in wy if a * b != 0)
    assert np.count_nonzero(grid) == 2 * nonzero_weights  # mass and volume words only


def test_fixed_point_is_order_independent_and_matches_fp64(orc, pkg):
    arr, config = pkg.scenes.water_oil_column(cells=24)
    st = pkg.layout.validate_input(arr, config)
    prm = orc.params_f32(24, 24, 1.0, 0.02, 1.0, 1.0 / 9600.0, len(arr))
    g1, f1 = orc.fixed_p2g_f32(st.packed, prm)
    perm = np.random.default_rng(3).permutation(len(arr))
    g2, f2 = orc.fixed_p2g_f32(st.packed[perm], prm)
    assert f1 == 0 and f2 == 0
    assert np.array_equal(g1, g2)
    dec = orc.decode_fixed_f32(g1, prm)
    parts = orc.particles_from_packed(st.packed)
    fp64 = orc.grid_f64(parts, orc.make_config(1.0 / 9600.0, 24, 24, 1.0, 0.02, 1.0))
    F = orc.GRID_FIELDS
    # decoded water mass / oil volume vs FP64 deposits: within 9 quanta of 2^-24 * scale per node
    assert np.max(np.abs(dec[:, 3] - fp64[:, F.index("wm")])) < 9 * 0.02 / 2 ** 24 + 1e-9
    assert np.max(np.abs(dec[:, 10] - fp64[:, F.index("oV")])) < 9 * 1.0 / 2 ** 24 + 1e-7
    assert np.max(np.abs(dec[:, 4] - fp64[:, F.index("wqx")])) < 1e-5


def test_fixed_point_overflow_flags(orc):
    """transfer-regression.ts:97-170: 512 coincident particles overflow the
    accumulator; one particle at 1e7 overflows a single contribution."""
    prm = orc.params_f32(80, 80, 0.005, 2.4955e-4, 2.5e-5, 1e-5, 512)
    rec = packed_one(0.1, 0.1, density=9.982, volume=2.5e-5, lam=1.0, param=0.0)
    many = np.tile(rec, (512, 1))
    _, flags = orc.fixed_p2g_f32(many, prm)
    assert flags & 65536
    one = packed_one(0.1, 0.1, vx=1e7, density=9.982, volume=2.5e-5, lam=1.0)
    _, flags = orc.fixed_p2g_f32(one, prm)
    assert flags & 65536
    _, flags = orc.fixed_p2g_f32(np.tile(rec, (4, 1)), prm)
    assert flags == 0


def test_binning_matches_brute_force(orc, pkg):
    arr, config = pkg.scenes.water_oil_column(cells=12)
    st = pkg.layout.validate_input(arr, config)
    prm = orc.params_f32(12, 12, 1.0, 0.02, 1.0, 1e-3, len(arr))
    key, start, perm, flags = orc.bin_particles(st.packed, prm)
    assert flags == 0
    pos = st.packed.view(np.float32)[:, :2].astype(np.float64)
    for k in range(prm.node_count):
        bx, by = k % 15 - 1, k // 15 - 1
        brute = [i for i in range(len(arr))
                 if math.floor(np.float32(pos[i, 0]) - 0.5) == bx and math.floor(np.float32(pos[i, 1]) - 0.5) == by]
        assert list(perm[start[k]:start[k + 1]]) == brute


GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "golden.npz"
GOLDEN_COLS = ("x", "y", "vx", "vy", "c00", "c01", "c10", "c11", "f00", "f01", "f10", "f11")


@pytest.mark.parametrize("name", ["five_1", "five_9", "five_45"])
def test_oracle_reproduces_golden_transfer_fixtures(orc, name):
    """The committed five-material fixtures (tests/golden/make_golden.py):
    8 FP64 steps reproduce the stored states."""
    g = np.load(GOLDEN)
    cfg = orc.make_config(1e-3, 48, 16, 1.0, 0.1, 0.05)
    out = orc.run_f64(orc.particles_from_packed(g[name + "_in"]), cfg, 8)
    got = np.stack([out[k] for k in GOLDEN_COLS], axis=1)
    assert np.max(np.abs(got - g[name + "_out"])) <= 1e-12


def test_oracle_reproduces_golden_fixed_point_words(orc):
    g = np.load(GOLDEN)
    prm = orc.params_f32(16, 16, 1.0, 0.02, 1.0, 1e-3, 3)
    words, flags = orc.fixed_p2g_f32(g["fixed3_in"], prm)
    assert flags == 0 and np.array_equal(words, g["fixed3_words"])
    # three particles of mass 0.02 = particleMass: the mass words of each sum to
    # 2^24 up to one rounding per support node
    assert abs(int(words[:, 3].sum()) - 3 * 2 ** 24) <= 27


def test_oracle_reproduces_golden_tank_macrostep(orc):
    g = np.load(GOLDEN)
    dt, steps, cell, mass, volume, gy = (float(v) for v in g["tank_cfg"])
    cfg = orc.make_config(dt, 80, 80, cell, mass, volume, (0.0, gy), 1.0)
    out = orc.run_f64(orc.particles_from_packed(g["tank_in"]), cfg, int(steps))
    got = np.stack([out[k] for k in GOLDEN_COLS], axis=1)
    assert np.max(np.abs(got - g["tank_out"])) <= 1e-12


@pytest.mark.parametrize("threads", [1, 3, 4])
def test_all_core_baseline_equals_the_serial_oracle(orc, pkg, threads):
    """orc_run_f64_parallel (the benchmark's all-core CPU baseline) steps the
    same scene as orc_run_f64, P2G sums reordered by row bands: equal to
    rounding, and an invalid particle fails it the same way."""
    rng = np.random.default_rng(11)
    arr, config = pkg.scenes.water_oil_column(cells=40)
    parts = orc.particles_from_packed(pkg.layout.validate_input(arr, config).packed)
    # a few solids among the liquids, to run every constitutive branch
    for k, model in enumerate((0, 1, 4)):
        idx = rng.choice(len(parts), 30, replace=False) + 0
        parts["model"][idx] = model
        parts["lambda_"][idx], parts["mu"][idx], parts["param"][idx] = 40.0, 20.0, 0.3 if model == 4 else 0.0
        parts["volume"][idx] = 1.0
    c = orc.make_config(1e-3, 40, 40, 1.0, 0.02, 1.0, (0.0, -9.81))
    serial = orc.run_f64(parts, c, 4)
    par = orc.run_f64_parallel(parts, c, 4, threads)
    for name in ("x", "y", "vx", "vy", "c00", "c01", "c10", "c11", "f00", "f01", "f10", "f11", "plastic_volume"):
        a, b = serial[name], par[name]
        assert np.allclose(a, b, rtol=1e-11, atol=1e-11 * max(1.0, float(np.abs(a).max()))), name
    assert np.array_equal(serial["model"], par["model"])
    bad = parts.copy()
    bad["x"][5] = -1.0  # outside the domain
    with pytest.raises(orc.OracleError):
        orc.run_f64_parallel(bad, c, 1, threads)
