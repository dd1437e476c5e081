"""GPU: step-path branches that the scene-level parity tests do not reach,
each checked through the C-ABI against the oracle or an exact restatement.

* drag (finishGridVelocity, kernels/shaders.ts:1748-1777): node velocities
  and the stepped particle against the FP64 oracle's drag;
* the maximum-speed reduction (kernels/shaders.ts:2463-2533): the compact
  readback's maxSpeed equals max |v| of the read-back records, bit for bit;
* the deformation retry (kernels/shaders.ts:2274-2286): a particle whose
  one-substep F update loses its orientation to float32 rounding is retried
  with two substeps, and its F equals a float32 restatement of that retry;
  one that loses it at every count up to 64 raises the deformation flag with
  the last trial's diagnostics, and one that needs more than 64 substeps the
  substep-limit flag with its index (kernels/shaders.ts:2253-2296);
* batches: appending a solid to one scene of a liquid-only batch gives every
  scene the stress cache, bit for bit the scenes stepped alone; fracture
  bonds on a batch scene are refused.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32


def particles_lone(pkg, x, y, F=(1.0, 0.0, 0.0, 1.0), C=(0.0, 0.0, 0.0, 0.0), v=(0.0, 0.0), model=0,
                   lam=20.0, mu=10.0, density=0.0, volume=1.0, param=0.0):
    n = len(x)
    return pkg.layout.ParticleArrays(
        x=np.asarray(x, np.float64), y=np.asarray(y, np.float64), vx=np.full(n, v[0]), vy=np.full(n, v[1]),
        model=np.full(n, model, np.uint32), lambda_=np.full(n, lam), mu=np.full(n, mu),
        density=np.full(n, density), volume=np.full(n, volume), param=np.full(n, param),
        C=np.tile(np.asarray(C, np.float64), (n, 1)), F=np.tile(np.asarray(F, np.float64), (n, 1)),
        object_id=np.zeros(n, np.uint32))


# ------------------------------------------------------------------ drag

@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_drag_matches_the_oracle(pkg, orc, mode):
    """A 6x6 block of Neo-Hookean particles moving at (0.6, -0.3) with the
    drag target inside it (stiffness 20, damping 8: MPM_MLS_GPU_DRAG_*,
    kernels/solver.ts:379-380).  One microstep: every node velocity and the
    particles against the oracle's FP64 step with the same drag."""
    xs, ys = np.meshgrid(np.arange(6) * 0.5 + 20.25, np.arange(6) * 0.5 + 20.25)
    arr = particles_lone(pkg, xs.ravel(), ys.ravel(), v=(0.6, -0.3), volume=0.25)
    config = {"grid": {"cellsX": 48, "cellsY": 48, "cellSize": 1.0}, "particleMass": 0.25, "particleVolume": 0.25,
              "material": {"model": "neo-hookean", "lambda": 20.0, "mu": 10.0}}
    dt = 1e-3
    target, radius = (21.6, 21.3), 1.7
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    s = factory(arr, config, device=0)
    try:
        s.ctx.set_interaction([target[0], target[1], radius, 20.0, 8.0, 0.0, 0.0, 0.0])
        s.step(dt, {"x": 0.0, "y": 0.0})
        s.check_numerical_status()
        grid = s.read_grid()
        got, _ = s.read_particles(packed=True)
    finally:
        s.dispose()
    state = pkg.layout.validate_input(arr, config)
    cfg = orc.make_config(float(F32(dt)), 48, 48, 1.0, 0.25, 0.25, (0.0, 0.0), 0.0,
                          drag=(float(F32(target[0])), float(F32(target[1])), float(F32(radius)), 20.0, 8.0))
    inp = orc.particles_from_packed(state.packed)
    g64 = orc.grid_f64(inp, cfg)
    G = orc.GRID_FIELDS
    touched = g64[:, G.index("m")] > 0
    want_v = g64[touched][:, [G.index("vx"), G.index("vy")]]
    have_v = grid.reshape(-1, 8)[touched][:, 0:2].astype(np.float64)
    W = 48 + 3
    keys = np.nonzero(touched)[0]
    node_xy = np.stack([(keys % W - 1).astype(float), (keys // W - 1).astype(float)], axis=1)
    dragged = np.hypot(node_xy[:, 0] - target[0], node_xy[:, 1] - target[1]) < radius
    assert 0 < dragged.sum() < len(keys)
    err = np.abs(have_v - want_v).max()
    print(f"drag ({mode}): {dragged.sum()} of {len(keys)} nodes dragged; max |dv_node| {err:.2e}")
    # float32 sums; the fixed-point path adds the 2^-24 quantum of each
    # contribution (relative to particleMass), divided by node masses as
    # small as 1/64 of a particle's
    tol = 1e-5 if mode == "gather" else 1e-4
    assert err <= tol * max(1.0, float(np.abs(want_v).max()))
    # the drag is visible: the dragged nodes moved away from the undragged velocity
    plain = orc.grid_f64(inp, orc.make_config(float(F32(dt)), 48, 48, 1.0, 0.25, 0.25))
    moved = np.abs(want_v - plain[touched][:, [G.index("vx"), G.index("vy")]]).max(axis=1)
    assert moved[dragged].min() > 1e-4 and moved[~dragged].max() == 0.0
    exp, _ = orc.step_f64(inp, cfg)
    have = got.view(np.float32)[:, :4].astype(np.float64)
    want = np.stack([exp[k] for k in ("x", "y", "vx", "vy")], axis=1)
    assert np.abs(have - want).max() <= tol


# ------------------------------------------------------------- max speed

def test_max_speed_is_max_norm_of_the_records(pkg):
    """measureParticleSpeedPartials + reduceParticleSpeedPartials: the
    compact readback's maxSpeed is the largest float32 |v| = sqrt(vx vx +
    vy vy) over the committed particles, bit for bit."""
    particles, config, macro, _ = pkg.scenes.tait_tank()
    for steps in (1, 40, 123):
        s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0)
        try:
            s.step_batch(steps, macro / 224, pkg.scenes.TAIT_TANK_GRAVITY, measure_speed=True)
            speed, done, status = s.ctx.read_compact()
            recs, _ = s.read_particles(packed=True)
        finally:
            s.dispose()
        assert status[0] == 0 and done == steps
        v = recs.view(np.float32)[:, 2:4]
        norm = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])  # float32 throughout
        print(f"max speed after {steps} microsteps: {speed!r} (records: {float(norm.max())!r})")
        assert F32(speed) == norm.max() and speed > 0


# ------------------------------------------------------ deformation retry

def _round32(q: Fraction) -> F32:
    """A rational rounded to the nearest float32 (ties to even), normal range."""
    if q == 0:
        return F32(0.0)
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    ulp = Fraction(2) ** (e - 23)
    m = q / ulp
    lo = m.numerator // m.denominator
    rem = m - lo
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2 == 1):
        lo += 1
    return F32(sign * float(lo * ulp))


def fma32(a, b, c) -> F32:
    return _round32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def integrate32(F, C, dt, substeps):
    """integrateDeformation in float32 (kernels/shaders.ts:1459-1483): F <-
    (I + h C) F, substeps times, each entry one fma over one product."""
    h = dt if substeps == 1 else F32(dt) / F32(substeps)
    u = [F32(1) + F32(h * C[0]), F32(h * C[1]), F32(h * C[2]), F32(1) + F32(h * C[3])]
    r = [F32(v) for v in F]
    for _ in range(substeps):
        r = [fma32(u[0], r[0], F32(u[1] * r[2])), fma32(u[0], r[1], F32(u[1] * r[3])),
             fma32(u[2], r[0], F32(u[3] * r[2])), fma32(u[2], r[1], F32(u[3] * r[3]))]
    return r


def det64(m):
    return float(m[0]) * float(m[3]) - float(m[1]) * float(m[2])


def project_liquid32(m):
    """projectLiquidDeformation in float32 (kernels/shaders.ts:474-488): the
    rotation of m scaled to the square root of its clamped compensated
    determinant (determinant2x2, :261-266)."""
    nx, ny = F32(m[0] + m[3]), F32(m[2] - m[1])
    sc = max(abs(nx), abs(ny))
    ax, ay = F32(nx / sc), F32(ny / sc)
    ln = np.sqrt(F32(F32(ax * ax) + F32(ay * ay)))
    rx, ry = F32(ax / ln), F32(ay / ln)
    sp = F32(m[1] * m[2])
    spe = fma32(m[1], m[2], -sp)
    vol = F32(fma32(m[0], m[3], -sp) - spe)
    s = np.sqrt(F32(min(max(vol, F32(0.5)), F32(2.0))))
    return [F32(s * rx), F32(s * -ry), F32(s * ry), F32(s * rx)]


def find_retry_case(F, dt):
    """A velocity gradient C (entries k / 2^16, so the lone particle's P2G and
    G2P reproduce it exactly) whose one-substep update of F has a
    non-positive determinant and whose two-substep update has a positive one."""
    rng = np.random.default_rng(7)
    for _ in range(2000):
        C = [F32(int(k) / 65536.0) for k in rng.integers(-2 ** 15, 2 ** 15, 4)]
        rowsum = max(abs(C[0]) + abs(C[1]), abs(C[2]) + abs(C[3]))
        if np.ceil(F32(F32(dt * rowsum) * F32(4.0))) > 1:
            continue
        if det64(integrate32(F, C, dt, 1)) <= 0 and det64(integrate32(F, C, dt, 2)) > 0:
            return C
    raise AssertionError("no retry case found")


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_deformation_retry_doubles_the_substeps(pkg, mode):
    """A water particle at rest on a node line (stencil fraction 1, dx 1, so
    every P2G/G2P sum is exact) with F = [[2^14, 2^14 - 1], [1, 1]] (det 1,
    so no pressure) and a chosen C.  Its one-substep F update rounds to a
    non-positive determinant; the kernel retries with two substeps, which
    keep the orientation, so the step commits with no flag and F equals the
    liquid projection of the two-substep update restated in float32, to 2
    ulp: the kernel's projection divides and takes square roots with the
    hardware reciprocal (one quotient correction) and square root, each
    within an ulp, where the restatement rounds correctly.  Without the
    retry the step would raise the deformation flag (2)."""
    a = F32(2 ** 14)
    F = [a, F32(a - F32(1)), F32(1), F32(1)]
    dt = F32(0.125)
    C = find_retry_case(F, dt)
    arr = particles_lone(pkg, [16.0], [16.0], F=[float(v) for v in F], C=[float(v) for v in C], model=2, lam=1.0,
                         mu=0.0, density=1.0, volume=1.0)
    config = {"grid": {"cellsX": 32, "cellsY": 32, "cellSize": 1.0}, "particleMass": 1.0, "particleVolume": 1.0,
              "material": {"model": "water", "lambda": 1.0, "mu": 0.0, "density": 1.0}}
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    print(f"retry case ({mode}): C = {[float(c) for c in C]}", flush=True)
    s = factory(arr, config, device=0)
    try:
        s.step(float(dt), {"x": 0.0, "y": 0.0})
        status = s.ctx.read_status()
        got, _ = s.read_particles(packed=True)
    finally:
        s.dispose()
    assert status[0] == 0, f"status {status}"
    out = got.view(np.float32)[0]
    assert list(out[4:8]) == C, "the lone particle's C should come back exactly"
    assert list(out[0:4]) == [16.0, 16.0, 0.0, 0.0]
    want = np.array(project_liquid32(integrate32(F, C, dt, 2)), np.float32)
    print(f"F out {list(out[8:12])} vs restated {list(want)}")
    assert np.all(np.abs(out[8:12] - want) <= 2 * np.spacing(np.abs(want))), (out[8:12], want)


def _lone_water(pkg, F, C):
    arr = particles_lone(pkg, [16.0], [16.0], F=[float(v) for v in F], C=[float(v) for v in C], model=2, lam=1.0,
                         mu=0.0, density=1.0, volume=1.0)
    config = {"grid": {"cellsX": 32, "cellsY": 32, "cellSize": 1.0}, "particleMass": 1.0, "particleVolume": 1.0,
              "material": {"model": "water", "lambda": 1.0, "mu": 0.0, "density": 1.0}}
    return arr, config


def _step_and_fail(pkg, mode, arr, config, dt):
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    s = factory(arr, config, device=0)
    try:
        before, _ = s.ctx.read_particles()
        s.step(float(dt), {"x": 0.0, "y": 0.0})
        with pytest.raises(pkg.MpmMlsGpuNumericalError) as e:
            s.check_numerical_status()
        after, _ = s.ctx.read_particles()
    finally:
        s.dispose()
    assert np.array_equal(before, after), "a failed microstep should leave the particles as they were"
    return e.value


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_deformation_substep_limit_names_the_particle(pkg, mode):
    """The lone water particle with C = [[200, 0], [0, 0]] and dt 1/8: the
    deformation increment dt * max row sum is 25, which needs
    ceil(25 / 0.25) = 100 substeps, above the 64 allowed
    (kernels/shaders.ts:2253-2262).  The step raises the substep-limit flag
    (32) with the particle's index and no diagnostics, and is rolled back."""
    F = [F32(1), F32(0), F32(0), F32(1)]
    arr, config = _lone_water(pkg, F, [F32(200), F32(0), F32(0), F32(0)])
    err = _step_and_fail(pkg, mode, arr, config, 0.125)
    assert err.flags == 32, err
    assert err.particle_index == 0
    assert err.diagnostics is None and err.projected_deformation is None and err.cfl_details is None


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_deformation_failure_after_every_retry_reports_the_last(pkg, mode):
    """The retry case's F = [[2^14, 2^14 - 1], [1, 1]] with a C (entries
    k / 2^16, so the lone particle returns it exactly) whose F update loses
    its orientation at every substep count 1, 2, ..., 64.  The kernel doubles
    the substeps up to 64 and then raises the deformation flag (2) with the
    particle's index and the diagnostics of the last trial
    (kernels/shaders.ts:2270-2296): its determinant (the FP64 determinant of
    the 64-substep float32 update, rounded to float32), the increment
    dt * max row sum, and det F before the step (1)."""
    a = F32(2 ** 14)
    F = [a, F32(a - F32(1)), F32(1), F32(1)]
    dt = F32(0.125)
    C = [F32(k / 65536.0) for k in (19016, -409, -21789, 15378)]
    rowsum = max(abs(C[0]) + abs(C[1]), abs(C[2]) + abs(C[3]))
    assert np.ceil(F32(F32(dt * rowsum) * F32(4.0))) == 1
    dets = [det64(integrate32(F, C, dt, n)) for n in (1, 2, 4, 8, 16, 32, 64)]
    assert all(d <= 0 for d in dets), dets
    arr, config = _lone_water(pkg, F, C)
    err = _step_and_fail(pkg, mode, arr, config, dt)
    assert err.flags == 2, err
    assert err.particle_index == 0
    d = err.diagnostics
    assert d is not None, err
    print(f"deformation failure ({mode}): {d} vs restated determinant {dets[-1]}")
    assert d["determinant"] == float(F32(dets[-1]))
    assert d["deformationIncrement"] == float(F32(dt * rowsum))
    assert d["precedingDeterminant"] == 1.0


# ------------------------------------------------------------------ batches

def _column(pkg, cells, seed):
    arr, config = pkg.scenes.water_oil_column(cells=cells, seed=seed)
    config = dict(config, capacity=cells * cells + 64)
    return arr, config


def _solids(pkg, n=12, base=(10.25, 30.25)):
    xs = base[0] + 0.5 * (np.arange(n) % 4)
    ys = base[1] + 0.5 * (np.arange(n) // 4)
    return pkg.layout.ParticleArrays(
        x=xs, y=ys, vx=np.full(n, 0.5), vy=np.zeros(n), model=np.zeros(n, np.uint32),
        lambda_=np.full(n, 2000.0), mu=np.full(n, 1000.0), density=np.zeros(n), volume=np.full(n, 1.0),
        param=np.zeros(n), object_id=np.full(n, 7, np.uint32))


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_append_of_solids_to_a_batch_scene_upgrades_the_batch(pkg, mode):
    """Two liquid-only columns batched together (no stress cache), 4
    microsteps; then Neo-Hookean particles are appended to the second scene,
    which needs the stress cache and the solid kernel variants for the whole
    batch; 4 more microsteps.  Each scene must equal the same scene stepped
    alone with the same append, bit for bit."""
    scenes = [_column(pkg, 40, 11), _column(pkg, 48, 12)]
    add = _solids(pkg)
    dt, g = 1.0 / 9600.0, pkg.scenes.WATER_OIL_GRAVITY
    p2g = pkg.native.MPM_P2G_GATHER if mode == "gather" else pkg.native.MPM_P2G_FIXED_POINT
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    print(f"batch append ({mode})", flush=True)
    b = pkg.MpmMlsGpuBatch(scenes, p2g_mode=p2g, device=0)
    try:
        b.step_batch(4, dt, g)
        b.scene(1).append_particles(pkg.layout.unpack_particles(
            pkg.layout.validate_input(add, dict(scenes[1][1], capacity=len(add))).packed))
        b.step_batch(4, dt, g)
        got = []
        for i in range(2):
            b.scene(i).check_numerical_status()
            got.append(b.scene(i).read_particles(packed=True)[0])
    finally:
        b.dispose()
    for i, (arr, config) in enumerate(scenes):
        s = factory(arr, config, device=0)
        try:
            s.step_batch(4, dt, g)
            if i == 1:
                s.append_particles(pkg.layout.unpack_particles(
                    pkg.layout.validate_input(add, dict(config, capacity=len(add))).packed))
            s.step_batch(4, dt, g)
            s.check_numerical_status()
            want = s.read_particles(packed=True)[0]
        finally:
            s.dispose()
        assert np.array_equal(got[i], want), f"scene {i}"
    assert len(got[1]) == 48 * 48 + len(add)


def test_fracture_bonds_on_a_batch_scene_are_refused(pkg):
    b = pkg.MpmMlsGpuBatch([_column(pkg, 16, 3)], device=0)
    try:
        bonds = np.zeros((1, 6), np.uint32)
        bonds[0, :2] = (0, 1)
        bonds[0, 2:4] = np.array([1.0, 0.1], np.float32).view(np.uint32)
        with pytest.raises(Exception, match="batch"):
            b.scene(0).ctx.set_fracture_bonds(bonds)
        b.step_batch(2, 1.0 / 9600.0, pkg.scenes.WATER_OIL_GRAVITY)
        b.scene(0).check_numerical_status()
    finally:
        b.dispose()


# ------------------------------------------------------------- Node addon

def test_node_addon_serialises_a_step_behind_a_pending_readback(pkg, tmp_path):
    """The reference lets its caller queue work while a mapAsync readback is
    pending.  Through the addon: stepBatch(4), readCompact() (a worker thread
    waits for the GPU), stepBatch(4) at once from the JS thread, readCompact()
    again.  The calls serialise on the context's lock, so the end state is
    eight microsteps, bit for bit the Python path's."""
    import json
    import pathlib
    import shutil
    import subprocess
    root = pathlib.Path(__file__).resolve().parent.parent
    if shutil.which("node") is None or not (root / "powder-play_amd" / "mpm355.node").exists():
        pytest.skip("node or the built Node-API addon is missing")
    arr, config = pkg.scenes.water_oil_column(cells=200)
    state = pkg.layout.validate_input(arr, config)
    g = pkg.scenes.WATER_OIL_GRAVITY
    params = pkg.layout.pack_step_parameters(state, 1.0 / 9600.0, g, len(arr), 0, None)
    (tmp_path / "packed.bin").write_bytes(np.ascontiguousarray(state.packed, np.uint32).tobytes())
    (tmp_path / "params.bin").write_bytes(params)
    script = """
const fs = require('fs');
const { MpmNativeRuntime } = require('./powder-play_amd/napi/index.js');
const dir = process.argv[1];
const b = fs.readFileSync(dir + '/packed.bin');
const packed = new Uint32Array(b.length / 4);
new Uint8Array(packed.buffer).set(b);
const rt = new MpmNativeRuntime({cellsX: 200, cellsY: 200, cellSize: 1, particleMass: 0.02, particleVolume: 1,
                                 device: 0}, packed, null, false);
const params = new Uint8Array(fs.readFileSync(dir + '/params.bin'));
rt.stepBatch(params, 4, true);
const first = rt.readCompact();
rt.stepBatch(params, 4, true);
const second = rt.readCompact();
Promise.all([first, second]).then(([a, c]) => {
  const r = rt.readParticles();
  fs.writeFileSync(dir + '/out.bin', Buffer.from(r.records.buffer));
  console.log(JSON.stringify({ first: a.completed, second: c.completed, status: [a.status[0], c.status[0]] }));
});
"""
    print("node addon: step queued behind a pending readback", flush=True)
    r = subprocess.run(["node", "-e", script, str(tmp_path)], cwd=root, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout)
    assert res["status"] == [0, 0] and res["second"] == 4
    node_recs = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.uint32).reshape(-1, 20)
    s = pkg.create_mpm_mls_gpu_solver(arr, config, device=0)
    try:
        s.step_batch(4, 1.0 / 9600.0, g, measure_speed=True)
        s.step_batch(4, 1.0 / 9600.0, g, measure_speed=True)
        recs, _ = s.read_particles(packed=True)
    finally:
        s.dispose()
    assert np.array_equal(node_recs, recs)
