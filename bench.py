"""Benchmark of the MLS-MPM step path on one MI355X.

Workload: the 1,000,000-particle water/oil column (1000x1000 cells of size 1,
60% water / 40% oil, Gate 4 liquids, velocities U(+-2), gravity 90), stepped
at the CFL microstep of a 1/240 s macrostep (the plan of its first
macrostep) with the CFL guard on, one batch per macrostep.  A "step" is one
microstep over all particles; value = particles x committed microsteps / s,
inputs resident in HBM, the timed region bracketed by stream synchronisation.

Printed JSON line (one): metric/value/unit of the float32 gather path plus
the fixed-point path's rate, the roofline of the dominant kernel (HIP events
on the solver's stream; algorithmic bytes per launch), the batched-scenes and
differentiable-rollout legs, the CPU baseline (the FP64 oracle step on a
bounded sample of the same scene, on every host CPU the process may use and
on one thread; and the FP64 oracle adjoint of the elastic block), and the
measured error of every configuration against the FP64 oracle ("parity",
the oracle as checker beside the CPU baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 256 << 20  # calibration copy of the counter runs
METRIC = "particle-steps/s @1M particles + %HBM peak; batched & adjoint vs plain step"
# Algorithmic bytes per microstep (float32 SoA): per particle, read x, v, C,
# F, volume and material id (56 B) and write x, v, C, F (48 B); per grid node
# and liquid phase present, 56 B (P2G writes m, qx, qy, V: 16; the grid
# update reads them, 16, and writes v, 8; G2P reads v, m, V, 16).  Each
# kernel is charged its part of that budget: the fused P2G + grid update the
# particle reads and 40 B per node and phase, G2P the particle writes and the
# 16 B per node and phase it reads.
PARTICLE_BYTES = 104
NODE_PHASE_BYTES = 56
P2G_PARTICLE_BYTES, P2G_NODE_PHASE_BYTES = 56, 40
G2P_PARTICLE_BYTES, G2P_NODE_PHASE_BYTES = 48, 16
TRAFFIC_PROFILE = ROOT / "profiles" / "pmc_traffic_water_oil_1m.md"


def measured_traffic(tuning: dict | None):
    """HBM bytes per launch of each step kernel from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the calibration copy), when
    it was measured with the launch shape this run uses; else an empty dict."""
    import re
    if not TRAFFIC_PROFILE.exists() or not tuning:
        return {}
    shape = ",".join(str(tuning[k]) for k in ("p2g_tile", "g2p_tile", "g2p_threads", "bucket_capacity"))
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            m = re.search(r"shape=([\d,]+)", line)
            if not m or m.group(1) != shape:
                return {}
            return {k: int(v) for k, v in re.findall(r"(\w+)=(\d+)", line) if k != "shape"}
    return {}


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def build_solver(pkg, arr, config, mode, timing=False, tuning=None):
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    s = factory(arr, config, device=0, tuning=tuning)
    if timing:
        s.ctx.set_timing(True)
    return s


def plan_for(pkg, solver, config):
    b = solver.cfl_bounds
    macro = pkg.scenes.WATER_OIL_MACRO_DT
    bound = pkg.cfl.calculate_velocity_bound(b.velocityBound, macro, 90.0)
    plan = pkg.cfl.calculate_microsteps(macro, config["grid"]["cellSize"], b.waveSpeedBound, bound, 1024)
    limit = pkg.cfl.calculate_velocity_limit(config["grid"]["cellSize"], b.waveSpeedBound, plan.microDt)
    return plan, limit


def run_steps(pkg, solver, plan, limit, steps):
    """Queue `steps` guarded microsteps in batches of one macrostep."""
    params = pkg.layout.pack_step_parameters(solver.state, plan.microDt, pkg.scenes.WATER_OIL_GRAVITY,
                                             solver.allocated_particle_count, 0,
                                             {"waveSpeedBound": solver.cfl_bounds.waveSpeedBound,
                                              "velocityBound": limit})
    left = steps
    while left > 0:
        k = min(left, plan.stepCount)
        solver.ctx.step_batch(params, k, True)
        left -= k


def timed(pkg, solver, plan, limit, steps, warmup):
    run_steps(pkg, solver, plan, limit, warmup)
    solver.ctx.read_compact()  # synchronise
    t0 = time.perf_counter()
    run_steps(pkg, solver, plan, limit, steps)
    speed, done, status = solver.ctx.read_compact()
    dt = time.perf_counter() - t0
    if status[0] != 0:
        raise RuntimeError(f"benchmark step failed with status {list(status)}")
    return dt, speed


def host_cpus() -> dict:
    """The host CPUs this process may use: the affinity mask, nproc, the
    lscpu model name, and the OpenMP thread budget of the environment."""
    import subprocess
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        info["nproc"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        fields = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        info["model"] = fields.get("Model name", "").strip() or None
        info["sockets"] = fields.get("Socket(s)", "").strip() or None
        info["threads_per_core"] = fields.get("Thread(s) per core", "").strip() or None
    except (OSError, subprocess.SubprocessError):
        info["model"] = None
    return info


def baseline_threads(info: dict) -> int:
    """Threads of the all-core CPU baseline: every CPU in the affinity mask,
    capped by OMP_NUM_THREADS when it is set (the CPU share the host grants
    one GPU's jobs)."""
    n = info["affinity_cpus"]
    env = info.get("omp_num_threads") or ""
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(pkg, arr, config, budget_s=12.0, threads=1):
    """The FP64 oracle step on the same scene: one thread (orc_run_f64) or
    `threads` OpenMP threads (orc_run_f64_parallel).  Returns
    (particle-steps/s, steps, seconds)."""
    from oracle import oracle as orc
    state = pkg.layout.validate_input(arr, config)
    parts = orc.particles_from_packed(state.packed)
    cfg = orc.make_config(1.0 / 9600.0, config["grid"]["cellsX"], config["grid"]["cellsY"], 1.0, 0.02, 1.0,
                          (0.0, 90.0))

    def step(p):
        return orc.run_f64(p, cfg, 1) if threads == 1 else orc.run_f64_parallel(p, cfg, 1, threads)

    parts = step(parts)  # warm-up step
    steps = 0
    t0 = time.perf_counter()
    while True:
        parts = step(parts)
        steps += 1
        el = time.perf_counter() - t0
        if (steps >= 3 and el > budget_s * 0.5) or el > budget_s:
            break
    return len(arr) * steps / el, steps, el


def cfl_macrosteps(pkg, solver, macro_dt, gravity, count):
    """`count` macrosteps through the adaptive CFL loop (stepCflBatchAdaptive,
    solver.ts:1724-2021), each planned from the bound the previous one
    recommends, as the reference's page loop does.  Returns the microsteps
    run."""
    cell = solver.state.config["grid"]["cellSize"]
    g = float(np.hypot(gravity["x"], gravity["y"]))
    bound = max(pkg.cfl.calculate_velocity_bound(solver.cfl_bounds.velocityBound, macro_dt, g),
                solver._last_scheduled_velocity_bound or 0.0)
    micro = 0
    for _ in range(count):
        plan = pkg.cfl.calculate_microsteps(macro_dt, cell, solver.cfl_bounds.waveSpeedBound, bound, 1024)
        m = solver.step_cfl_batch_adaptive(1, macro_dt, plan, bound, 1024, gravity)
        micro += m.attemptedMicroStepCount
        # the next bound covers what the GPU measured and never drops below
        # the envelope the solver last scheduled (a retry or continuation
        # inside the adaptive loop may have raised it)
        bound = max(bound, pkg.cfl.calculate_velocity_bound(m.recommendedVelocityBound, macro_dt, g),
                    solver._last_scheduled_velocity_bound or 0.0)
    solver.check_numerical_status()
    return micro


def small_scenes_leg(pkg, macrosteps: int = 12):
    """The reference's own scenes, which are launch-bound at this size: the SI
    Tait tank (1,600 particles, 1/120 s macrosteps, both P2G modes) and the
    Gate 4 scene (3,674 particles with bonded pebbles and a fracture solid),
    on the reference's benchmark grid with fixed 1/240 s steps and on the
    MPM page's 220 x 600 grid with CFL macrosteps.  Each is stepped through the adaptive CFL loop with
    its host work and status readback per macrostep, after one warm-up
    macrostep; rate = particles x microsteps / wall seconds."""
    out = {}
    particles, config, macro_dt, _ = pkg.scenes.tait_tank()
    n = len(particles)
    for mode in ("gather", "fixed"):
        factory = (pkg.create_mpm_mls_gpu_solver if mode == "gather"
                   else pkg.create_experimental_mpm_mls_atomic_gpu_solver)
        s = factory(particles, config, device=0)
        try:
            cfl_macrosteps(pkg, s, macro_dt, pkg.scenes.TAIT_TANK_GRAVITY, 1)
            t0 = time.perf_counter()
            micro = cfl_macrosteps(pkg, s, macro_dt, pkg.scenes.TAIT_TANK_GRAVITY, macrosteps)
            el = time.perf_counter() - t0
        finally:
            s.dispose()
        out[f"tait_tank_{mode}"] = {"value": n * micro / el, "unit": "particle-steps/s", "particles": n,
                                    "macrosteps": macrosteps, "microsteps": micro,
                                    "ms_per_microstep": 1e3 * el / micro, "ms_per_macrostep": 1e3 * el / macrosteps}
        log(f"[bench] Tait tank ({mode}): {micro} microsteps in {1e3 * el:.1f} ms, {n * micro / el:.3e} particle-steps/s")
    # Gate 4 with the reference's own batch protocol (docs/PERFORMANCE-MLS-MPM.md:
    # 622-636): the 220 x 150 grid of cell size 10, liquid area 100, fixed
    # dt = 1/240 and gravity (0, 90); 64 warm-up steps, then 12 batches of 8
    # steps, each synchronised; median and p95 per step
    particles, config, gravity = pkg.scenes.gate4_scene(mpm_grid=False, liquid_area=100.0)
    s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0)
    try:
        n = s.particle_count
        for _ in range(8):
            s.step_batch(8, 1.0 / 240.0, gravity)
        s.ctx.read_compact()
        per_step = []
        for _ in range(12):
            t0 = time.perf_counter()
            s.step_batch(8, 1.0 / 240.0, gravity)
            s.ctx.read_compact()
            per_step.append((time.perf_counter() - t0) / 8)
        s.check_numerical_status()
    finally:
        s.dispose()
    med = float(np.median(per_step))
    out["gate4"] = {"value": n / med, "unit": "particle-steps/s", "particles": n,
                    "ms_per_step_median": 1e3 * med, "ms_per_step_p95": 1e3 * float(np.quantile(per_step, 0.95)),
                    "protocol": "220x150 grid, cell 10, liquid area 100, dt 1/240, gravity (0, 90); 64 warm-up "
                                "steps, 12 batches of 8 steps, wall clock per batch including its synchronisation"}
    log(f"[bench] Gate 4: {1e3 * med:.4f} ms per step (median of 12 batches of 8), {n / med:.3e} particle-steps/s")
    # Gate 4 as the MPM page runs it: the 220 x 600 grid of cell size 10,
    # liquid area 1,600, 1/240 s macrosteps through the adaptive CFL loop
    particles, config, gravity = pkg.scenes.gate4_scene()
    s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0)
    try:
        n = s.particle_count
        cfl_macrosteps(pkg, s, 1.0 / 240.0, gravity, 1)
        t0 = time.perf_counter()
        micro = cfl_macrosteps(pkg, s, 1.0 / 240.0, gravity, macrosteps)
        el = time.perf_counter() - t0
    finally:
        s.dispose()
    out["gate4_mpm_grid"] = {"value": n * micro / el, "unit": "particle-steps/s", "particles": n,
                             "macrosteps": macrosteps, "microsteps": micro, "ms_per_microstep": 1e3 * el / micro,
                             "ms_per_macrostep": 1e3 * el / macrosteps,
                             "protocol": "220x600 grid, cell 10, liquid area 1,600, gravity (0, 90); 1/240 s "
                                         "macrosteps through the adaptive CFL loop after one warm-up macrostep, "
                                         "wall clock including host planning and status readbacks"}
    log(f"[bench] Gate 4 (220x600): {micro} microsteps in {1e3 * el:.1f} ms, {n * micro / el:.3e} particle-steps/s")
    out["reference_published"] = {
        "tait_tank": "23.89e6 particle-steps/s, 1,600 particles, 423-microstep batch, Apple Metal 3 "
                     "(docs/PERFORMANCE-MLS-MPM.md:837-851); ~4.8e7 with fixed-point P2G "
                     "(docs/REPORT-FLUID-SOLVER-PROTOTYPES.md:207-229)",
        "gate4": "0.216618 ms per step (median, GPU timestamps) on Apple Metal 3, same protocol "
                 "(docs/PERFORMANCE-MLS-MPM.md:622-636); CPU reference 4.643 ms per step, one JS thread "
                 "(docs/PERFORMANCE-MLS-MPM.md:778-790)"}
    return out


def batched_leg(pkg, steps: int, how: str = "auto", sequential: bool = True):
    """256 independent Tait tanks (scenes.batched_tanks): `steps` microsteps
    of all of them in one batch, against the same scenes as separate solvers
    stepped one after another (each synchronised before the next starts)."""
    tanks = pkg.scenes.batched_tanks()
    scenes = [(a, c) for a, c, _ in tanks]
    gravity = [g for _, _, g in tanks]
    total = sum(len(a) for a, _, _ in tanks)
    dt = pkg.scenes.BATCH_DT
    log(f"[bench] batched: {len(tanks)} tanks, {total} particles, {steps} microsteps")
    if how == "auto":
        t0 = time.perf_counter()
        tuned = pkg.autotune.autotune_batch(scenes, dt, gravity, device=0)
        log(f"[bench] batched: autotuned {tuned['tuning']} in {time.perf_counter() - t0:.1f} s")
    else:
        keys = ("p2g_tile", "g2p_tile", "g2p_threads", "bucket_capacity", "tile_queue")
        tuned = {"tuning": dict(zip(keys, (int(v) for v in how.split(",")))), "candidates": None}
    batch = pkg.MpmMlsGpuBatch(scenes, device=0, tuning=tuned["tuning"])
    batch.step_batch(8, dt, gravity)
    batch.scene(0).ctx.read_compact()
    t0 = time.perf_counter()
    batch.step_batch(steps, dt, gravity)
    batch.scene(0).ctx.read_compact()  # waits for the batch's stream
    t_batch = time.perf_counter() - t0
    bad = [i for i in range(len(batch)) if batch.scene(i).ctx.read_compact()[2][0] != 0]
    if bad:
        batch.dispose()
        raise RuntimeError(f"batched scenes {bad[:8]} raised numerical flags")
    # the same batch with the FP64 mass/momentum reductions of every scene on
    for i in range(len(batch)):
        batch.scene(i).set_diagnostics(True)
    batch.step_batch(8, dt, gravity)
    batch.scene(0).ctx.read_compact()
    t0 = time.perf_counter()
    batch.step_batch(steps, dt, gravity)
    batch.scene(0).ctx.read_compact()
    t_diag = time.perf_counter() - t0
    drift = 0.0
    for i in range(len(batch)):
        d = batch.scene(i).read_diagnostics()
        drift = max(drift, abs(d["outputMass"] - d["inputMass"]) / d["inputMass"])
    batch.dispose()
    t_seq = float("nan")
    if sequential:
        solvers = [pkg.create_mpm_mls_gpu_solver(a, c, device=0) for a, c in scenes]
        for s, g in zip(solvers, gravity):
            s.step_batch(8, dt, g)
            s.ctx.read_compact()
        t0 = time.perf_counter()
        for s, g in zip(solvers, gravity):
            s.step_batch(steps, dt, g)
            s.ctx.read_compact()
        t_seq = time.perf_counter() - t0
        for s in solvers:
            s.dispose()
    rate = total * steps / t_batch
    rate_seq = total * steps / t_seq if sequential else None
    log(f"[bench] batched: {1e3 * t_batch / steps:.4f} ms per microstep of all scenes ({rate:.3e} particle-steps/s); "
        f"with FP64 reductions {1e3 * t_diag / steps:.4f} ms (mass drift {drift:.1e})"
        + (f"; sequential {1e3 * t_seq / steps:.4f} ms ({rate_seq:.3e}); x{rate / rate_seq:.1f}" if sequential else ""))
    return {"value": rate, "unit": "particle-steps/s", "ms_per_step": 1e3 * t_batch / steps, "scenes": len(tanks),
            "particles": total, "steps": steps, "sequential_value": rate_seq,
            "sequential_ms_per_step": 1e3 * t_seq / steps if sequential else None,
            "speedup_vs_sequential": rate / rate_seq if sequential else None,
            "fp64_reductions": {"value": total * steps / t_diag, "ms_per_step": 1e3 * t_diag / steps,
                                "max_relative_mass_drift": drift},
            "tuning": tuned,
            "workload": f"{len(tanks)} SI-material Tait tanks, {total} particles (log-uniform 500-50,000 per tank), "
                        "64-512 cell grids, dam-break blocks, random gravity directions, float32 gather P2G"}


def adjoint_leg(pkg, reps: int):
    """The differentiable elastic block (scenes.elastic_block: 4,096
    fixed-corotated particles, 256x256 cells, 200 steps of 1e-3, checkpoints
    every 20): device time of the forward rollout and of the reverse pass
    (HIP events on the rollout's stream around the kernel sequences; host
    copies excluded), median of `reps` forward+backward pairs after one
    warm-up pair.  The reverse pass includes the recomputation of every
    segment from its checkpoint."""
    state, lam, mu, config, target = pkg.scenes.elastic_block()
    n, T = len(lam), config["steps"]
    r = pkg.adjoint.ElasticRollout(n, config, segment_steps=20, device=0)
    fwd, bwd = [], []
    try:
        for i in range(reps + 1):
            final = r.forward(state, lam, mu)
            d = final[:, :2].astype(np.float64).mean(axis=0) - target
            bar = np.zeros_like(final)
            bar[:, :2] = 2.0 * d / n
            r.backward(bar)
            t = r.timing()
            if i:
                fwd.append(t["forward_ms"])
                bwd.append(t["backward_ms"])
        layout = r.layout()
    finally:
        r.close()
    f_ms, b_ms = float(np.median(fwd)), float(np.median(bwd))
    log(f"[bench] adjoint: forward {f_ms:.3f} ms, reverse {b_ms:.3f} ms for {T} steps of {n} particles "
        f"(reverse / forward = {b_ms / f_ms:.2f})")
    return {"workload": f"fixed-corotated elastic block, {n} particles, {config['cellsX']}x{config['cellsY']} cells, "
                        f"{T} steps of {config['dt']}, checkpoints every {layout['segment_steps']}",
            "forward_ms": f_ms, "backward_ms": b_ms, "forward_value": n * T / (f_ms * 1e-3),
            "backward_value": n * T / (b_ms * 1e-3), "unit": "particle-steps/s", "backward_over_forward": b_ms / f_ms,
            "device_bytes": layout["device_bytes"], "reps": reps}


def adjoint_cpu(pkg):
    """FP64 oracle forward + reverse of the same elastic block, one thread.
    Returns (report, (final, bar, dL/dstate0, dL/dlambda, dL/dmu))."""
    from oracle import oracle as orc
    state, lam, mu, config, target = pkg.scenes.elastic_block()
    cfg = orc.diff_config(config)
    t0 = time.perf_counter()
    final = orc.diff_forward(state, lam, mu, cfg)
    t1 = time.perf_counter()
    _, bar = orc.com_loss(final, target)
    b0, gl, gm, _ = orc.diff_backward(state, lam, mu, cfg, bar)
    t2 = time.perf_counter()
    n, T = len(lam), config["steps"]
    return ({"forward_value": n * T / (t1 - t0), "backward_value": n * T / (t2 - t1), "unit": "particle-steps/s",
             "cores": 1, "kind": "port", "sample": f"one FP64 forward and reverse of the {n}-particle block "
                                                   f"({t1 - t0:.2f} s + {t2 - t1:.2f} s; the reverse re-runs the "
                                                   "forward)"},
            (final, bar, b0, gl, gm))


def _rel_l2(a, b) -> float:
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _oracle_config(orc, config, dt, gravity):
    g = config["grid"]
    return orc.make_config(float(np.float32(dt)), g["cellsX"], g["cellsY"], g["cellSize"], config["particleMass"],
                           config["particleVolume"], (gravity["x"], gravity["y"]), config.get("wallFriction", 0.0))


def _step_error(orc, solver, config, dt, gravity, bodies=False):
    """One microstep of `solver` against one FP64 oracle microstep from the
    same (the GPU's own) input state: relative L2 error of the velocities and
    positions and the largest position error.  Without bodies, also the
    velocity error over the particles whose stencil avoids every node where
    the water/oil contact decides within float32 rounding of a threshold
    (oracle.contact_near_threshold), and how many particles touch such nodes."""
    before, oid = solver.read_particles(packed=True)
    cfg = _oracle_config(orc, config, dt, gravity)
    if bodies:
        bc = orc.BodyConfig(float(np.float32(config.get("bondRestLength", 0.0))),
                            float(np.float32(config.get("bondStiffness", 0.35))),
                            float(np.float32(config.get("bondReferenceDt", dt))))
        bonds = None
        if solver.state.fracture_bonds is not None:
            b = solver.read_fracture_bonds()
            bonds = np.zeros(len(b), orc.BOND_DTYPE)
            for k, d in enumerate(b):
                bonds[k] = (d["firstParticle"], d["secondParticle"], d["restLength"], d["failureStrain"],
                            d["damage"], 1 if d["broken"] else 0, 0)
        exp, _, _ = orc.step_bodies_f64(orc.particles_from_packed(before), oid, cfg, bc, bonds)
    else:
        ref_in = orc.particles_from_packed(before)
        exp, _ = orc.step_f64(ref_in, cfg)
        near, _ = orc.contact_near_threshold(orc.grid_f64(ref_in, cfg), cfg)
        excluded = orc.particles_touching(near, ref_in["x"], ref_in["y"], float(cfg.cell_size))
    solver.step(dt, gravity)
    solver.check_numerical_status()
    got, _ = solver.read_particles(packed=True)
    have = got.view(np.float32)[:, :4].astype(np.float64)
    want = np.stack([exp["x"], exp["y"], exp["vx"], exp["vy"]], axis=1)
    out = {"rel_l2_v": _rel_l2(have[:, 2:4], want[:, 2:4]), "rel_l2_x": _rel_l2(have[:, :2], want[:, :2]),
           "max_abs_x": float(np.abs(have[:, :2] - want[:, :2]).max())}
    if not bodies:
        keep = ~excluded
        out["rel_l2_v_away_from_contact_thresholds"] = _rel_l2(have[keep, 2:4], want[keep, 2:4])
        out["particles_at_contact_thresholds"] = int(excluded.sum())
    return out


def parity_leg(pkg, arr, config, fp64_adjoint):
    """The measured error of every configuration against the FP64 oracle (the
    oracle as checker, on the host, beside the CPU baseline).  Per microstep
    the oracle takes one FP64 step from the GPU's own input state; the bar of
    the float32 gather path is a relative L2 velocity error of 1e-4 per
    microstep, and the fixed-point grid words must be bit-exact."""
    import hashlib
    from oracle import oracle as orc
    t_start = time.perf_counter()
    out = {"bar": "per microstep from a common state: relative L2 velocity error <= 1e-4 (float32 paths) over the "
                  "particles whose stencil avoids the nodes where the water/oil contact decides within float32 "
                  "rounding of a threshold (rel_l2_v_away_from_contact_thresholds; rel_l2_v covers every "
                  "particle); int32 fixed-point grid words bit-exact; adjoint gradients relative L2 <= 1e-3"}

    # C1: SI Tait tank, first microstep and one whole 1/120 s macrostep
    particles, config1, macro_dt, _ = pkg.scenes.tait_tank()
    g1 = pkg.scenes.TAIT_TANK_GRAVITY
    c1 = {}
    for mode in ("gather", "fixed"):
        factory = (pkg.create_mpm_mls_gpu_solver if mode == "gather"
                   else pkg.create_experimental_mpm_mls_atomic_gpu_solver)
        s = factory(particles, config1, device=0)
        try:
            plan = pkg.cfl.calculate_microsteps(macro_dt, config1["grid"]["cellSize"], s.cfl_bounds.waveSpeedBound,
                                                pkg.cfl.calculate_velocity_bound(s.cfl_bounds.velocityBound, macro_dt,
                                                                                 9.81), 1024)
            first = _step_error(orc, s, config1, plan.microDt, g1)
            s.step_batch(plan.stepCount - 1, plan.microDt, g1)
            s.check_numerical_status()
            got, _ = s.read_particles(packed=True)
        finally:
            s.dispose()
        ref = orc.run_f64(orc.particles_from_packed(pkg.layout.validate_input(particles, config1).packed),
                          _oracle_config(orc, config1, plan.microDt, g1), plan.stepCount)
        have = got.view(np.float32)[:, :4].astype(np.float64)
        want = np.stack([ref["x"], ref["y"], ref["vx"], ref["vy"]], axis=1)
        c1[mode] = {"microstep_1": first,
                    "macrostep": {"microsteps": plan.stepCount,
                                  "max_abs_x_m": float(np.abs(have[:, :2] - want[:, :2]).max()),
                                  "max_abs_v_ms": float(np.abs(have[:, 2:4] - want[:, 2:4]).max()),
                                  "rel_l2_v": _rel_l2(have[:, 2:4], want[:, 2:4])}}
    out["C1_tait_tank"] = c1
    log(f"[bench] parity C1: {c1}")

    # C2: Gate 4 on the 220 x 600 grid (bodies, contact, fracture bonds)
    particles, config2, gravity2 = pkg.scenes.gate4_scene()
    s = pkg.create_mpm_mls_gpu_solver(particles, config2, device=0)
    c2 = {}
    try:
        dt2 = 1.0 / 2400.0
        c2["microstep_1"] = _step_error(orc, s, config2, dt2, gravity2, bodies=True)
        s.step_batch(8, dt2, gravity2)
        c2["microstep_10"] = _step_error(orc, s, config2, dt2, gravity2, bodies=True)
        c2["dt"] = dt2
        c2["note"] = ("max_abs_x is led by liquid particles that start inside the pebbles: the projection out of "
                      "the triangles moves them by up to ~54 units along a normal interpolated from the float32 "
                      "state, so their position error scales with that displacement")
    finally:
        s.dispose()
    out["C2_gate4"] = c2
    log(f"[bench] parity C2: {c2}")

    # C3: the 1M column at microsteps 1, 20 and 40 of a 1/240 s macrostep,
    # both modes; fixed-point grid words at microstep 20
    dt3, g3 = 1.0 / 9600.0, pkg.scenes.WATER_OIL_GRAVITY
    cells = config["grid"]["cellsX"]
    c3 = {}
    for mode in ("gather", "fixed"):
        factory = (pkg.create_mpm_mls_gpu_solver if mode == "gather"
                   else pkg.create_experimental_mpm_mls_atomic_gpu_solver)
        s = factory(arr, config, device=0, keep_fixed_grid=(mode == "fixed"))
        rows = {}
        try:
            done = 0
            for k in (1, 20, 40):
                if k - 1 > done:
                    s.step_batch(k - 1 - done, dt3, g3)
                    done = k - 1
                if mode == "fixed" and k == 20:
                    before, _ = s.read_particles(packed=True)
                    prm = orc.params_f32(cells, cells, 1.0, 0.02, 1.0, dt3, len(arr))
                    want, _ = orc.fixed_p2g_f32(before, prm)
                rows[f"microstep_{k}"] = _step_error(orc, s, config, dt3, g3)
                done = k
                if mode == "fixed" and k == 20:
                    got = s.read_fixed_grid()
                    rows["fixed_words_microstep_20"] = {
                        "words": int(got.size), "differ": int(np.count_nonzero(got != want)),
                        "sha256_gpu": hashlib.sha256(np.ascontiguousarray(got).tobytes()).hexdigest()[:16],
                        "sha256_cpu": hashlib.sha256(np.ascontiguousarray(want).tobytes()).hexdigest()[:16]}
        finally:
            s.dispose()
        c3[mode] = rows
    out["C3_column_1m"] = c3
    log(f"[bench] parity C3: {c3}")

    # C4: the 256-tank batch; after 8 batched microsteps, the smallest, a
    # median and the largest tank against one oracle microstep each
    tanks = pkg.scenes.batched_tanks()
    sizes = np.array([len(a) for a, _, _ in tanks])
    order = np.argsort(sizes)
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    batch = pkg.MpmMlsGpuBatch([(a, c) for a, c, _ in tanks], device=0)
    gs = [g for _, _, g in tanks]
    c4 = {}
    try:
        batch.step_batch(8, pkg.scenes.BATCH_DT, gs)
        before = {i: batch.scene(i).read_particles(packed=True)[0] for i in pick}
        batch.step_batch(1, pkg.scenes.BATCH_DT, gs)
        for i in pick:
            sc = batch.scene(i)
            sc.check_numerical_status()
            got = sc.read_particles(packed=True)[0]
            exp, _ = orc.step_f64(orc.particles_from_packed(before[i]),
                                  _oracle_config(orc, tanks[i][1], pkg.scenes.BATCH_DT, gs[i]))
            have = got.view(np.float32)[:, :4].astype(np.float64)
            c4[f"tank_{i}_{int(sizes[i])}_particles"] = {
                "rel_l2_v": _rel_l2(have[:, 2:4], np.stack([exp["vx"], exp["vy"]], 1)),
                "rel_l2_x": _rel_l2(have[:, :2], np.stack([exp["x"], exp["y"]], 1))}
    finally:
        batch.dispose()
    out["C4_batched"] = c4
    log(f"[bench] parity C4: {c4}")

    # C5: GPU forward and adjoint of the elastic block against the FP64 ones
    if fp64_adjoint is not None:
        final64, bar, b0, gl, gm = fp64_adjoint
        state, lam, mu, config5, _ = pkg.scenes.elastic_block()
        r = pkg.adjoint.ElasticRollout(len(lam), config5, segment_steps=20, device=0)
        try:
            final = r.forward(state, lam, mu)
            g0, ggl, ggm = r.backward(bar.astype(np.float32))
        finally:
            r.close()
        out["C5_adjoint"] = {"max_abs_x_final": float(np.abs(final[:, :2] - final64[:, :2]).max()),
                             "rel_l2_grad_x0": _rel_l2(g0[:, 0:2], b0[:, 0:2]),
                             "rel_l2_grad_v0": _rel_l2(g0[:, 2:4], b0[:, 2:4]),
                             "rel_l2_grad_lambda": _rel_l2(ggl, gl), "rel_l2_grad_mu": _rel_l2(ggm, gm)}
        log(f"[bench] parity C5: {out['C5_adjoint']}")
    out["seconds"] = round(time.perf_counter() - t_start, 1)
    return out


def choose_tuning(pkg, arr, config, mode, how):
    """The launch shape a mode is measured with: "default", "auto" (the
    autotuner on a scratch copy of the scene, before and outside every timed
    region) or explicit "p2g_tile,g2p_tile,g2p_threads,bucket_capacity".
    Returns (tuning dict or None, report)."""
    if how == "default":
        return None, {"source": "default", "tuning": dict(pkg.autotune.DEFAULT_TUNING)}
    if how != "auto":
        keys = ("p2g_tile", "g2p_tile", "g2p_threads", "bucket_capacity", "tile_queue")
        t = dict(zip(keys, (int(v) for v in how.split(","))))
        return t, {"source": "explicit", "tuning": t}
    code = pkg.native.MPM_P2G_GATHER if mode == "gather" else pkg.native.MPM_P2G_FIXED_POINT
    t0 = time.perf_counter()
    r = pkg.autotune.autotune(arr, config, code, gravity=pkg.scenes.WATER_OIL_GRAVITY, steps=20, warmup=4)
    el = time.perf_counter() - t0
    cands = {part: {k: round(v["p2g_us" if part == "p2g" else "g2p_us" if part == "g2p" else "step_us"], 2)
                    for k, v in r["candidates"][part].items()} for part in r["candidates"]}
    log(f"[bench] {mode}: autotuned {r['tuning']} in {el:.1f} s")
    return r["tuning"], {"source": "autotune", "tuning": r["tuning"], "seconds": round(el, 2),
                         "class": r["class"], "device": r["device"], "candidates_us": cands}


def pmc_sample(pkg, arr, config, mode, tuning_how="default"):
    """Workload of the FETCH_SIZE / WRITE_SIZE counter runs.  The copy kernel
    reads and writes COPY_BYTES each, so its counter values per byte scale
    the counters of the step kernels to HBM bytes."""
    import torch
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"[bench] calibration copy of {COPY_BYTES} bytes done")
    for m in (["gather", "fixed"] if mode == "both" else [mode]):
        tuning, _ = choose_tuning(pkg, arr, config, m, tuning_how)
        solver = build_solver(pkg, arr, config, m, tuning=tuning)
        plan, limit = plan_for(pkg, solver, config)
        run_steps(pkg, solver, plan, limit, 8)
        solver.ctx.read_compact()
        solver.dispose()
        log(f"[bench] {m}: 8 microsteps sampled")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--cells", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["both", "gather", "fixed"], default="both")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batched-steps", type=int, default=480)
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched leg without the scenes stepped one after another")
    ap.add_argument("--only-batched", action="store_true",
                    help="run the batched leg alone (a profiling workload); prints its JSON object")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout leg")
    ap.add_argument("--adjoint-reps", type=int, default=5)
    ap.add_argument("--no-small", action="store_true", help="skip the Tait tank and Gate 4 leg")
    ap.add_argument("--no-parity", action="store_true", help="skip the error-vs-FP64 checks of the CPU leg")
    ap.add_argument("--batch-tuning", default="auto",
                    help='launch shape of the batched leg: "auto" or "p2g_tile,g2p_tile,g2p_threads,bucket_capacity'
                         '[,tile_queue]"')
    ap.add_argument("--tuning", default="auto",
                    help='launch shape: "auto" (autotune before timing), "default", or '
                         '"p2g_tile,g2p_tile,g2p_threads,bucket_capacity[,tile_queue]"')
    ap.add_argument("--pmc-sample", action="store_true",
                    help="counter-collection workload: one 256 MiB copy kernel (the HBM byte calibration), then "
                         "8 untimed microsteps per mode; prints nothing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (each GPU box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    pkg = __graft_entry__.load_package()
    if args.only_batched:
        print(json.dumps(batched_leg(pkg, args.batched_steps, args.batch_tuning, not args.no_sequential)), flush=True)
        return
    log(f"[bench] building scene: {args.cells}x{args.cells} cells")
    arr, config = pkg.scenes.water_oil_column(cells=args.cells)
    if args.pmc_sample:
        pmc_sample(pkg, arr, config, args.mode, args.tuning)
        return
    n = len(arr)
    nodes = (args.cells + 3) ** 2

    results = {}
    kernel_times = {}
    tunings = {}
    microdt = None
    for mode in (["gather", "fixed"] if args.mode == "both" else [args.mode]):
        tuning, tunings[mode] = choose_tuning(pkg, arr, config, mode, args.tuning)
        solver = build_solver(pkg, arr, config, mode, tuning=tuning)
        tunings[mode]["effective"] = solver.ctx.get_tuning().as_dict()
        plan, limit = plan_for(pkg, solver, config)
        microdt = plan.microDt
        log(f"[bench] {mode}: plan {plan.stepCount} microsteps of {plan.microDt:.4e} s; timing {args.steps} steps")
        el, speed = timed(pkg, solver, plan, limit, args.steps, args.warmup)
        results[mode] = {"seconds": el, "steps": args.steps, "rate": n * args.steps / el,
                         "ms_per_step": 1e3 * el / args.steps, "max_speed": speed}
        log(f"[bench] {mode}: {1e3 * el / args.steps:.4f} ms/step, {n * args.steps / el:.4e} particle-steps/s")
        solver.dispose()
        # per-kernel times with HIP events on the solver's stream (separate run)
        solver = build_solver(pkg, arr, config, mode, timing=True, tuning=tuning)
        run_steps(pkg, solver, plan, limit, min(args.steps, 2 * plan.stepCount))
        solver.ctx.read_compact()
        kt = {}
        for k in ("p2g_grid", "p2g_grid_fixed", "g2p", "measure_speed", "fix_rollback", "commit_last",
                  "batch_begin"):
            mean, launches = solver.ctx.read_timing(k)
            if launches:
                kt[k] = {"mean_us": round(mean, 3), "launches": launches}
        kernel_times[mode] = kt
        log(f"[bench] {mode} kernel means (us): " + ", ".join(f"{k}={v['mean_us']}" for k, v in kt.items()))
        solver.dispose()

    main_mode = "gather" if "gather" in results else "fixed"
    r = results[main_mode]
    kt = kernel_times[main_mode]
    p2g_name = "p2g_grid" if main_mode == "gather" else "p2g_grid_fixed"
    phases = 2  # water and oil
    p2g_bytes = P2G_PARTICLE_BYTES * n + P2G_NODE_PHASE_BYTES * phases * nodes
    g2p_bytes = G2P_PARTICLE_BYTES * n + G2P_NODE_PHASE_BYTES * phases * nodes
    cand = {p2g_name: (p2g_bytes, kt.get(p2g_name, {}).get("mean_us")),
            "g2p": (g2p_bytes, kt.get("g2p", {}).get("mean_us"))}
    dom = max(cand, key=lambda k: cand[k][1] or 0.0)
    dom_bytes, dom_us = cand[dom]
    achieved = dom_bytes / (dom_us * 1e-6) / 1e9 if dom_us else None
    traffic = measured_traffic(tunings[main_mode].get("effective")).get(dom) if args.cells == 1000 else None
    step_bytes = PARTICLE_BYTES * n + NODE_PHASE_BYTES * phases * nodes
    step_gbs = step_bytes / (r["ms_per_step"] * 1e-3) / 1e9

    batched = (None if args.no_batched or args.cells != 1000
               else batched_leg(pkg, args.batched_steps, args.batch_tuning, not args.no_sequential))
    adjoint = None if args.no_adjoint else adjoint_leg(pkg, args.adjoint_reps)
    small = None if args.no_small else small_scenes_leg(pkg)

    cpu = None
    parity = None
    if not args.no_cpu:
        hosts = host_cpus()
        threads = baseline_threads(hosts)
        log(f"[bench] CPU baseline: FP64 oracle step, {threads} threads (host: {hosts})")
        rate, steps, el = cpu_baseline(pkg, arr, config, budget_s=8.0, threads=threads)
        log("[bench] CPU baseline: FP64 oracle step, one thread")
        rate1, steps1, el1 = cpu_baseline(pkg, arr, config)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{steps} FP64 oracle microsteps of the full {n}-particle scene on {threads} OpenMP "
                         f"threads ({el:.1f} s, after 1 warm-up step)",
               "host": hosts,
               "single_thread": {"value": rate1, "cores": 1,
                                 "sample": f"{steps1} FP64 oracle microsteps, one thread ({el1:.1f} s)"},
               "gpu_over_cpu": None}
        fp64_adjoint = None
        if adjoint is not None:
            log("[bench] CPU baseline: FP64 oracle adjoint of the elastic block, one thread")
            cpu["adjoint"], fp64_adjoint = adjoint_cpu(pkg)
        if not args.no_parity:
            parity = parity_leg(pkg, arr, config, fp64_adjoint)

    if cpu is not None:
        cpu["gpu_over_cpu"] = r["rate"] / cpu["value"]
    line = {
        "metric": METRIC,
        "value": r["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"water/oil column, {n} particles (60% water / 40% oil), {args.cells}x{args.cells} "
                               "cells of size 1, CFL microsteps of a 1/240 s macrostep, float32 gather P2G",
                   "particles": n, "grid_nodes": nodes, "microdt": microdt},
        "fixed_point": None if "fixed" not in results else {
            "value": results["fixed"]["rate"], "ms_per_step": results["fixed"]["ms_per_step"]},
        "step_hbm": {"algorithmic_bytes_per_step": step_bytes, "achieved_GBs": step_gbs,
                     "frac": step_gbs / HBM_PEAK_GBS},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None if achieved is None else achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom_bytes, "mean_launch_us": dom_us,
                     "traffic_source": str(TRAFFIC_PROFILE.relative_to(ROOT)) if traffic else None},
        "kernels_us": kernel_times,
        "tuning": tunings,
        "batched": batched,
        "adjoint": adjoint,
        "small_scenes": small,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
