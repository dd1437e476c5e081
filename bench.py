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
differentiable-rollout legs, and the CPU baseline (the FP64 oracle step,
single thread, on a bounded sample of the same scene; and the FP64 oracle
adjoint of the elastic block).
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


def baseline_threads() -> int:
    """Host threads of the all-core CPU baseline: OMP_NUM_THREADS when set
    (the GPU host's CPU share), else the CPU count, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, os.cpu_count() or 1))


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
    Gate 4 scene (3,674 particles with bonded pebbles and a fracture solid,
    1/240 s macrosteps).  Each is stepped through the adaptive CFL loop with
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
    out["reference_published"] = {
        "tait_tank": "23.89e6 particle-steps/s, 1,600 particles, 423-microstep batch, Apple Metal 3 "
                     "(docs/PERFORMANCE-MLS-MPM.md:837-851); ~4.8e7 with fixed-point P2G "
                     "(docs/REPORT-FLUID-SOLVER-PROTOTYPES.md:207-229)",
        "gate4": "0.216618 ms per step (median, GPU timestamps) on Apple Metal 3, same protocol "
                 "(docs/PERFORMANCE-MLS-MPM.md:622-636); CPU reference 4.643 ms per step, one JS thread "
                 "(docs/PERFORMANCE-MLS-MPM.md:778-790)"}
    return out


def batched_leg(pkg, steps: int, how: str = "auto"):
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
        keys = ("p2g_tile", "g2p_tile", "g2p_threads", "bucket_capacity")
        tuned = {"tuning": dict(zip(keys, (int(v) for v in how.split(",")))), "candidates": None}
    batch = pkg.MpmMlsGpuBatch(scenes, device=0, tuning=tuned["tuning"])
    batch.step_batch(8, dt, gravity)
    batch.scene(0).ctx.read_compact()
    t0 = time.perf_counter()
    batch.step_batch(steps, dt, gravity)
    batch.scene(0).ctx.read_compact()  # waits for the batch's stream
    t_batch = time.perf_counter() - t0
    bad = [i for i in range(len(batch)) if batch.scene(i).ctx.read_compact()[2][0] != 0]
    batch.dispose()
    if bad:
        raise RuntimeError(f"batched scenes {bad[:8]} raised numerical flags")
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
    rate, rate_seq = total * steps / t_batch, total * steps / t_seq
    log(f"[bench] batched: {1e3 * t_batch / steps:.4f} ms per microstep of all scenes ({rate:.3e} particle-steps/s); "
        f"sequential {1e3 * t_seq / steps:.4f} ms ({rate_seq:.3e}); x{rate / rate_seq:.1f}")
    return {"value": rate, "unit": "particle-steps/s", "ms_per_step": 1e3 * t_batch / steps, "scenes": len(tanks),
            "particles": total, "steps": steps, "sequential_value": rate_seq,
            "sequential_ms_per_step": 1e3 * t_seq / steps, "speedup_vs_sequential": rate / rate_seq,
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
    """FP64 oracle forward + reverse of the same elastic block, one thread."""
    from oracle import oracle as orc
    state, lam, mu, config, target = pkg.scenes.elastic_block()
    cfg = orc.diff_config(config)
    t0 = time.perf_counter()
    final = orc.diff_forward(state, lam, mu, cfg)
    t1 = time.perf_counter()
    _, bar = orc.com_loss(final, target)
    orc.diff_backward(state, lam, mu, cfg, bar)
    t2 = time.perf_counter()
    n, T = len(lam), config["steps"]
    return {"forward_value": n * T / (t1 - t0), "backward_value": n * T / (t2 - t1), "unit": "particle-steps/s",
            "cores": 1, "kind": "port", "sample": f"one FP64 forward and reverse of the {n}-particle block "
                                                   f"({t1 - t0:.2f} s + {t2 - t1:.2f} s; the reverse re-runs the forward)"}


def choose_tuning(pkg, arr, config, mode, how):
    """The launch shape a mode is measured with: "default", "auto" (the
    autotuner on a scratch copy of the scene, before and outside every timed
    region) or explicit "p2g_tile,g2p_tile,g2p_threads,bucket_capacity".
    Returns (tuning dict or None, report)."""
    if how == "default":
        return None, {"source": "default", "tuning": dict(pkg.autotune.DEFAULT_TUNING)}
    if how != "auto":
        keys = ("p2g_tile", "g2p_tile", "g2p_threads", "bucket_capacity")
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
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout leg")
    ap.add_argument("--adjoint-reps", type=int, default=5)
    ap.add_argument("--no-small", action="store_true", help="skip the Tait tank and Gate 4 leg")
    ap.add_argument("--batch-tuning", default="auto",
                    help='launch shape of the batched leg: "auto" or "p2g_tile,g2p_tile,g2p_threads,bucket_capacity"')
    ap.add_argument("--tuning", default="auto",
                    help='launch shape: "auto" (autotune before timing), "default", or '
                         '"p2g_tile,g2p_tile,g2p_threads,bucket_capacity"')
    ap.add_argument("--pmc-sample", action="store_true",
                    help="counter-collection workload: one 256 MiB copy kernel (the HBM byte calibration), then "
                         "8 untimed microsteps per mode; prints nothing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (each GPU box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    pkg = __graft_entry__.load_package()
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
               else batched_leg(pkg, args.batched_steps, args.batch_tuning))
    adjoint = None if args.no_adjoint else adjoint_leg(pkg, args.adjoint_reps)
    small = None if args.no_small else small_scenes_leg(pkg)

    cpu = None
    if not args.no_cpu:
        threads = baseline_threads()
        log(f"[bench] CPU baseline: FP64 oracle step, {threads} threads")
        rate, steps, el = cpu_baseline(pkg, arr, config, budget_s=8.0, threads=threads)
        log("[bench] CPU baseline: FP64 oracle step, one thread")
        rate1, steps1, el1 = cpu_baseline(pkg, arr, config)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{steps} FP64 oracle microsteps of the full {n}-particle scene on {threads} OpenMP "
                         f"threads ({el:.1f} s, after 1 warm-up step)",
               "single_thread": {"value": rate1, "cores": 1,
                                 "sample": f"{steps1} FP64 oracle microsteps, one thread ({el1:.1f} s)"},
               "gpu_over_cpu": None}
        if adjoint is not None:
            log("[bench] CPU baseline: FP64 oracle adjoint of the elastic block, one thread")
            cpu["adjoint"] = adjoint_cpu(pkg)

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
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
