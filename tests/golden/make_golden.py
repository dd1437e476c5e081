"""Write tests/golden/golden.npz: input/output vectors of the CPU oracle.

Run from the repository root after building the oracle:
    python tests/golden/make_golden.py

The reference commits no golden vectors for this path (its parity checks
compare live CPU and GPU runs), so these fixtures come from this repository's
FP64 and integer oracle, whose functions are pinned independently by the
closed forms and invariants in tests/test_oracle.py.  They freeze the oracle
against later edits and give the GPU tests committed targets:

* five_*: the five-material transfer fixture (experiments/transfer-regression.ts
  :33-95), 8 steps of dt 1e-3 without gravity, for 1, 9 and 45 particles;
* fixed3_*: the int32 2^24 fixed-point P2G words of three water particles
  (shaders.ts:971-1070), small enough to check by hand;
* tank_*: one 1/120 s macrostep (224 microsteps) of the SI Tait tank.
"""
from __future__ import annotations

import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import __graft_entry__  # noqa: E402
import oracle as orc  # noqa: E402

STATE_COLS = ("x", "y", "vx", "vy", "c00", "c01", "c10", "c11", "f00", "f01", "f10", "f11")


def state(parts):
    return np.stack([parts[k] for k in STATE_COLS], axis=1)


def main():
    pkg = __graft_entry__.load_package()
    out = {}
    particles, config = pkg.scenes.five_model_block()
    packed = pkg.layout.validate_input(particles, config).packed
    cfg = orc.make_config(1e-3, 48, 16, 1.0, 0.1, 0.05)
    for name, sel in (("five_1", [22]), ("five_9", list(range(18, 27))), ("five_45", list(range(45)))):
        inp = np.ascontiguousarray(packed[sel])
        out[name + "_in"] = inp
        out[name + "_out"] = state(orc.run_f64(orc.particles_from_packed(inp), cfg, 8))

    recs = []
    for x, y, vx in ((5.5, 6.25, 0.0), (6.1, 6.9, 0.5), (7.75, 5.6, -0.25)):
        r = np.zeros(20, np.float32)
        r[0], r[1], r[2] = x, y, vx
        r[8], r[11], r[12] = 1.0, 1.0, 1.0
        r[14], r[16], r[17] = 8.0, 0.02, 1.0  # lambda, density, volume (water, linear EOS)
        u = r.view(np.uint32)
        u[13] = 2
        recs.append(u)
    fixed_in = np.stack(recs)
    prm = orc.params_f32(16, 16, 1.0, 0.02, 1.0, 1e-3, 3)
    words, flags = orc.fixed_p2g_f32(fixed_in, prm)
    assert flags == 0
    out["fixed3_in"] = fixed_in
    out["fixed3_grid"] = np.array([16, 16], np.int32)
    out["fixed3_words"] = words

    tparticles, tconfig, macro_dt, _ = pkg.scenes.tait_tank()
    tpacked = pkg.layout.validate_input(tparticles, tconfig).packed
    steps = 224
    dt = float(np.float32(macro_dt / steps))
    tcfg = orc.make_config(dt, 80, 80, float(np.float32(0.0025)), float(np.float32(tconfig["particleMass"])),
                           float(np.float32(tconfig["particleVolume"])), (0.0, float(np.float32(-9.81))), 1.0)
    out["tank_in"] = tpacked
    # dt, microsteps, cell size, particleMass, particleVolume, gravity y (as the oracle takes them)
    out["tank_cfg"] = np.array([dt, steps, float(np.float32(0.0025)), float(np.float32(tconfig["particleMass"])),
                                float(np.float32(tconfig["particleVolume"])), float(np.float32(-9.81))], np.float64)
    out["tank_out"] = state(orc.run_f64(orc.particles_from_packed(tpacked), tcfg, steps))

    np.savez_compressed(ROOT / "tests" / "golden" / "golden.npz", **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
