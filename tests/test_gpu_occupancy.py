"""GPU: the surface occupancy gather (gatherSurfaceOccupancy,
kernels/shaders.ts:710-806) against a CPU restatement of the reference's
current-area occupancy (reference/occupancy.ts:70-140) in FP64, scaled to the
display area (2 V det F, shaders.ts:760-763).

The reference sums each node's contributions in linked-list order, which is
not deterministic; the GPU here sums them in the P2G gather's fixed order.
Both are float32 sums of the same terms, so the bar is a float32 one: every
node within 1e-5 of the largest occupancy plus 1e-5 relative."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def occupancy_restated(recs: np.ndarray, cells_x: int, cells_y: int, dx: float) -> np.ndarray:
    """gatherCurrentArea (reference/occupancy.ts:70-140) of the liquid
    particles, in FP64 from the float32 grid position, times the display
    factor 2 of the GPU field."""
    W, H = cells_x + 3, cells_y + 3
    f = recs.view(np.float32).astype(np.float64)
    out = np.zeros((W * H, 2))
    for phase, model in ((0, 2), (1, 3)):
        m = recs[:, 13] == model
        if not m.any():
            continue
        # the grid position in float32, as the kernel forms it (shaders.ts:754);
        # the rest in FP64
        gx = (recs[m, 0].view(np.float32) / np.float32(dx)).astype(np.float64)
        gy = (recs[m, 1].view(np.float32) / np.float32(dx)).astype(np.float64)
        bx, by = np.floor(gx - 0.5), np.floor(gy - 0.5)
        fx, fy = gx - bx, gy - by
        wx = [0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1) ** 2, 0.5 * (fx - 0.5) ** 2]
        wy = [0.5 * (1.5 - fy) ** 2, 0.75 - (fy - 1) ** 2, 0.5 * (fy - 0.5) ** 2]
        area = f[m, 17] * (f[m, 8] * f[m, 11] - f[m, 9] * f[m, 10])
        for oy in range(3):
            for ox in range(3):
                key = ((by + 1 + oy) * W + (bx + 1 + ox)).astype(np.int64)
                np.add.at(out[:, phase], key, 2.0 * wx[ox] * wy[oy] * area / (dx * dx))
    return out


def check(got, want, label):
    scale = float(np.abs(want).max())
    err = float(np.abs(got.astype(np.float64) - want).max())
    print(f"{label}: max occupancy {scale:.4g}, max |GPU - FP64| {err:.3e}", flush=True)
    assert np.all(np.abs(got - want) <= 1e-5 * scale + 1e-5 * np.abs(want))
    assert np.count_nonzero(want[:, 0]) > 0


@pytest.mark.parametrize("mode", ["gather", "fixed"])
def test_occupancy_of_the_water_oil_column(pkg, mode):
    arr, config = pkg.scenes.water_oil_column(cells=300)
    factory = pkg.create_mpm_mls_gpu_solver if mode == "gather" else pkg.create_experimental_mpm_mls_atomic_gpu_solver
    s = factory(arr, config, device=0)
    try:
        for steps in (0, 20):
            if steps:
                s.step_batch(steps, 1.0 / 9600.0, pkg.scenes.WATER_OIL_GRAVITY)
            occ = s.gather_surface_occupancy()
            recs, _ = s.read_particles(packed=True)
            check(occ, occupancy_restated(recs, 300, 300, 1.0), f"column {mode} after {steps} microsteps")
            assert np.array_equal(occ, s.gather_surface_occupancy())  # reproducible
    finally:
        s.dispose()


def test_occupancy_of_the_tait_tank_and_gate4(pkg):
    particles, config, macro, _ = pkg.scenes.tait_tank()
    s = pkg.create_mpm_mls_gpu_solver(particles, config, device=0)
    try:
        s.step_batch(50, macro / 224, pkg.scenes.TAIT_TANK_GRAVITY)
        recs, _ = s.read_particles(packed=True)
        check(s.gather_surface_occupancy(), occupancy_restated(recs, 80, 80, float(np.float32(0.0025))), "Tait tank")
    finally:
        s.dispose()
    parts, cfg, gravity = pkg.scenes.gate4_scene()
    s = pkg.create_mpm_mls_gpu_solver(parts, cfg, device=0)
    try:
        s.step_batch(3, 1.0 / 2400.0, gravity)
        recs, _ = s.read_particles(packed=True)
        occ = s.gather_surface_occupancy()
        check(occ, occupancy_restated(recs, cfg["grid"]["cellsX"], cfg["grid"]["cellsY"],
                                      float(cfg["grid"]["cellSize"])), "Gate 4")
        assert np.count_nonzero(occ[:, 1]) > 0
    finally:
        s.dispose()


def test_occupancy_refused_while_the_status_is_set(pkg):
    arr, config = pkg.scenes.water_oil_column(cells=32)
    s = pkg.create_mpm_mls_gpu_solver(arr, config, device=0)
    try:
        s.step_batch(2, 0.001, {"x": 0.0, "y": 0.0}, cfl_velocity_bound=1e-6)  # the CFL guard fails
        with pytest.raises(pkg.native.NativeError, match="numerical status"):
            s.gather_surface_occupancy()
        s.clear_numerical_status()
        assert s.gather_surface_occupancy().shape == (35 * 35, 2)
    finally:
        s.dispose()
