"""GPU: the allocation-free surface occupancy gather
(mpm_gather_occupancy_static, whose per-node scratch is a static device
array of the library) returns bit for bit what mpm_gather_occupancy returns,
called through the C-ABI on two scenes in a row, the second on a smaller
grid, and then on the first again; grids beyond the static scratch are
refused with an error."""
from __future__ import annotations

import numpy as np
import pytest

from test_gpu_interaction import blocks_scene

pytestmark = pytest.mark.gpu


def test_static_occupancy_matches_the_allocating_gather(pkg):
    parts, config, gravity = pkg.scenes.gate4_scene()            # 220 x 600 cells
    big = pkg.create_mpm_mls_gpu_solver(parts, config, device=0)
    bparts, bconfig = blocks_scene(pkg)                          # 64 x 64 cells
    small = pkg.create_mpm_mls_gpu_solver(bparts, bconfig, device=0)
    try:
        big.step_batch(6, 1.0 / 2400.0, gravity)
        small.step_batch(5, 1e-3, {"x": 0.0, "y": -9.8})
        for label, s in (("Gate 4, 220 x 600", big), ("blocks, 64 x 64", small), ("Gate 4 again", big)):
            print(f"static occupancy: {label}", flush=True)
            nodes = s.ctx.counts()[2]
            got = s.ctx.gather_occupancy_static()
            want = s.ctx.gather_occupancy()
            again = s.ctx.gather_occupancy_static()
            assert got.shape == want.shape == (nodes, 2)
            assert np.count_nonzero(want[:, 0]) > 0
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), label
            assert np.array_equal(again.view(np.uint32), want.view(np.uint32)), label
    finally:
        small.dispose()
        big.dispose()


def test_static_occupancy_refuses_grids_beyond_the_scratch(pkg):
    bparts, bconfig = blocks_scene(pkg)
    bconfig = dict(bconfig, grid={"cellsX": 1030, "cellsY": 1020, "cellSize": 1.0})   # 1033 x 1023 nodes
    s = pkg.create_mpm_mls_gpu_solver(bparts, bconfig, device=0)
    try:
        assert s.ctx.counts()[2] == 1033 * 1023 > 1048576
        with pytest.raises(Exception, match="static scratch holds at most 1048576"):
            s.ctx.gather_occupancy_static()
    finally:
        s.dispose()
