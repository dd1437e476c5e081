"""Host logic of the interaction operations (pick, outlet removal, clone and
drag through a selection) with a fake of the native context, as the
reference's spec mocks its GPU (kernels/solver.spec.ts): validation before any
device work, selection resolution against the object ranges, the refusals
of clone, and the drag uniform (solver.ts:1389-1401, stiffness 20, damping
8).  CPU only."""
from __future__ import annotations

import numpy as np
import pytest

from test_adaptive_cfl import FakeContext

CONFIG = {"grid": {"cellsX": 16, "cellsY": 8, "cellSize": 1}, "particleMass": 0.1, "particleVolume": 0.01,
          "material": {"lambda": 20, "mu": 10}, "capacity": 8}


def particle(x, y, oid, model="neo-hookean"):
    mat = {"model": model, "lambda": 20.0, "mu": 10.0 if model != "water" else 0.0}
    if model == "water":
        mat["density"] = 1.0
    return {"position": {"x": x, "y": y}, "velocity": {"x": 0, "y": 0}, "C": {"m00": 0, "m01": 0, "m10": 0, "m11": 0},
            "F": {"m00": 1, "m01": 0, "m10": 0, "m11": 1}, "objectId": oid, "material": mat}


class InteractionContext(FakeContext):
    def __init__(self, n, pick_result=None, removed=0):
        super().__init__([], n=n, capacity=8)
        self.interaction = []
        self.pick_result = pick_result
        self.removed = removed
        self.cloned = []

    def set_interaction(self, values):
        self.interaction.append(list(values))

    def pick(self, x, y, r):
        return np.array(self.pick_result, np.uint32)

    def remove_in_outlet(self, bounds):
        self.bounds = list(bounds)
        return self.removed

    def clone(self, start, count, ox, oy, oid):
        self.cloned.append((start, count, ox, oy, oid))
        self.n += count


def solver(pkg, ctx):
    parts = [particle(2, 2, 3), particle(2.5, 2, 3), particle(9, 4, 7, "water")]
    state = pkg.layout.validate_input(parts, CONFIG)
    return pkg.MpmMlsGpuSolver._over_context(state, None, ctx, pkg.native.MPM_P2G_GATHER)


def f32bits(v):
    return int(np.float32(v).view(np.uint32))


def test_pick_resolves_the_gpu_result(pkg):
    ctx = InteractionContext(3, [1, f32bits(2.5), f32bits(2.0), 3])
    s = solver(pkg, ctx)
    sel = s.pick({"x": 2.4, "y": 2.1}, 0.5)
    assert sel == {"particleIndex": 1, "objectId": 3, "anchor": {"x": 2.5, "y": 2.0}}
    ctx.pick_result = [0xFFFFFFFF, 0, 0, 0]
    assert s.pick({"x": 2.4, "y": 2.1}, 0.5) is None
    ctx.pick_result = [2, f32bits(9.0), f32bits(4.0), 3]  # object 3 does not hold particle 2
    with pytest.raises(RuntimeError, match="inconsistent"):
        s.pick({"x": 9, "y": 4}, 0.5)


@pytest.mark.parametrize("point,radius,err", [
    ({"x": -1, "y": 2}, 1.0, "RangeError"), ({"x": 2, "y": 9}, 1.0, "RangeError"),
    ({"x": float("nan"), "y": 2}, 1.0, "TypeError"), ({"x": 2, "y": 2}, -0.5, "RangeError"),
    ({"x": 2, "y": 2}, 17.0, "RangeError"), ([2, 2], 1.0, "TypeError")])
def test_pick_validates_before_any_device_work(pkg, point, radius, err):
    ctx = InteractionContext(3, [0, 0, 0, 0])
    s = solver(pkg, ctx)
    exc = pkg.RangeError if err == "RangeError" else TypeError
    with pytest.raises(exc):
        s.pick(point, radius)


def test_drag_writes_the_reference_uniform(pkg):
    ctx = InteractionContext(3)
    s = solver(pkg, ctx)
    sel = {"particleIndex": 0, "objectId": 3, "anchor": {"x": 2.0, "y": 2.0}}
    s.begin_drag(sel, {"x": 5.0, "y": 3.0}, 2.0)
    s.update_drag({"x": 6.0, "y": 3.5})
    s.end_drag()
    assert ctx.interaction == [[5.0, 3.0, 2.0, 20.0, 8.0, 0.0, 0.0, 0.0], [6.0, 3.5, 2.0, 20.0, 8.0, 0.0, 0.0, 0.0],
                               [0.0, 0.0, 0.0, 20.0, 8.0, 0.0, 0.0, 0.0]]
    with pytest.raises(pkg.RangeError, match="drag radius must be greater than zero"):
        s.begin_drag(sel, {"x": 5.0, "y": 3.0}, 0.0)
    with pytest.raises(pkg.RangeError, match="does not match"):
        s.begin_drag({**sel, "objectId": 7}, {"x": 5.0, "y": 3.0}, 1.0)
    with pytest.raises(pkg.RangeError, match="active MPM particle"):
        s.begin_drag({**sel, "particleIndex": 3}, {"x": 5.0, "y": 3.0}, 1.0)


def test_outlet_counts_and_validation(pkg):
    ctx = InteractionContext(3, removed=1)
    s = solver(pkg, ctx)
    with pytest.raises(pkg.RangeError, match="must not exceed the maximum"):
        s.remove_particles_in_outlet({"minX": 5, "minY": 1, "maxX": 4, "maxY": 6})
    assert s.remove_particles_in_outlet({"minX": 8, "minY": 3, "maxX": 10, "maxY": 5}) == 1
    assert ctx.bounds == [8.0, 3.0, 10.0, 5.0]
    assert s.particle_count == 2 and s.allocated_particle_count == 3


def test_clone_bookkeeping_and_refusals(pkg):
    ctx = InteractionContext(3)
    s = solver(pkg, ctx)
    sel = {"particleIndex": 1, "objectId": 3, "anchor": {"x": 2.5, "y": 2.0}}
    assert s.clone(sel, {"x": 4.0, "y": 1.0}) is True
    assert ctx.cloned == [(0, 2, 4.0, 1.0, 8)]        # next object id after 3 and 7
    assert s.particle_count == 5 and s._object_ranges[8] == {"objectId": 8, "start": 3, "count": 2}
    copy = {"particleIndex": 4, "objectId": 8, "anchor": {"x": 6.5, "y": 3.0}}
    assert s.clone(copy, {"x": 1.0, "y": 0.0}) is True  # a clone can be cloned
    assert ctx.cloned[-1] == (3, 2, 1.0, 0.0, 9)
    assert s.clone(sel, {"x": 1.0, "y": 0.0}) is False  # 7 slots in use of 8: no room for 2
    with pytest.raises(TypeError):
        s.clone(sel, {"x": float("inf"), "y": 0.0})
    ctx2 = InteractionContext(3, removed=1)
    t = solver(pkg, ctx2)
    t.remove_particles_in_outlet({"minX": 8, "minY": 3, "maxX": 10, "maxY": 5})
    assert t.clone(sel, {"x": 1.0, "y": 0.0}) is False and ctx2.cloned == []
