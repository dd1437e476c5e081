"""The float32 restatements that tests/test_gpu_paths.py checks the GPU's
deformation retry against, pinned on the reference's own known answer for a
near-singular update (kernels/solver.spec.ts:197-260): with the fused
products the kernels use, the two-substep update keeps a positive
determinant where unfused products lose it.  CPU only."""
from __future__ import annotations

import numpy as np

from test_gpu_paths import F32, det64, fma32, integrate32, project_liquid32

INITIAL = [F32(0.06047571077942848), F32(0.17931117117404938), F32(0.17301689088344574), F32(0.5129971504211426)]
GRADIENT = [F32(0.1195390224456787), F32(0.15536771714687347), F32(0.040612027049064636), F32(0.3162637948989868)]


def test_near_singular_update_known_answer():
    row = max(F32(abs(GRADIENT[0]) + abs(GRADIENT[1])), F32(abs(GRADIENT[2]) + abs(GRADIENT[3])))
    assert float(row) == 0.35687583684921265
    substeps = max(1, int(np.ceil(row / F32(0.25))))
    assert substeps == 2
    assert 0 < det64(INITIAL) < 1e-8
    fused = integrate32(INITIAL, GRADIENT, F32(1.0), substeps)
    assert det64(fused) > 0
    # unfused: each entry rounds the first product before the sum
    h = F32(1.0) / F32(substeps)
    u = [F32(1) + F32(h * GRADIENT[0]), F32(h * GRADIENT[1]), F32(h * GRADIENT[2]), F32(1) + F32(h * GRADIENT[3])]
    r = list(INITIAL)
    for _ in range(substeps):
        r = [F32(F32(u[0] * r[0]) + F32(u[1] * r[2])), F32(F32(u[0] * r[1]) + F32(u[1] * r[3])),
             F32(F32(u[2] * r[0]) + F32(u[3] * r[2])), F32(F32(u[2] * r[1]) + F32(u[3] * r[3]))]
    assert det64(r) <= 0


def test_fma32_is_the_correctly_rounded_fused_product():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b, c = (F32(v) for v in rng.standard_normal(3))
        exact = float(a) * float(b) + float(c)   # a*b is exact in FP64; the sum rounds once at most
        got = fma32(a, b, c)
        assert abs(float(got) - exact) <= 0.5 * float(np.spacing(np.float32(abs(got)))) * (1 + 1e-9)


def test_liquid_projection_keeps_the_rotation():
    th = 0.3
    R = [F32(np.cos(th)), F32(-np.sin(th)), F32(np.sin(th)), F32(np.cos(th))]
    out = project_liquid32(R)
    np.testing.assert_allclose([float(v) for v in out], [float(v) for v in R], rtol=0, atol=2e-7)
