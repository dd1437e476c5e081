"""Closed-form pins of the FP64 oracle's constitutive stresses, its G2P
density correction and its drag, from the mathematics of the method rather
than from the oracle's own code.  CPU only.

The lone-particle identity.  A single particle away from the walls, with no
gravity, deposits m_i = w_i m and q_i = w_i (m v + A d_i) on the nodes of its
stencil (d_i = x_i - x_p), so every node moves at v + A d_i / m.  Gathered
back, the quadratic B-spline's zero first moment (sum w_i d_i = 0) and its
second moment (sum w_i d_i d_i^T = dx^2/4 I) give

    v' = v      and      C' = (4/dx^2) sum w_i v_i d_i^T = A / m,

with A = m C + (-dt V 4/dx^2) tau (reference/solver.ts:418-479, :610-773),
tau = P F^T the Kirchhoff stress of the model.  Each test below computes tau
from the textbook formula of its model, by numpy's SVD where a rotation is
needed, at dx != 1, and asserts the oracle's C' against it.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

DX = 0.37          # a cell size away from 1, so 4/dx^2, 4/dx and 1 differ
CELLS = 64         # the particle sits ~30 cells from every wall
DT = 1e-3


def particle(orc, F, model, lam=20.0, mu=10.0, density=0.0, volume=0.05, param=0.0, pv=1.0,
             C=(0.3, -0.2, 0.15, 0.1), v=(0.4, -0.25), x=11.83, y=12.07):
    p = np.zeros(1, orc.PARTICLE_DTYPE)
    p["x"], p["y"], p["vx"], p["vy"] = x, y, v[0], v[1]
    for k, name in enumerate(("c00", "c01", "c10", "c11")):
        p[name] = C[k]
    for k, name in enumerate(("f00", "f01", "f10", "f11")):
        p[name] = np.asarray(F, np.float64).reshape(4)[k]
    p["plastic_volume"], p["lambda_"], p["mu"] = pv, lam, mu
    p["density"], p["volume"], p["param"], p["model"] = density, volume, param, model
    return p


def config(orc, mass=0.3, g=(0.0, 0.0), drag=None, dx=DX, cells=CELLS):
    return orc.make_config(DT, cells, cells, dx, mass, 0.05, g, 0.0, drag)


def affine_after_one_step(orc, p, cfg):
    out, _ = orc.step_f64(p, cfg)
    C = np.array([[out["c00"][0], out["c01"][0]], [out["c10"][0], out["c11"][0]]])
    return C, np.array([out["vx"][0], out["vy"][0]])


def expected_affine(p, tau, m, dx=DX):
    C = np.array([[p["c00"][0], p["c01"][0]], [p["c10"][0], p["c11"][0]]])
    return C + (-DT * float(p["volume"][0]) * 4.0 / (dx * dx) / m) * tau


def check_identity(orc, p, cfg, tau, m, dx=DX):
    C1, v1 = affine_after_one_step(orc, p, cfg)
    want = expected_affine(p, tau, m, dx)
    stress_part = want - np.array([[p["c00"][0], p["c01"][0]], [p["c10"][0], p["c11"][0]]])
    # the stress term must be far above the tolerance for the check to mean anything
    assert np.abs(stress_part).max() > 1e-5
    np.testing.assert_allclose(C1, want, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(v1, [p["vx"][0], p["vy"][0]], rtol=0, atol=1e-12)


def rotation(theta):
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[c, -s], [s, c]])


def polar_rotation(F):
    U, _, Vt = np.linalg.svd(F)
    R = U @ Vt
    if np.linalg.det(R) < 0:  # F here has det > 0, so this never triggers
        U[:, 1] *= -1
        R = U @ Vt
    return R


DEFORMATIONS = [
    np.array([[1.08, 0.0], [0.0, 1.0]]),                                  # uniaxial stretch
    np.array([[0.93, 0.04], [-0.02, 0.97]]),                              # compression with shear
    rotation(0.4) @ np.array([[1.05, 0.0], [0.0, 0.9]]) @ rotation(-0.9),  # rotated, anisotropic
]


@pytest.mark.parametrize("F", DEFORMATIONS)
@pytest.mark.parametrize("damage", [0.0, 0.25])
def test_neo_hookean_kirchhoff_stress(orc, F, damage):
    """tau = (1 - d) (mu (F F^T - I) + lambda ln J I): the compressible
    Neo-Hookean Kirchhoff stress, scaled by the fracture damage
    (reference/solver.ts:2036-2061, :2204-2212)."""
    lam, mu, m = 20.0, 10.0, 0.3
    J = np.linalg.det(F)
    tau = (1 - damage) * (mu * (F @ F.T - np.eye(2)) + lam * math.log(J) * np.eye(2))
    check_identity(orc, particle(orc, F, 0, lam, mu, param=damage), config(orc, mass=m), tau, m)


def test_neo_hookean_uniaxial_closed_form(orc):
    """F = diag(s, 1): tau = diag(mu (s^2 - 1) + lambda ln s, lambda ln s)."""
    s, lam, mu, m = 1.12, 20.0, 10.0, 0.3
    tau = np.diag([mu * (s * s - 1) + lam * math.log(s), lam * math.log(s)])
    check_identity(orc, particle(orc, np.diag([s, 1.0]), 0, lam, mu), config(orc, mass=m), tau, m)


@pytest.mark.parametrize("F", DEFORMATIONS)
@pytest.mark.parametrize("pv", [0.9, 1.0, 1.3])
def test_snow_corotated_stress_with_hardening(orc, F, pv):
    """tau = 2 mu h (F - R) F^T + lambda h J (J - 1) I with the hardening
    h = exp(10 (1 - J_p)) (reference/solver.ts:2171-2203, materials.ts:46-53):
    a plastic volume below one hardens, above one softens."""
    lam, mu, m = 20.0, 10.0, 0.3
    h = math.exp(10.0 * (1.0 - pv))
    J = np.linalg.det(F)
    R = polar_rotation(F)
    tau = 2 * mu * h * (F - R) @ F.T + lam * h * J * (J - 1) * np.eye(2)
    check_identity(orc, particle(orc, F, 1, lam, mu, pv=pv), config(orc, mass=m), tau, m)


@pytest.mark.parametrize("J", [0.97, 0.995, 1.02])
@pytest.mark.parametrize("model", [2, 3])
def test_tait_pressure(orc, J, model):
    """Tait: p = B ((rho/rho0)^n - 1) = B (J^-n - 1) (tait-eos.ts:26-86), and
    the Kirchhoff stress is -J p I (reference/solver.ts:2150-2160)."""
    B, n, density, volume = 40.0, 7.0, 2.0, 0.05
    F = np.diag([J, 1.0])
    m = density * volume
    tau = -J * B * (J ** -n - 1.0) * np.eye(2)
    check_identity(orc, particle(orc, F, model, lam=B, mu=0.0, density=density, volume=volume, param=n),
                   config(orc), tau, m)


@pytest.mark.parametrize("J", [0.96, 1.03])
def test_linear_liquid_is_tension_free(orc, J):
    """param = 0: tau = lambda min(J - 1, 0) I (reference/solver.ts:2150-2160)."""
    lam, density, volume = 40.0, 2.0, 0.05
    F = rotation(0.3) @ np.diag([J, 1.0])
    m = density * volume
    tau = lam * min(J - 1.0, 0.0) * np.eye(2)
    p = particle(orc, F, 2, lam=lam, mu=0.0, density=density, volume=volume, param=0.0)
    if J > 1:
        C1, _ = affine_after_one_step(orc, p, config(orc))
        np.testing.assert_allclose(C1, expected_affine(p, tau, m), rtol=1e-12, atol=1e-13)
    else:
        check_identity(orc, p, config(orc), tau, m)


def powder_case(F, lam, mu, friction):
    """Hencky stress in the principal frame of F = U S V^T, with the
    deviatoric part capped at friction * pressure (the granular model of
    reference/solver.ts:2222-2265): returns (tau, pressure, deviatoric norm)."""
    U, S, _ = np.linalg.svd(F)
    l1, l2 = np.log(S)
    lv = l1 + l2
    pressure = max(0.0, -(lam + mu) * max(math.log(0.5), lv))
    dev = 0.5 * (l1 - l2)
    shear = min(2 * mu * dev, friction * pressure)
    mean = (lam + mu) * lv
    tau = U @ np.diag([mean + shear, mean - shear]) @ U.T
    return tau, pressure, shear


@pytest.mark.parametrize("F,friction,yields", [
    (np.diag([0.9, 0.88]), 0.5, False),
    (rotation(0.7) @ np.diag([0.93, 0.9]) @ rotation(-0.2), 0.5, False),
    (np.diag([1.1, 0.8]), 0.5, True),
    (rotation(-0.5) @ np.diag([1.06, 0.78]) @ rotation(1.1), 0.5, True),
    (np.diag([0.9, 0.5]), 0.2, True),   # volume below the 0.5 pressure floor
    (np.diag([1.05, 1.02]), 0.5, True),  # tension: no pressure, no shear
])
def test_powder_hencky_stress_and_friction_cap(orc, F, friction, yields):
    lam, mu, m = 20.0, 10.0, 0.3
    tau, pressure, shear = powder_case(F, lam, mu, friction)
    U, S, _ = np.linalg.svd(F)
    dev = 0.5 * math.log(S[0] / S[1])
    assert (2 * mu * dev > friction * pressure) == yields
    if not yields:  # unyielded: the plain Hencky stress 2 mu ln S + lambda tr(ln S) I
        hencky = U @ np.diag(2 * mu * np.log(S) + lam * np.log(S).sum()) @ U.T
        np.testing.assert_allclose(tau, hencky, rtol=1e-12, atol=1e-12)
    check_identity(orc, particle(orc, F, 4, lam, mu, param=friction), config(orc, mass=m), tau, m)


def test_identity_holds_at_another_cell_size(orc):
    """The same Neo-Hookean case at dx = 2.5: the stress term scales as
    1/dx^2 (a 4/dx scale would give a different C' at any dx != 1)."""
    F = DEFORMATIONS[2]
    lam, mu, m, dx = 20.0, 10.0, 0.3, 2.5
    tau = mu * (F @ F.T - np.eye(2)) + lam * math.log(np.linalg.det(F)) * np.eye(2)
    p = particle(orc, F, 0, lam, mu, volume=2.0, x=80.3, y=79.6)
    check_identity(orc, p, config(orc, mass=m, dx=dx), tau, m, dx=dx)


# ------------------------------------------------------------ buoyancy

@pytest.mark.parametrize("rho_a,rho_b", [(3.0, 1.0), (1.0, 1.6)])
def test_local_density_correction_is_buoyancy(orc, rho_a, rho_b):
    """Two water particles at one point, same volume, densities rho_a and
    rho_b, at rest with F = I under gravity g.  Every node they touch holds
    mass w (rho_a + rho_b) V and volume 2 w V, so the local density is the
    mean (rho_a + rho_b)/2 and each node moves at dt g.  The density
    correction v += dt g (rho_p - rho_local)/rho_p
    (reference/solver.ts:687-689) then gives each particle
    v' = dt g (1 + (rho_p - rho_local)/rho_p): the heavier one falls faster,
    the lighter one slower than free fall."""
    g = (0.0, -9.81)
    V = 0.05
    ps = np.concatenate([particle(orc, np.eye(2), 2, lam=40.0, mu=0.0, density=rho_a, volume=V, C=(0, 0, 0, 0),
                                  v=(0, 0)),
                         particle(orc, np.eye(2), 2, lam=40.0, mu=0.0, density=rho_b, volume=V, C=(0, 0, 0, 0),
                                  v=(0, 0))])
    out, _ = orc.step_f64(ps, config(orc, g=g))
    local = 0.5 * (rho_a + rho_b)
    for k, rho in enumerate((rho_a, rho_b)):
        want = DT * g[1] * (1.0 + (rho - local) / rho)
        assert out["vy"][k] == pytest.approx(want, rel=1e-12, abs=1e-15)
        assert out["vx"][k] == 0.0
    heavier = 0 if rho_a > rho_b else 1
    assert out["vy"][heavier] < DT * g[1] < out["vy"][1 - heavier]


# ------------------------------------------------------------ drag

def drag_node_velocity(v, node_xy, target, radius, k, c, dt=DT):
    """The spring-damper of finishGridVelocity (kernels/shaders.ts:1748-1777)
    at one node: within the radius, v += dt (1 - |d|/r) (k d - c v) with
    d = target - node."""
    d = np.asarray(target, np.float64) - np.asarray(node_xy, np.float64)
    dist = math.hypot(*d)
    if dist >= radius:
        return np.asarray(v, np.float64)
    return np.asarray(v, np.float64) + dt * (1.0 - dist / radius) * (k * d - c * np.asarray(v, np.float64))


@pytest.mark.parametrize("v", [(0.0, 0.0), (0.6, -0.3)])
def test_drag_pulls_nodes_towards_the_target(orc, v):
    """A lone Neo-Hookean particle at rest in F (so tau = 0) moving at v with
    C = 0: every node of its stencil holds velocity v before the drag.  With
    the target inside the stencil and a radius that reaches some nodes but not
    others, each node's velocity must equal the spring-damper closed form,
    nodes outside the radius keep v, and the gathered particle velocity is the
    weighted mean of the node velocities."""
    target, radius, k, c = (11.95, 12.2), 0.5, 20.0, 8.0
    p = particle(orc, np.eye(2), 0, C=(0, 0, 0, 0), v=v)
    cfg = config(orc, drag=(target[0], target[1], radius, k, c))
    grid = orc.grid_f64(p, cfg)
    F = orc.GRID_FIELDS
    W = CELLS + 3
    touched = np.nonzero(grid[:, F.index("m")] > 0)[0]
    assert len(touched) == 9
    inside = 0
    for key in touched:
        sx, sy = key % W, key // W
        node = ((sx - 1) * DX, (sy - 1) * DX)
        want = drag_node_velocity(v, node, target, radius, k, c)
        got = grid[key, [F.index("vx"), F.index("vy")]]
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
        inside += math.hypot(target[0] - node[0], target[1] - node[1]) < radius
    assert 0 < inside < 9
    # and with the drag off every node keeps v
    plain = orc.grid_f64(p, config(orc))
    np.testing.assert_allclose(plain[touched][:, [F.index("vx"), F.index("vy")]],
                               np.tile(v, (9, 1)), rtol=0, atol=1e-14)


def test_drag_acts_on_every_phase_and_skips_halo_nodes(orc):
    """The drag applies to the solid, water and oil velocities alike, and only
    at nodes inside the domain (0 <= node <= cells): a particle next to the
    corner with the target on it leaves the halo nodes undragged."""
    target, radius, k, c = (0.0, 0.0), 2.0, 20.0, 8.0
    ps = np.concatenate([particle(orc, np.eye(2), 0, C=(0, 0, 0, 0), v=(0, 0), x=0.1, y=0.12),
                         particle(orc, np.eye(2), 2, lam=40.0, mu=0.0, density=2.0, C=(0, 0, 0, 0), v=(0, 0),
                                  x=0.1, y=0.12),
                         particle(orc, np.eye(2), 3, lam=40.0, mu=0.0, density=1.5, C=(0, 0, 0, 0), v=(0, 0),
                                  x=0.1, y=0.12)])
    cfg = config(orc, drag=(target[0], target[1], radius, k, c))
    grid = orc.grid_f64(ps, cfg)
    F = orc.GRID_FIELDS
    W = CELLS + 3
    for mname, vx, vy in (("m", "vx", "vy"), ("wm", "wvx", "wvy"), ("om", "ovx", "ovy")):
        for key in np.nonzero(grid[:, F.index(mname)] > 0)[0]:
            xi, yi = key % W - 1, key // W - 1
            got = grid[key, [F.index(vx), F.index(vy)]]
            if xi < 0 or yi < 0:
                want = np.zeros(2)
            else:
                want = drag_node_velocity((0.0, 0.0), (xi * DX, yi * DX), target, radius, k, c)
                # the wall cushion scales the velocity components towards a wall
                cushion = min(8.0, CELLS / 4)
                want = want * np.array([1.0 if want[0] >= 0 else min(1.0, xi / cushion),
                                        1.0 if want[1] >= 0 else min(1.0, yi / cushion)])
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
