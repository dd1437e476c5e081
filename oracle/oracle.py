"""ctypes front end of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module, and only as the checker.  The product (the HIP library and
its host layer under ``powder-play_amd/``) never imports it.

The C sources (``mpm_oracle.c``) restate ``stepMpm``
(src/lib/mpm/reference/solver.ts:396-824) in FP64 and the WGSL fixed-point
P2G (src/lib/mpm/kernels/shaders.ts:808-1070) as an ordered f32 sequence.
``mpm_adjoint.c`` holds the FP64 differentiable rollout of the fixed-corotated
solid and its hand-written reverse pass (the checker of the GPU adjoint).
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libmpm_oracle.so"
# A deliberately broken copy of the oracle, built by tests/test_oracle_mutants.py
# to show that the closed-form tests catch it, is loaded in its place when this
# names it.
_LIB_OVERRIDE = os.environ.get("MPM_ORACLE_LIB")

PARTICLE_DTYPE = np.dtype(
    [
        ("x", "f8"), ("y", "f8"), ("vx", "f8"), ("vy", "f8"),
        ("c00", "f8"), ("c01", "f8"), ("c10", "f8"), ("c11", "f8"),
        ("f00", "f8"), ("f01", "f8"), ("f10", "f8"), ("f11", "f8"),
        ("plastic_volume", "f8"), ("lambda_", "f8"), ("mu", "f8"),
        ("density", "f8"), ("volume", "f8"), ("param", "f8"),
        ("model", "i4"), ("bonded_vertex", "i4"),
    ],
    align=True,
)
assert PARTICLE_DTYPE.itemsize == 152

MODEL_NEO_HOOKEAN, MODEL_SNOW, MODEL_WATER, MODEL_OIL, MODEL_POWDER = 0, 1, 2, 3, 4
MODEL_REMOVED = -1

ERRORS = {0: "ok", -1: "invalid input", -2: "non-finite", -3: "non-positive determinant",
          -4: "substep limit", -5: "unsupported"}


class Config(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("cells_x", ctypes.c_int32), ("cells_y", ctypes.c_int32),
        ("cell_size", ctypes.c_double), ("particle_mass", ctypes.c_double),
        ("particle_volume", ctypes.c_double), ("gravity_x", ctypes.c_double),
        ("gravity_y", ctypes.c_double), ("wall_friction", ctypes.c_double),
        ("drag_x", ctypes.c_double), ("drag_y", ctypes.c_double), ("drag_radius", ctypes.c_double),
        ("drag_stiffness", ctypes.c_double), ("drag_damping", ctypes.c_double),
    ]


class Diagnostics(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in (
        "input_mass", "grid_mass", "output_mass", "input_px", "input_py",
        "grid_px", "grid_py", "output_px", "output_py")]


class ParamsF32(ctypes.Structure):
    _fields_ = [
        ("cells_x", ctypes.c_uint32), ("cells_y", ctypes.c_uint32),
        ("node_width", ctypes.c_uint32), ("node_height", ctypes.c_uint32),
        ("particle_count", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
        ("cell_size", ctypes.c_float), ("particle_mass", ctypes.c_float),
        ("particle_volume", ctypes.c_float), ("dt", ctypes.c_float),
        ("gradient_scale", ctypes.c_float),
    ]


class OracleError(RuntimeError):
    def __init__(self, code: int):
        super().__init__(f"oracle step failed: {ERRORS.get(code, code)}")
        self.code = code


_lib = None


def build() -> pathlib.Path:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if _LIB_OVERRIDE:
            _lib = ctypes.CDLL(_LIB_OVERRIDE)
        else:
            if not LIB_PATH.exists():
                build()
            _lib = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        _lib.orc_step_f64.argtypes = [P, ctypes.c_int32, ctypes.POINTER(Config), P, ctypes.POINTER(Diagnostics)]
        _lib.orc_run_f64.argtypes = [P, ctypes.c_int32, ctypes.POINTER(Config), ctypes.c_int32]
        _lib.orc_grid_f64.argtypes = [P, ctypes.c_int32, ctypes.POINTER(Config), P]
        _lib.orc_prepare_f32.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ParamsF32), P]
        _lib.orc_prepare_f32.restype = ctypes.c_uint32
        _lib.orc_fixed_p2g_f32.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ParamsF32), P]
        _lib.orc_fixed_p2g_f32.restype = ctypes.c_uint32
        _lib.orc_decode_fixed_f32.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ParamsF32), P]
        _lib.orc_bin.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ParamsF32), P, P, P]
        _lib.orc_bin.restype = ctypes.c_uint32
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_config(dt, cells_x, cells_y, cell_size, particle_mass, particle_volume,
                gravity=(0.0, 0.0), wall_friction=0.0, drag=None) -> Config:
    """``drag`` = (target x, target y, radius, stiffness, damping) or None."""
    d = (0.0, 0.0, 0.0, 0.0, 0.0) if drag is None else tuple(float(v) for v in drag)
    return Config(dt, cells_x, cells_y, cell_size, particle_mass, particle_volume,
                  gravity[0], gravity[1], wall_friction, *d)


def particles_from_packed(packed20: np.ndarray) -> np.ndarray:
    """20-word f32 AoS records (layout.ts:385-437) -> FP64 oracle particles."""
    f = packed20.reshape(-1, 20).view(np.float32).astype(np.float64)
    u = packed20.reshape(-1, 20).view(np.uint32)
    p = np.zeros(f.shape[0], PARTICLE_DTYPE)
    names = ["x", "y", "vx", "vy", "c00", "c01", "c10", "c11", "f00", "f01", "f10", "f11", "plastic_volume"]
    for k, name in enumerate(names):
        p[name] = f[:, k]
    p["lambda_"], p["mu"], p["density"], p["volume"], p["param"] = f[:, 14], f[:, 15], f[:, 16], f[:, 17], f[:, 18]
    p["model"] = u[:, 13].astype(np.int64).astype(np.int32)
    p["bonded_vertex"] = u[:, 19].astype(np.int32)
    return p


def step_f64(particles: np.ndarray, cfg: Config):
    out = np.zeros_like(particles)
    diag = Diagnostics()
    code = lib().orc_step_f64(_ptr(particles), len(particles), ctypes.byref(cfg), _ptr(out), ctypes.byref(diag))
    if code:
        raise OracleError(code)
    return out, diag


def run_f64(particles: np.ndarray, cfg: Config, steps: int) -> np.ndarray:
    p = particles.copy()
    code = lib().orc_run_f64(_ptr(p), len(p), ctypes.byref(cfg), steps)
    if code:
        raise OracleError(code)
    return p


def run_f64_parallel(particles: np.ndarray, cfg: Config, steps: int, threads: int) -> np.ndarray:
    """run_f64 on `threads` OpenMP threads (the all-core CPU baseline; equal
    to run_f64 to rounding)."""
    L = lib()
    L.orc_run_f64_parallel.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Config), ctypes.c_int32,
                                       ctypes.c_int32]
    p = particles.copy()
    code = L.orc_run_f64_parallel(_ptr(p), len(p), ctypes.byref(cfg), steps, threads)
    if code:
        raise OracleError(code)
    return p


GRID_FIELDS = ["m", "qx", "qy", "vx", "vy", "wm", "wqx", "wqy", "wvx", "wvy", "wV",
               "om", "oqx", "oqy", "ovx", "ovy", "oV"]


def grid_f64(particles: np.ndarray, cfg: Config) -> np.ndarray:
    nodes = (cfg.cells_x + 3) * (cfg.cells_y + 3)
    out = np.zeros((nodes, len(GRID_FIELDS)), np.float64)
    code = lib().orc_grid_f64(_ptr(particles), len(particles), ctypes.byref(cfg), _ptr(out))
    if code:
        raise OracleError(code)
    return out


# Rounding scale of the float32 node sums relative to their magnitude, used
# to tell a threshold decision of the contact that float32 may take the other
# way (the sums of up to ~20 rounded products each carry a few 1e-7).
CONTACT_ROUNDING = 1e-6


def contact_near_threshold(grid: np.ndarray, cfg: Config, rounding: float = CONTACT_ROUNDING):
    """Nodes whose water/oil contact decision lies within float32 rounding of
    its threshold, from the FP64 node sums of grid_f64 (the P2G of one
    microstep).  Restates in FP64 the three decisions of solveLiquidGridContact
    (src/lib/mpm/reference/solver.ts:1293-1409, the contact branch of
    grid_node in mpm_oracle.c; WGSL kernels/shaders.ts:1527-1594):

    * the length of the volume-gradient difference g = grad V_oil - grad V_water
      (the float32 kernel keeps nodes above 1e-8 of the unscaled difference,
      the reference above 1e-12 of g);
    * the normal's orientation flip, contrast * align < -1e-3;
    * the impulse, rel < minsep.

    The normal n = g / |g| of nearly cancelling gradients is ill-conditioned:
    float32 volumes with relative error `rounding` move it by about
    rounding * (|grad V_w| + |grad V_o| + V_w + V_o) / |g|.  A node is flagged
    when a decision's margin is within that perturbation of the quantities it
    compares (or the normal is undetermined).  Returns (mask of shape
    (cells_y + 3, cells_x + 3), number of contact nodes)."""
    W, H = cfg.cells_x + 3, cfg.cells_y + 3
    f = {n: grid[:, i].reshape(H, W) for i, n in enumerate(GRID_FIELDS)}
    dx, dt = float(cfg.cell_size), float(cfg.dt)
    gvx, gvy = float(cfg.gravity_x), float(cfg.gravity_y)
    vs = 1.0 / (2 * dx * dx)

    def grad(V, axis):  # central differences, zero volume outside the node grid
        P = np.pad(V, 1)
        return ((P[1:-1, 2:] - P[1:-1, :-2]) if axis == 0 else (P[2:, 1:-1] - P[:-2, 1:-1])) * vs

    wV, oV, wm, om = f["wV"], f["oV"], f["wm"], f["om"]
    wgx, wgy, ogx, ogy = grad(wV, 0), grad(wV, 1), grad(oV, 0), grad(oV, 1)
    gx, gy = ogx - wgx, ogy - wgy
    gl = np.hypot(gx, gy)
    yi, xi = np.mgrid[0:H, 0:W] - 1
    contact = ((wm > 0) & (om > 0) & (wV > 0) & (oV > 0) & (xi >= 0) & (xi <= cfg.cells_x) & (yi >= 0) &
               (yi <= cfg.cells_y))
    with np.errstate(all="ignore"):
        nx, ny = gx / gl, gy / gl
        wd, od = wm / wV, om / oV
        contrast = (od - wd) / np.maximum(wd, od)
        gm = np.hypot(gvx, gvy)
        align = (nx * gvx + ny * gvy) / gm if gm > 0 else np.zeros_like(nx)
        m_flip = contrast * align + 1e-3
        side = np.where(m_flip < 0, -1.0, 1.0)
        vwx, vwy = f["wqx"] / wm + dt * gvx, f["wqy"] / wm + dt * gvy
        vox, voy = f["oqx"] / om + dt * gvx, f["oqy"] / om + dt * gvy
        rel = ((vox - vwx) * nx + (voy - vwy) * ny) * side
        overlap = np.minimum(1.0, 2 * np.minimum(wV, oV) / (wV + oV))
        minsep = np.maximum(0.25 * overlap * dx / 0.05, np.maximum(-rel, 0.0))
        m_imp = rel - minsep
        dn = rounding * (np.abs(wgx) + np.abs(wgy) + np.abs(ogx) + np.abs(ogy) + vs * (wV + oV)) / gl
        speed = np.hypot(vwx, vwy) + np.hypot(vox, voy)
        near = ((np.abs(m_flip) <= np.abs(contrast) * dn + rounding) |
                (np.abs(m_imp) <= np.hypot(vox - vwx, voy - vwy) * dn + rounding * (speed + 1.0)) |
                ~np.isfinite(dn) | (dn > 0.1) | (2 * gl * dx * dx <= 1e-6))
    return contact & near, int(contact.sum())


def particles_touching(mask: np.ndarray, x: np.ndarray, y: np.ndarray, cell_size: float) -> np.ndarray:
    """Whether each particle's 3x3 stencil (base floor(x/dx - 0.5), nodes
    base + 1 .. base + 3 in node storage) holds a masked node."""
    bx = np.floor(np.asarray(x, np.float64) / cell_size - 0.5).astype(np.int64)
    by = np.floor(np.asarray(y, np.float64) / cell_size - 0.5).astype(np.int64)
    H, W = mask.shape
    hit = np.zeros(bx.shape, bool)
    for oy in range(3):
        for ox in range(3):
            sx, sy = np.clip(bx + 1 + ox, 0, W - 1), np.clip(by + 1 + oy, 0, H - 1)
            hit |= mask[sy, sx]
    return hit


def params_f32(cells_x, cells_y, cell_size, particle_mass, particle_volume, dt, n) -> ParamsF32:
    cs = np.float32(cell_size)
    return ParamsF32(cells_x, cells_y, cells_x + 3, cells_y + 3, n, (cells_x + 3) * (cells_y + 3),
                     cs, np.float32(particle_mass), np.float32(particle_volume), np.float32(dt),
                     np.float32(4.0 / (float(cell_size) * float(cell_size))))


def prepare_f32(packed20: np.ndarray, prm: ParamsF32):
    n = packed20.size // 20
    a = np.zeros((n, 4), np.float32)
    flags = lib().orc_prepare_f32(_ptr(np.ascontiguousarray(packed20, np.uint32)), n, ctypes.byref(prm), _ptr(a))
    return a, flags


def fixed_p2g_f32(packed20: np.ndarray, prm: ParamsF32):
    n = packed20.size // 20
    g = np.zeros((prm.node_count, 11), np.int32)
    flags = lib().orc_fixed_p2g_f32(_ptr(np.ascontiguousarray(packed20, np.uint32)), n, ctypes.byref(prm), _ptr(g))
    return g, flags


def decode_fixed_f32(grid: np.ndarray, prm: ParamsF32) -> np.ndarray:
    out = np.zeros((prm.node_count, 11), np.float32)
    lib().orc_decode_fixed_f32(_ptr(np.ascontiguousarray(grid, np.int32)), prm.node_count, ctypes.byref(prm), _ptr(out))
    return out


def bin_particles(packed20: np.ndarray, prm: ParamsF32):
    n = packed20.size // 20
    key = np.zeros(n, np.uint32)
    start = np.zeros(prm.node_count + 1, np.uint32)
    perm = np.zeros(n, np.uint32)
    flags = lib().orc_bin(_ptr(np.ascontiguousarray(packed20, np.uint32)), n, ctypes.byref(prm),
                          _ptr(key), _ptr(start), _ptr(perm))
    return key, start, perm, flags


# ---------------- differentiable rollout (mpm_adjoint.c) ----------------

class DiffConfig(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("cell_size", ctypes.c_double), ("particle_mass", ctypes.c_double),
        ("particle_volume", ctypes.c_double), ("gravity_x", ctypes.c_double), ("gravity_y", ctypes.c_double),
        ("wall_friction", ctypes.c_double), ("cells_x", ctypes.c_int32), ("cells_y", ctypes.c_int32),
        ("steps", ctypes.c_int32),
    ]


def diff_config(cfg: dict, steps: int | None = None) -> DiffConfig:
    """DiffConfig from an elastic_block-style config dict."""
    return DiffConfig(cfg["dt"], cfg["cellSize"], cfg["particleMass"], cfg["particleVolume"],
                      cfg["gravity"][0], cfg["gravity"][1], cfg.get("wallFriction", 0.0),
                      cfg["cellsX"], cfg["cellsY"], cfg["steps"] if steps is None else steps)


def _diff_lib():
    L = lib()
    if not getattr(L, "_diff_ready", False):
        P = ctypes.c_void_p
        C = ctypes.POINTER(DiffConfig)
        L.orc_diff_step.argtypes = [C, ctypes.c_int32, P, P, P, P]
        L.orc_diff_forward.argtypes = [C, ctypes.c_int32, P, P, P, P]
        L.orc_diff_backward.argtypes = [C, ctypes.c_int32, P, P, P, P, P, P, P, P]
        L._diff_ready = True
    return L


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float64)


def diff_step(state: np.ndarray, lam, mu, cfg: DiffConfig) -> np.ndarray:
    s, la, m = _f64(state), _f64(lam), _f64(mu)
    out = np.zeros_like(s)
    code = _diff_lib().orc_diff_step(ctypes.byref(cfg), len(la), _ptr(s), _ptr(la), _ptr(m), _ptr(out))
    if code:
        raise OracleError(code)
    return out


def diff_forward(state0: np.ndarray, lam, mu, cfg: DiffConfig) -> np.ndarray:
    s, la, m = _f64(state0), _f64(lam), _f64(mu)
    out = np.zeros_like(s)
    code = _diff_lib().orc_diff_forward(ctypes.byref(cfg), len(la), _ptr(s), _ptr(la), _ptr(m), _ptr(out))
    if code:
        raise OracleError(code)
    return out


def diff_backward(state0: np.ndarray, lam, mu, cfg: DiffConfig, bar_final: np.ndarray):
    """Returns (dL/dstate0 (n, 12), dL/dlambda (n,), dL/dmu (n,), confinements)."""
    s, la, m, b = _f64(state0), _f64(lam), _f64(mu), _f64(bar_final)
    bar0 = np.zeros_like(s)
    gl = np.zeros(len(la))
    gm = np.zeros(len(la))
    conf = ctypes.c_int64(0)
    code = _diff_lib().orc_diff_backward(ctypes.byref(cfg), len(la), _ptr(s), _ptr(la), _ptr(m), _ptr(b),
                                         _ptr(bar0), _ptr(gl), _ptr(gm), ctypes.byref(conf))
    if code:
        raise OracleError(code)
    return bar0, gl, gm, conf.value


def com_loss(state: np.ndarray, target) -> tuple[float, np.ndarray]:
    """L = |mean(x) - target|^2 for equal-mass particles and dL/dstate."""
    d = state[:, :2].mean(axis=0) - np.asarray(target, np.float64)
    bar = np.zeros_like(state)
    bar[:, :2] = 2.0 * d / state.shape[0]
    return float(d @ d), bar


# ---------------- config-2 bodies and fracture (orc_step_bodies_f64) ----------------

class BodyConfig(ctypes.Structure):
    _fields_ = [("bond_rest_length", ctypes.c_double), ("bond_stiffness", ctypes.c_double),
                ("bond_reference_dt", ctypes.c_double)]


BOND_DTYPE = np.dtype([("first", "i4"), ("second", "i4"), ("rest_length", "f8"), ("failure_strain", "f8"),
                       ("damage", "f8"), ("broken", "i4"), ("pad", "i4")], align=True)
assert BOND_DTYPE.itemsize == 40


def _bodies_lib():
    L = lib()
    if not getattr(L, "_bodies_ready", False):
        P = ctypes.c_void_p
        L.orc_step_bodies_f64.argtypes = [P, ctypes.c_int32, P, ctypes.POINTER(Config), ctypes.POINTER(BodyConfig),
                                          P, ctypes.c_int32, P, ctypes.POINTER(Diagnostics)]
        L.orc_triangle_sample.argtypes = [ctypes.c_double, ctypes.c_double, P, P]
        L.orc_triangle_pair_contact.argtypes = [P, P, ctypes.c_double, P]
        L._bodies_ready = True
    return L


def step_bodies_f64(particles: np.ndarray, object_ids, cfg: Config, body_cfg: BodyConfig, bonds=None):
    """One FP64 step with bonded triangles and fracture bonds.  Returns
    (particles out, bonds out, diagnostics)."""
    oid = np.ascontiguousarray(object_ids, np.uint32)
    b = np.zeros(0, BOND_DTYPE) if bonds is None else np.ascontiguousarray(bonds, BOND_DTYPE).copy()
    out = np.zeros_like(particles)
    diag = Diagnostics()
    code = _bodies_lib().orc_step_bodies_f64(_ptr(particles), len(particles), _ptr(oid), ctypes.byref(cfg),
                                             ctypes.byref(body_cfg), _ptr(b) if len(b) else None, len(b), _ptr(out),
                                             ctypes.byref(diag))
    if code:
        raise OracleError(code)
    return out, b, diag


def triangle_sample(px: float, py: float, tri) -> np.ndarray:
    t = np.ascontiguousarray(tri, np.float64).reshape(6)
    out = np.zeros(6)
    code = _bodies_lib().orc_triangle_sample(px, py, _ptr(t), _ptr(out))
    if code:
        raise OracleError(code)
    return out


def triangle_pair_contact(a, b, skin: float):
    A = np.ascontiguousarray(a, np.float64).reshape(6)
    B = np.ascontiguousarray(b, np.float64).reshape(6)
    out = np.zeros(3)
    hit = _bodies_lib().orc_triangle_pair_contact(_ptr(A), _ptr(B), skin, _ptr(out))
    return (out if hit else None)
