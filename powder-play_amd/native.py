"""ctypes binding of ``libmpm355.so`` (declared in include/mpm355.h).

The library is the product: every solver call in this package goes through it.
There is no CPU fallback — if the library is missing or cannot load, the
import of the solver raises.
"""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libmpm355.so"

MPM_OK, MPM_EINVAL, MPM_EDEVICE, MPM_ESTATE, MPM_ECAPACITY = 0, 1, 2, 3, 4
MPM_P2G_GATHER, MPM_P2G_FIXED_POINT = 0, 1
MPM_FLAG_KEEP_FIXED_GRID = 1

# Every symbol include/mpm355.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
SIGNATURES = {
    "mpm_create": (ctypes.c_int, [_P, _P, _P, _U32, ctypes.POINTER(_P)]),
    "mpm_step_batch": (ctypes.c_int, [_P, _P, _U32, _U32]),
    "mpm_read_compact": (ctypes.c_int, [_P, _P, _P, _P]),
    "mpm_read_status": (ctypes.c_int, [_P, _P]),
    "mpm_clear_status": (ctypes.c_int, [_P]),
    "mpm_read_particles": (ctypes.c_int, [_P, _P, _U32, _P, _P]),
    "mpm_append_particles": (ctypes.c_int, [_P, _P, _P, _U32]),
    "mpm_set_interaction": (ctypes.c_int, [_P, _P]),
    "mpm_set_fracture_bonds": (ctypes.c_int, [_P, _P, _U32]),
    "mpm_read_fracture_bonds": (ctypes.c_int, [_P, _P, _U32]),
    "mpm_get_counts": (ctypes.c_int, [_P, _P, _P, _P]),
    "mpm_remove_in_outlet": (ctypes.c_int, [_P, _P, _P]),
    "mpm_gather_occupancy": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "mpm_gather_occupancy_static": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "mpm_pick": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
    "mpm_clone": (ctypes.c_int, [_P, _U32, _U32, ctypes.c_float, ctypes.c_float, _U32]),
    "mpm_read_fixed_grid": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "mpm_read_grid": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "mpm_read_bins": (ctypes.c_int, [_P, _P, _U32, _P, ctypes.c_size_t]),
    "mpm_get_buckets": (ctypes.c_int, [_P, _P, _P, _P]),
    "mpm_set_diagnostics": (ctypes.c_int, [_P, _U32]),
    "mpm_read_diagnostics": (ctypes.c_int, [_P, _P]),
    "mpm_set_timing": (ctypes.c_int, [_P, _U32]),
    "mpm_read_timing": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_uint64)]),
    "mpm_set_tuning": (ctypes.c_int, [_P, _P]),
    "mpm_get_tuning": (ctypes.c_int, [_P, _P]),
    "mpm_batch_set_tuning": (ctypes.c_int, [_P, _P]),
    "mpm_batch_get_tuning": (ctypes.c_int, [_P, _P]),
    "mpm_device_arch": (ctypes.c_int, [ctypes.c_int32, ctypes.c_char_p, _U32, _P]),
    "mpm_batch_create": (ctypes.c_int, [_P, _P, _P, _U32, ctypes.POINTER(_P)]),
    "mpm_batch_step": (ctypes.c_int, [_P, _P, _U32, _U32]),
    "mpm_batch_scene": (_P, [_P, _U32]),
    "mpm_batch_count": (_U32, [_P]),
    "mpm_batch_destroy": (None, [_P]),
    "mpm_batch_last_error": (ctypes.c_char_p, [_P]),
    "mpm_diff_create": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "mpm_diff_forward": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "mpm_diff_backward": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "mpm_diff_read_status": (ctypes.c_int, [_P, _P]),
    "mpm_diff_read_timing": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "mpm_diff_get_layout": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(ctypes.c_size_t)]),
    "mpm_diff_destroy": (None, [_P]),
    "mpm_diff_last_error": (ctypes.c_char_p, [_P]),
    "mpm_destroy": (None, [_P]),
    "mpm_last_error": (ctypes.c_char_p, [_P]),
}

MPM_DIFF_STATE_FLOATS = 12
MPM_DIFF_STENCIL, MPM_DIFF_DEFORMATION, MPM_DIFF_NONFINITE, MPM_DIFF_SUBSTEPS, MPM_DIFF_CONFINED = 1, 2, 4, 8, 16
MPM_DIFF_NO_GRAPHS = 1


class MpmDiffConfig(ctypes.Structure):
    _fields_ = [
        ("particle_count", _U32), ("cells_x", _U32), ("cells_y", _U32), ("steps", _U32),
        ("segment_steps", _U32), ("cell_size", ctypes.c_float), ("dt", ctypes.c_float),
        ("particle_mass", ctypes.c_float), ("particle_volume", ctypes.c_float),
        ("gravity_x", ctypes.c_float), ("gravity_y", ctypes.c_float), ("wall_friction", ctypes.c_float),
        ("device", ctypes.c_int32), ("flags", _U32),
    ]


class MpmConfig(ctypes.Structure):
    _fields_ = [
        ("cells_x", _U32), ("cells_y", _U32), ("cell_size", ctypes.c_float),
        ("particle_mass", ctypes.c_float), ("particle_volume", ctypes.c_float),
        ("wall_friction", ctypes.c_float), ("capacity", _U32), ("p2g_mode", _U32),
        ("flags", _U32), ("device", ctypes.c_int32), ("bucket_slots", _U32),
    ]


class NativeError(RuntimeError):
    """A C-ABI call failed; ``code`` is the MPM_E* value."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        # One HIP runtime per process.  PyTorch wheels bundle their own
        # libamdhip64 (same soname as ROCm's); loaded first, it is the one this
        # library binds to.  Loaded after this library, torch would bring a
        # second runtime that cannot open the GPU this one holds.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """Owns one ``mpm_ctx``; thin, checked wrappers around the C-ABI."""

    @classmethod
    def borrowed(cls, handle, cfg: MpmConfig):
        """A view of a context owned elsewhere (a scene of a batch)."""
        self = cls.__new__(cls)
        self.lib = load()
        self.handle = ctypes.c_void_p(handle)
        self.cfg = cfg
        self.owned = False
        return self

    def __init__(self, cfg: MpmConfig, packed: np.ndarray, object_ids: np.ndarray | None):
        self.owned = True
        self.lib = load()
        packed = np.ascontiguousarray(packed, dtype=np.uint32).reshape(-1, 20)
        ids = None if object_ids is None else np.ascontiguousarray(object_ids, dtype=np.uint32)
        handle = ctypes.c_void_p()
        code = self.lib.mpm_create(ctypes.byref(cfg), _ptr(packed), _ptr(ids), packed.shape[0], ctypes.byref(handle))
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_last_error(None).decode())
        self.handle = handle
        self.cfg = cfg

    def _check(self, code: int):
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_last_error(self.handle).decode())

    def step_batch(self, params: bytes, steps: int, measure: bool):
        buf = (ctypes.c_uint8 * 144).from_buffer_copy(params)
        self._check(self.lib.mpm_step_batch(self.handle, ctypes.addressof(buf), steps, 1 if measure else 0))

    def read_compact(self):
        speed = np.zeros(1, np.float32)
        done = np.zeros(1, np.uint32)
        status = np.zeros(8, np.uint32)
        self._check(self.lib.mpm_read_compact(self.handle, _ptr(speed), _ptr(done), _ptr(status)))
        return float(speed[0]), int(done[0]), status

    def read_status(self) -> np.ndarray:
        status = np.zeros(8, np.uint32)
        self._check(self.lib.mpm_read_status(self.handle, _ptr(status)))
        return status

    def clear_status(self):
        self._check(self.lib.mpm_clear_status(self.handle))

    def counts(self):
        n, cap, nodes = (ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32())
        self._check(self.lib.mpm_get_counts(self.handle, ctypes.byref(n), ctypes.byref(cap), ctypes.byref(nodes)))
        return n.value, cap.value, nodes.value

    def read_particles(self, indices=None):
        if indices is None:
            n = self.counts()[0]
            idx = None
        else:
            idx = np.ascontiguousarray(indices, dtype=np.uint32)
            n = idx.size
        out = np.zeros((n, 20), np.uint32)
        ids = np.zeros(n, np.uint32)
        self._check(self.lib.mpm_read_particles(self.handle, _ptr(idx), n, _ptr(out), _ptr(ids)))
        return out, ids

    def append(self, packed: np.ndarray, object_ids: np.ndarray | None):
        packed = np.ascontiguousarray(packed, dtype=np.uint32).reshape(-1, 20)
        ids = None if object_ids is None else np.ascontiguousarray(object_ids, dtype=np.uint32)
        self._check(self.lib.mpm_append_particles(self.handle, _ptr(packed), _ptr(ids), packed.shape[0]))

    def set_fracture_bonds(self, records: np.ndarray):
        rec = np.ascontiguousarray(records, dtype=np.uint32).reshape(-1, 6)
        self._check(self.lib.mpm_set_fracture_bonds(self.handle, _ptr(rec), rec.shape[0]))

    def read_fracture_bonds(self, count: int) -> np.ndarray:
        out = np.zeros((count, 6), np.uint32)
        self._check(self.lib.mpm_read_fracture_bonds(self.handle, _ptr(out), count))
        return out

    def set_interaction(self, values):
        arr = np.zeros(8, np.float32)
        arr[: len(values)] = values
        self._check(self.lib.mpm_set_interaction(self.handle, _ptr(arr)))

    def gather_occupancy(self) -> np.ndarray:
        """(node_count, 2) float32: water and oil surface occupancy per node."""
        nodes = self.counts()[2]
        out = np.zeros((nodes, 2), np.float32)
        self._check(self.lib.mpm_gather_occupancy(self.handle, _ptr(out), out.size))
        return out

    def gather_occupancy_static(self) -> np.ndarray:
        """gather_occupancy through the allocation-free entry point (the
        per-node scratch is a static device array of the library)."""
        nodes = self.counts()[2]
        out = np.zeros((nodes, 2), np.float32)
        self._check(self.lib.mpm_gather_occupancy_static(self.handle, _ptr(out), out.size))
        return out

    def remove_in_outlet(self, bounds) -> int:
        b = np.ascontiguousarray(bounds, np.float32).reshape(4)
        out = np.zeros(1, np.uint32)
        self._check(self.lib.mpm_remove_in_outlet(self.handle, _ptr(b), _ptr(out)))
        return int(out[0])

    def pick(self, x: float, y: float, radius: float) -> np.ndarray:
        """The 16-byte pick result: index (0xFFFFFFFF = none), x bits, y bits, object id."""
        out = np.zeros(4, np.uint32)
        self._check(self.lib.mpm_pick(self.handle, x, y, radius, _ptr(out)))
        return out

    def clone(self, start: int, count: int, offset_x: float, offset_y: float, object_id: int):
        self._check(self.lib.mpm_clone(self.handle, start, count, offset_x, offset_y, object_id))

    def read_fixed_grid(self) -> np.ndarray:
        nodes = self.counts()[2]
        out = np.zeros((nodes, 11), np.int32)
        self._check(self.lib.mpm_read_fixed_grid(self.handle, _ptr(out), out.size))
        return out

    def read_grid(self) -> np.ndarray:
        nodes = self.counts()[2]
        out = np.zeros((nodes, 8), np.float32)
        self._check(self.lib.mpm_read_grid(self.handle, _ptr(out), out.size))
        return out

    def buckets(self):
        bx, by, cap = (ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32())
        self._check(self.lib.mpm_get_buckets(self.handle, ctypes.byref(bx), ctypes.byref(by), ctypes.byref(cap)))
        return bx.value, by.value, cap.value

    def read_bins(self):
        n = self.counts()[0]
        bx, by, _ = self.buckets()
        keys = np.zeros(n, np.uint32)
        counts = np.zeros(bx * by, np.uint32)
        self._check(self.lib.mpm_read_bins(self.handle, _ptr(keys), n, _ptr(counts), counts.size))
        return keys, counts

    def set_diagnostics(self, enable: bool):
        self._check(self.lib.mpm_set_diagnostics(self.handle, 1 if enable else 0))

    def read_diagnostics(self) -> np.ndarray:
        out = np.zeros(9, np.float64)
        self._check(self.lib.mpm_read_diagnostics(self.handle, _ptr(out)))
        return out

    def set_timing(self, enable: bool):
        self._check(self.lib.mpm_set_timing(self.handle, 1 if enable else 0))

    def read_timing(self, kernel: str):
        mean = ctypes.c_double()
        launches = ctypes.c_uint64()
        self._check(self.lib.mpm_read_timing(self.handle, kernel.encode(), ctypes.byref(mean), ctypes.byref(launches)))
        return mean.value, launches.value

    def set_tuning(self, tuning: "MpmTuning"):
        self._check(self.lib.mpm_set_tuning(self.handle, ctypes.byref(tuning)))

    def get_tuning(self) -> "MpmTuning":
        out = MpmTuning()
        self._check(self.lib.mpm_get_tuning(self.handle, ctypes.byref(out)))
        return out

    def close(self):
        if getattr(self, "handle", None):
            if getattr(self, "owned", True):
                self.lib.mpm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MpmTuning(ctypes.Structure):
    """``mpm_tuning``: launch shape of the step kernels and bucket capacity."""
    _fields_ = [("p2g_tile", _U32), ("g2p_tile", _U32), ("g2p_threads", _U32), ("bucket_capacity", _U32),
                ("tile_queue", _U32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}

    @classmethod
    def from_dict(cls, d: dict) -> "MpmTuning":
        return cls(*(int(d.get(k, 0)) for k, _ in cls._fields_))


def device_arch(device: int = 0) -> tuple[str, int]:
    """(architecture name, compute units) of a HIP device."""
    lib = load()
    buf = ctypes.create_string_buffer(128)
    cus = ctypes.c_uint32()
    code = lib.mpm_device_arch(device, buf, 128, ctypes.byref(cus))
    if code != MPM_OK:
        raise NativeError(code, lib.mpm_last_error(None).decode())
    return buf.value.decode(), int(cus.value)


class Batch:
    """Owns one ``mpm_batch``: scenes stepped by the same kernel launches."""

    def __init__(self, cfgs: list[MpmConfig], packed: list[np.ndarray]):
        self.lib = load()
        n = len(cfgs)
        arrs = [np.ascontiguousarray(p, dtype=np.uint32).reshape(-1, 20) for p in packed]
        self._keep = arrs
        cfg_arr = (MpmConfig * n)(*cfgs)
        ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in arrs])
        counts = np.array([a.shape[0] for a in arrs], np.uint32)
        handle = ctypes.c_void_p()
        code = self.lib.mpm_batch_create(ctypes.addressof(cfg_arr), ctypes.addressof(ptrs), _ptr(counts), n,
                                         ctypes.byref(handle))
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_batch_last_error(None).decode())
        self.handle = handle
        self.cfgs = cfgs
        self.scenes = [Context.borrowed(self.lib.mpm_batch_scene(handle, i), cfgs[i]) for i in range(n)]

    def step(self, params: list[bytes], steps: int, measure: bool):
        blob = b"".join(params)
        if len(blob) != 144 * len(self.scenes):
            raise ValueError("one 144-byte parameter block per scene")
        buf = (ctypes.c_uint8 * len(blob)).from_buffer_copy(blob)
        code = self.lib.mpm_batch_step(self.handle, ctypes.addressof(buf), steps, 1 if measure else 0)
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_batch_last_error(self.handle).decode())

    def set_tuning(self, tuning: "MpmTuning"):
        code = self.lib.mpm_batch_set_tuning(self.handle, ctypes.byref(tuning))
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_batch_last_error(self.handle).decode())

    def get_tuning(self) -> "MpmTuning":
        out = MpmTuning()
        code = self.lib.mpm_batch_get_tuning(self.handle, ctypes.byref(out))
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_batch_last_error(self.handle).decode())
        return out

    def close(self):
        if getattr(self, "handle", None):
            for s in self.scenes:
                s.handle = None
            self.lib.mpm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DiffRollout:
    """Owns one ``mpm_diff`` (the differentiable rollout).  Array arguments
    are raw addresses of host or device memory; callers size them."""

    def __init__(self, cfg: MpmDiffConfig):
        self.lib = load()
        handle = ctypes.c_void_p()
        code = self.lib.mpm_diff_create(ctypes.byref(cfg), ctypes.byref(handle))
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_diff_last_error(None).decode())
        self.handle = handle
        self.cfg = cfg

    def _check(self, code: int):
        if code != MPM_OK:
            raise NativeError(code, self.lib.mpm_diff_last_error(self.handle).decode())

    def forward(self, state0: int, lam: int, mu: int, state_final: int | None):
        self._check(self.lib.mpm_diff_forward(self.handle, state0, lam, mu, state_final))

    def backward(self, bar_final: int, bar_state0: int | None, grad_lambda: int | None, grad_mu: int | None):
        self._check(self.lib.mpm_diff_backward(self.handle, bar_final, bar_state0, grad_lambda, grad_mu))

    def status(self) -> int:
        out = np.zeros(1, np.uint32)
        self._check(self.lib.mpm_diff_read_status(self.handle, _ptr(out)))
        return int(out[0])

    def timing(self) -> tuple[float, float]:
        f, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mpm_diff_read_timing(self.handle, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def layout(self) -> tuple[int, int, int]:
        seg, ck = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        nbytes = ctypes.c_size_t()
        self._check(self.lib.mpm_diff_get_layout(self.handle, _ptr(seg), _ptr(ck), ctypes.byref(nbytes)))
        return int(seg[0]), int(ck[0]), nbytes.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.mpm_diff_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
