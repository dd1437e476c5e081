"""mpm355 — MI355X-native 2D MLS-MPM step path (host side).

The compute path is ``libmpm355.so`` (hand-written gfx950 kernels behind the
C-ABI in include/mpm355.h); this package is the host mirror of the reference's
solver API (src/lib/mpm/kernels/solver.ts:106-144) plus its layout, CFL
scheduling, scene generators, snapshot files, the simulation worker's
request protocol (``worker``), and the differentiable rollout of a
fixed-corotated solid (``adjoint``).  Import it as ``powder_play_amd`` (the
directory name carries a hyphen; see ``load_package`` in ``__graft_entry__``).
"""
from . import adjoint, autotune, cfl, layout, native, scenes, snapshot, worker  # noqa: F401
from .layout import ParticleArrays, RangeError  # noqa: F401
from .solver import (  # noqa: F401
    MpmMlsGpuBatch,
    MpmMlsGpuNumericalError,
    MpmMlsGpuSolver,
    create_experimental_mpm_mls_atomic_gpu_solver,
    create_mpm_mls_gpu_solver,
)

__all__ = [
    "adjoint", "autotune", "cfl", "layout", "native", "scenes", "snapshot", "worker", "ParticleArrays", "RangeError", "MpmMlsGpuNumericalError",
    "MpmMlsGpuSolver", "MpmMlsGpuBatch", "create_mpm_mls_gpu_solver", "create_experimental_mpm_mls_atomic_gpu_solver",
]
