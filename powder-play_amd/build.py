"""Build step of the native library: hipcc for gfx950, in-tree output.

``python powder-play_amd/build.py`` (or ``__graft_entry__.build()``) compiles
``csrc/*.hip`` into ``libmpm355.so`` next to this file.  Flags:

* ``--offload-arch=gfx950`` — MI355X only.
* ``-ffp-contract=off`` — every f32 product and sum is rounded on its own, the
  fixed-point P2G's bit-exact contract depends on it (explicit ``fmaf`` calls
  stay fused).
* ``-Xclang -target-feature -Xclang -packed-fp32-ops`` — no packed FP32 ALU
  forms in generated code.
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
LIB = HERE / "libmpm355.so"
ADDON = HERE / "mpm355.node"
NODE_INCLUDE = pathlib.Path(os.environ.get("NODE_INCLUDE", "/usr/include/node"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPM355_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False) -> pathlib.Path:
    """Compile every HIP source and link the shared library (always rebuilds)."""
    objs = []
    for src in sources():
        obj = HERE / "_build" / (src.stem + ".o")
        obj.parent.mkdir(exist_ok=True)
        cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_addon(verbose: bool = False) -> pathlib.Path | None:
    """Compile the Node-API binding (napi/mpm355_napi.c) against node_api.h
    into mpm355.node beside libmpm355.so (found through $ORIGIN).  Returns
    None, with a message, when this machine has no Node-API headers."""
    if not (NODE_INCLUDE / "node_api.h").exists():
        print(f"[build] {NODE_INCLUDE}/node_api.h not found: Node-API addon not built", file=sys.stderr)
        return None
    tmp = ADDON.with_suffix(".node.tmp")
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-Wall", f"-I{NODE_INCLUDE}", f"-I{ROOT / 'include'}", "-o", str(tmp),
           str(HERE / "napi" / "mpm355_napi.c"), f"-L{HERE}", "-lmpm355", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, ADDON)
    return ADDON


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_addon(verbose="-v" in sys.argv)
    print(LIB)
