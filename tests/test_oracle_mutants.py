"""The closed-form pins of tests/test_oracle_pins.py must catch plausible
mistakes in the oracle.  Each case below plants one mistake in a private copy
of oracle/mpm_oracle.c (a wrong scale, a dropped factor, a flipped sign, a
transposed operand), builds it with the oracle's own flags, and runs the pin
tests against it in a child process: they must fail.  An unmodified copy,
built the same way, must pass.  CPU only; the mistakes live in the oracle's
copy, never in the product."""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle"
PINS = ROOT / "tests" / "test_oracle_pins.py"
CFLAGS = ["-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-fopenmp", "-shared"]

# (name, text in mpm_oracle.c, replacement, which occurrence)
MUTANTS = [
    ("stress scale 4/dx instead of 4/dx^2", "double ss = -c->dt * p->volume * gscale;",
     "double ss = -c->dt * p->volume * (4.0 / dx);", 0),
    ("stress scale doubled", "double ss = -c->dt * p->volume * gscale;",
     "double ss = -2.0 * c->dt * p->volume * gscale;", 0),
    ("Neo-Hookean pressure lambda (J - 1)", "double pressure = p->lambda * log(J);",
     "double pressure = p->lambda * (J - 1.0);", 0),
    ("Neo-Hookean P F instead of P F^T", "mat2 Ft = {F.m00, F.m10, F.m01, F.m11};", "mat2 Ft = F;", 0),
    ("Tait stress without its J factor", "? -J * p->lambda * (pow(1.0 / J, p->param) - 1.0)",
     "? -p->lambda * (pow(1.0 / J, p->param) - 1.0)", 0),
    ("snow hardening sign flipped", "SNOW_HARDENING * (1 - p->plastic_volume)",
     "SNOW_HARDENING * (p->plastic_volume - 1)", 0),
    ("snow volumetric term without J", "vol = lambda * J * (J - 1);", "vol = lambda * (J - 1);", 0),
    ("powder friction cap doubled", "double maxdev = (p->param * pressure) / (2 * p->mu);",
     "double maxdev = (p->param * pressure) / p->mu;", 0),
    ("G2P density correction sign flipped", "double rel = (pd - ld) / pd;", "double rel = (ld - pd) / pd;", 0),
    ("drag damping sign flipped", "- vel[ph][0] * c->drag_damping;", "+ vel[ph][0] * c->drag_damping;", 0),
    ("drag without falloff", "vel[ph][0] += c->dt * falloff * ax;", "vel[ph][0] += c->dt * ax;", 0),
]


def _build(tmp: pathlib.Path, mutate=None) -> pathlib.Path:
    src = tmp / "src"
    src.mkdir()
    for name in ("mpm_oracle.c", "mpm_oracle.h", "mpm_adjoint.c", "mpm_adjoint.h"):
        shutil.copy(ORACLE / name, src / name)
    if mutate is not None:
        old, new, which = mutate
        text = (src / "mpm_oracle.c").read_text()
        assert text.count(old) > which, f"mutation target not found: {old!r}"
        parts = text.split(old)
        text = old.join(parts[:which + 1]) + new + old.join(parts[which + 1:])
        (src / "mpm_oracle.c").write_text(text)
    lib = tmp / "libmutant.so"
    subprocess.run(["gcc", *CFLAGS, "-o", str(lib), str(src / "mpm_oracle.c"), str(src / "mpm_adjoint.c"), "-lm"],
                   check=True)
    return lib


def _pins_pass(lib: pathlib.Path) -> tuple[bool, str]:
    env = dict(os.environ, MPM_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", str(PINS)],
                       cwd=str(ROOT), env=env, capture_output=True, text=True, timeout=300)
    return r.returncode == 0, r.stdout[-2000:]


@pytest.fixture(scope="module")
def gcc():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")


def test_unmodified_copy_passes_the_pins(gcc, tmp_path):
    ok, out = _pins_pass(_build(tmp_path))
    assert ok, out


@pytest.mark.parametrize("name,old,new,which", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_fails_a_pin(gcc, tmp_path, name, old, new, which):
    ok, out = _pins_pass(_build(tmp_path, (old, new, which)))
    assert not ok, f"the pins did not catch: {name}\n{out}"
