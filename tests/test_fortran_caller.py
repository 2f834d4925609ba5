"""A Fortran caller of the 1-based entry points through ISO_C_BINDING
(tests/fortran/sp3mm_amd.f90: the binding module; tests/fortran/rap_caller.f90:
the caller), the way an AMG4PSBLAS-style code passes 1-based CSR
(code/src/Fortran_C_indexing.md).  The caller checks the pair form (with
spmmRowByRow_1, with spmmRowByRow_SymbNum_1 and with no SpMM) and the merged
form against the closed-form Galerkin operator of a 2D Laplacian with 2x2
aggregates.

CPU: the module and the caller compile with ROCm's flang and link with
-lsp3mm_amd.  GPU: the caller runs and prints "fortran caller ok"."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "sp3mm4amg_amd")
SRC = [os.path.join(ROOT, "tests", "fortran", f) for f in ("sp3mm_amd.f90", "rap_caller.f90")]


def _flang():
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang"), shutil.which("gfortran")):
        if c and os.access(c, os.X_OK):
            return c
    return None


needs_build = pytest.mark.skipif(_flang() is None or not os.path.exists(os.path.join(LIBDIR, "libsp3mm_amd.so")),
                                 reason="no Fortran compiler, or the built library is missing")


def _build(tmp_path) -> str:
    fc = _flang()
    out = str(tmp_path / "rap_caller_f")
    mod = "-J" if os.path.basename(fc).startswith("gfortran") else "-module-dir"
    cmd = [fc, "-O2", mod, str(tmp_path)] + SRC + ["-L", LIBDIR, "-lsp3mm_amd", f"-Wl,-rpath,{LIBDIR}", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return out


@needs_build
def test_fortran_caller_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.access(exe, os.X_OK)


@needs_build
@pytest.mark.gpu
def test_fortran_caller_runs(tmp_path):
    exe = _build(tmp_path)
    print("running the Fortran caller", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran caller ok" in r.stdout
    assert r.stdout.count("ok: ") == 4
