"""A compiled C caller of the drop-in entry points (tests/c/rap_caller.c),
written like the reference's test driver: it includes include/sp3mm_amd.h,
links with -lsp3mm_amd and checks the pair, merged and 1-based products of a
2D Laplacian with 2x2 aggregates against their closed form (aggregate sums).

CPU: the caller compiles against the header and links against the library.
GPU: it runs, as a child process, and exits 0."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "sp3mm4amg_amd")
SRC = os.path.join(ROOT, "tests", "c", "rap_caller.c")


def _build(out: str) -> None:
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lsp3mm_amd", f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=120)


needs_build = pytest.mark.skipif(not shutil.which("gcc") or not os.path.exists(os.path.join(LIBDIR, "libsp3mm_amd.so")),
                                 reason="gcc or the built library is missing")


@needs_build
def test_c_caller_compiles_and_links(tmp_path):
    _build(str(tmp_path / "rap_caller"))
    assert os.access(tmp_path / "rap_caller", os.X_OK)


@needs_build
@pytest.mark.gpu
def test_c_caller_runs(tmp_path):
    exe = str(tmp_path / "rap_caller")
    _build(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "rap_caller ok" in r.stdout
