"""The reference's own test driver on the GPU drop-in.

oracle/Makefile target `harness` compiles code/test/Sp3MM_test.c unchanged,
with the reference's utility units (Matrix Market reader, result check,
OpenMP helpers) and none of its Sp3MM units, and links it with
-lsp3mm_amd: every Sp3MM entry point the driver calls (the four UB SpMMs in
the pair product, the merged product, the two symbolic-numeric SpMMs in both
symbolic modes) runs on the GPU.  The MOCK_FORTRAN_INDEXING build of the same
driver shifts its matrices to 1-based and drives the _1 set.  The binaries
live in oracle/_ref/ (built where the reference sources exist); the tests skip
without them.

Inputs are the reference's known-answer group code/test/{r,ac,p,ac_next},
written back to Matrix Market from tests/golden/fixture_2744.npz.  The driver
compares every product with ac_next (its own check, spmatDiff) and exits 0
only if all of them match.
"""
import os
import subprocess

import numpy as np
import pytest

from sp3mm4amg_amd import mmio
from sp3mm4amg_amd.csr import CSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
GOLDEN = os.path.join(ROOT, "tests", "golden", "fixture_2744.npz")
# products the driver times: 4 UB SpMMs in the pair form, the merged form,
# 2 symbolic-numeric SpMMs in the pair form for each of the 2 symbolic modes
PRODUCTS = 4 + 1 + 2 + 2


def _binary(name):
    path = os.path.join(REF, name)
    if not os.access(path, os.X_OK):
        pytest.skip(f"oracle/_ref/{name} is not built (make -C oracle harness)")
    return path


def _write_group(d):
    z = np.load(GOLDEN)
    paths = []
    for k in ("r", "ac", "p", "ac_next"):
        M, N = (int(x) for x in z[k + "_shape"])
        m = CSR(M, N, z[k + "_irp"], z[k + "_ja"].astype(np.int64), z[k + "_vals"])
        p = os.path.join(d, k)
        mmio.write(p, m)
        paths.append(p)
    return paths


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sp3mm_test_amd", "sp3mm_test_amd_fortran"])
def test_reference_driver_on_dropin(tmp_path, name):
    exe = _binary(name)
    args = _write_group(str(tmp_path))
    env = dict(os.environ, OMP_NUM_THREADS="4", TMPDIR=str(tmp_path))
    print(f"running {name} on the known-answer group", flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=180, env=env, cwd=str(tmp_path))
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("timeAvg:") == PRODUCTS
    assert "COARSENING AC: 2744x2744 ---> 343x343" in r.stdout
