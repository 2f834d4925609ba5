"""CPU checks of the workload generators and of the C ABI library (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from sp3mm4amg_amd import problems
from sp3mm4amg_amd.csr import CSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sp3mm_amd.h")
LIB = os.path.join(ROOT, "sp3mm4amg_amd", "libsp3mm_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # preprocessor lines (macro values such as SP3MM_NOALLOC_SCRATCH_BYTES) declare no symbols
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text))
    names -= {"if", "while", "for", "return", "sizeof"}
    # function-pointer typedef names are not symbols
    names -= {"SPMM_INTERF", "SP3MM_INTERF"}
    return sorted(n for n in names if not n.startswith("__"))


def test_header_declares_the_reference_entry_points():
    names = set(declared_functions())
    for base in (0, 1):
        for n in ("sp3mmRowByRowPair_", "sp3mmRowByRowMerged_", "spmmSerial_", "spmmRowByRow_",
                  "spmmRowByRow1DBlocks_", "spmmRowByRow2DBlocks_",
                  "spmmRowByRow2DBlocksAllocated_", "spmmRowByRow_SymbNum_",
                  "spmmRowByRow1DBlocks_SymbNum_", "spmmRowByRow2DBlocks_SymbNum_",
                  "spMMSizeUpperbound_", "spMMSizeUpperboundColParts_", "SpMM_Symb___",
                  "SpMM_Symb_OutIdxs__", "Sp3MM_Symb_OutIdxs_", "SpMM_Symb__ColParts_",
                  "SpMM_Symb_OutIdxs_ColParts_"):
            assert f"{n}{base}" in names
    for n in ("allocSpMatrix", "freeSpmat", "freeSpmatInternal", "C_FortranShiftIdxs",
              "Fortran_C_ShiftIdxs"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    for g in ("Start", "End", "Elapsed", "ElapsedInternal"):
        ctypes.c_double.in_dll(lib, g)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_struct_layouts_match_reference():
    from sp3mm4amg_amd import lib as L
    # spmat: idx_t NZ,M,N; double* AS; idx_t* JA; idx_t* IRP; idx_t MAX_ROW_NZ
    assert ctypes.sizeof(L.Spmat) == 7 * 8
    assert [f for f, _ in L.Spmat._fields_] == ["NZ", "M", "N", "AS", "JA", "IRP", "MAX_ROW_NZ"]
    # CONFIG: ushort, ushort, enum, uint, void*
    assert ctypes.sizeof(L.Config) == 24
    assert L.Config.chunkDistrbFunc.offset == 16
    # sp3mm_tuning (this library's own): eight int32 fields, two uint32 hash multipliers
    assert ctypes.sizeof(L.Tuning) == 40
    assert L.Tuning.hash1.offset == 32 and L.Tuning.hash2.offset == 36


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_host_helpers_without_gpu():
    from sp3mm4amg_amd import lib as L
    lib = L.lib()
    m = lib.allocSpMatrix(5, 7)
    assert m and m.contents.M == 5 and m.contents.N == 7
    irp = np.ctypeslib.as_array(m.contents.IRP, shape=(6,))
    assert not irp.any()
    lib.C_FortranShiftIdxs(m)
    assert (irp == 1).all()
    lib.Fortran_C_ShiftIdxs(m)
    assert not irp.any()
    lib.freeSpmat(m)


def test_poisson2d64_sizes():
    p = problems.poisson2d_sa(64)
    R, A, P = p["R"], p["A"], p["P"]
    for m in (R, A, P):
        m.check()
    assert (R.M, R.N, A.M, P.N) == (1024, 4096, 4096, 1024)
    assert (R.nnz, A.nnz, P.nnz) == (12032, 20224, 12032)
    assert O.ub(R, A)[-1] == 59656
    RA = O.spmm(R, A)
    assert RA.nnz == 23812 and O.ub(RA, P)[-1] == 70428
    C = O.sp3mm_pair(R, A, P)
    assert C.nnz == 12676 and np.diff(C.irp).min() == 6 and np.diff(C.irp).max() == 13


def test_prolongator_properties():
    """Interior rows of A have zero row sum, so P = (I - w D^-1 A) P0 keeps the
    constant vector there: (P 1)_i = 1; R = P^T exactly."""
    p = problems.poisson3d_sa(16)
    R, A, P = p["R"], p["A"], p["P"]
    rs = np.add.reduceat(A.vals, A.irp[:-1])
    interior = np.isclose(rs, 0)
    prs = np.add.reduceat(P.vals, P.irp[:-1])
    np.testing.assert_allclose(prs[interior], 1.0, rtol=0, atol=1e-14)
    RT = P.transpose()
    assert RT.pattern_equal(R) and np.array_equal(RT.vals, R.vals)
    # row i of P touches at most 4 aggregates (its own and 3 face neighbours)
    assert np.diff(P.irp).max() == 4


@pytest.mark.slow
def test_poisson3d128_sizes():
    p = problems.poisson3d_sa(128)
    R, A, P = p["R"], p["A"], p["P"]
    assert (R.nnz, A.nnz, P.nnz) == (8290304, 14581760, 8290304)
    assert (R.M, P.N) == (262144, 262144)
    assert O.ub(R, A)[-1] == 57740288
    C = O.sp3mm_merged(R, A, P)
    assert C.nnz == 8382712


def test_stencil27_symmetry_and_diagonal():
    p = problems.stencil27_sa(12, seed=3)
    A = p["A"]
    AT = A.transpose()
    assert AT.pattern_equal(A) and np.array_equal(AT.vals, A.vals)
    rows = A.row_ids()
    d = A.vals[A.ja == rows]
    off = np.abs(A.vals).copy()
    off[A.ja == rows] = 0
    np.testing.assert_allclose(d, 1.001 * np.add.reduceat(off, A.irp[:-1]), rtol=1e-15)


def test_irregular_generator_shape():
    p = problems.irregular_sa(4000, seed=4)
    R, A, P = p["R"], p["A"], p["P"]
    for m in (R, A, P):
        m.check()
    rows = A.row_ids()
    assert (A.ja == rows).sum() == A.M  # every row has its diagonal
    lens = np.diff(A.irp) - 1
    assert lens.min() >= 1 and lens.max() <= 4096
    # aggregates of 4..12 rows (the last one may be smaller)
    assert 4000 // 12 <= P.N <= 4000 // 4 + 1


def test_batch_generator():
    probs = problems.poisson2d_batch(8, seed=5)
    sides = [p["side"] for p in probs]
    assert all(16 <= s <= 256 for s in sides)
    for p in probs:
        assert p["R"].N == p["A"].M and p["A"].N == p["P"].M
