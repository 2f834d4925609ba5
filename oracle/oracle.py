"""ctypes binding of the CPU oracle (oracle/liboracle.so built from sp3mm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path never loads this module.

Every function takes 0-based ``sp3mm4amg_amd.csr.CSR`` matrices; ``base=1``
runs the oracle on 1-based copies of the inputs and returns the 1-based result
shifted back to 0-based, so a 1-based run is checked against the same arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from sp3mm4amg_amd.csr import CSR

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _In(C.Structure):
    _fields_ = [("M", C.c_int64), ("N", C.c_int64), ("NZ", C.c_int64),
                ("irp", C.c_void_p), ("ja", C.c_void_p), ("as_", C.c_void_p)]


class _Out(C.Structure):
    _fields_ = [("M", C.c_int64), ("N", C.c_int64), ("NZ", C.c_int64),
                ("irp", C.POINTER(C.c_int64)), ("ja", C.POINTER(C.c_int64)),
                ("as_", C.POINTER(C.c_double))]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER
        for name in ("orc_ub", "orc_symb"):
            getattr(L, name).argtypes = [P(_In), P(_In), C.c_int, C.c_void_p]
        L.orc_ub_colparts.argtypes = [P(_In), P(_In), C.c_int, C.c_int, C.c_void_p]
        L.orc_symb3.argtypes = [P(_In), P(_In), P(_In), C.c_int, C.c_void_p]
        L.orc_spmm.argtypes = [P(_In), P(_In), C.c_int, C.c_int, P(_Out)]
        for name in ("orc_sp3mm_pair", "orc_sp3mm_merged"):
            getattr(L, name).argtypes = [P(_In), P(_In), P(_In), C.c_int, C.c_int, P(_Out)]
        L.orc_sp3mm_pair_ld.argtypes = [P(_In), P(_In), P(_In), C.c_int, P(_Out)]
        L.orc_free.argtypes = [P(_Out)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_wtime.restype = C.c_double
        _LIB = L
    return _LIB


class _Held:
    """Keeps the shifted int64/float64 arrays alive while C reads them."""

    def __init__(self, m: CSR, base: int):
        self.irp = np.ascontiguousarray(m.irp + base, dtype=np.int64)
        self.ja = np.ascontiguousarray(m.ja + base, dtype=np.int64)
        self.vals = np.ascontiguousarray(m.vals, dtype=np.float64)
        self.s = _In(m.M, m.N, m.nnz, self.irp.ctypes.data, self.ja.ctypes.data,
                     self.vals.ctypes.data)


def _take(out: _Out, base: int) -> CSR:
    M, nz = out.M, out.NZ
    irp = np.ctypeslib.as_array(out.irp, shape=(M + 1,)).copy() - base
    ja = np.ctypeslib.as_array(out.ja, shape=(max(nz, 1),))[:nz].copy() - base
    vals = np.ctypeslib.as_array(out.as_, shape=(max(nz, 1),))[:nz].copy()
    lib().orc_free(C.byref(out))
    return CSR(M, out.N, irp, ja, vals)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def wtime() -> float:
    return lib().orc_wtime()


def ub(A: CSR, B: CSR, base: int = 0) -> np.ndarray:
    a, b = _Held(A, base), _Held(B, base)
    out = np.zeros(A.M + 1, dtype=np.int64)
    lib().orc_ub(C.byref(a.s), C.byref(b.s), base, out.ctypes.data)
    return out


def ub_colparts(A: CSR, B: CSR, grid_cols: int, base: int = 0) -> np.ndarray:
    a, b = _Held(A, base), _Held(B, base)
    out = np.zeros(A.M * grid_cols + 1, dtype=np.int64)
    rc = lib().orc_ub_colparts(C.byref(a.s), C.byref(b.s), base, grid_cols, out.ctypes.data)
    if rc:
        raise RuntimeError("orc_ub_colparts failed")
    return out


def symb(A: CSR, B: CSR, base: int = 0) -> np.ndarray:
    a, b = _Held(A, base), _Held(B, base)
    out = np.zeros(A.M + 1, dtype=np.int64)
    if lib().orc_symb(C.byref(a.s), C.byref(b.s), base, out.ctypes.data):
        raise RuntimeError("orc_symb failed")
    return out


def symb3(R: CSR, A: CSR, P: CSR, base: int = 0) -> np.ndarray:
    r, a, p = _Held(R, base), _Held(A, base), _Held(P, base)
    out = np.zeros(R.M + 1, dtype=np.int64)
    if lib().orc_symb3(C.byref(r.s), C.byref(a.s), C.byref(p.s), base, out.ctypes.data):
        raise RuntimeError("orc_symb3 failed")
    return out


def spmm(A: CSR, B: CSR, base: int = 0, faithful: bool = False) -> CSR:
    a, b = _Held(A, base), _Held(B, base)
    out = _Out()
    rc = lib().orc_spmm(C.byref(a.s), C.byref(b.s), base, int(faithful), C.byref(out))
    if rc:
        lib().orc_free(C.byref(out))
        raise RuntimeError(f"orc_spmm failed ({rc})")
    return _take(out, base)


def _three(fn, R, A, P, base, *extra):
    r, a, p = _Held(R, base), _Held(A, base), _Held(P, base)
    out = _Out()
    rc = fn(C.byref(r.s), C.byref(a.s), C.byref(p.s), base, *extra, C.byref(out))
    if rc:
        lib().orc_free(C.byref(out))
        raise RuntimeError(f"oracle triple product failed ({rc})")
    return _take(out, base)


def sp3mm_pair(R: CSR, A: CSR, P: CSR, base: int = 0, faithful: bool = False) -> CSR:
    """(R*A)*P in the reference's pair order (code/src/Sp3MM_CSR_OMP_UB_Generic.c:456-494)."""
    return _three(lib().orc_sp3mm_pair, R, A, P, base, int(faithful))


def sp3mm_merged(R: CSR, A: CSR, P: CSR, base: int = 0, faithful: bool = False) -> CSR:
    """Direct row forwarding (code/src/Sp3MM_CSR_OMP_UB_Generic.c:498-573)."""
    return _three(lib().orc_sp3mm_merged, R, A, P, base, int(faithful))


def sp3mm_pair_ld(R: CSR, A: CSR, P: CSR, base: int = 0) -> CSR:
    """Pair product accumulated in x87 long double, rounded once to double."""
    return _three(lib().orc_sp3mm_pair_ld, R, A, P, base)


def abs_product(R: CSR, A: CSR, P: CSR) -> CSR:
    """|R| |A| |P| on the pattern of R*A*P: the magnitude scale of every entry."""
    return sp3mm_pair(R.with_vals(np.abs(R.vals)), A.with_vals(np.abs(A.vals)),
                      P.with_vals(np.abs(P.vals)))


def rel_errors(got: CSR, ref: CSR, scale: CSR | None = None) -> np.ndarray:
    """Per-nonzero relative error |got - ref| / max(|ref|, eps * scale).

    `scale` (same pattern, e.g. abs_product) guards entries that cancel to
    near zero; without it the denominator is |ref| (and 1 for exact zeros)."""
    if not got.pattern_equal(ref):
        raise AssertionError("patterns differ")
    diff = np.abs(got.vals.astype(np.float64) - ref.vals)
    den = np.abs(ref.vals)
    if scale is not None:
        if not scale.pattern_equal(ref):
            raise AssertionError("scale pattern differs")
        den = np.maximum(den, np.finfo(np.float64).eps * scale.vals)
    den = np.where(den == 0, 1.0, den)
    return diff / den


def error_bound(R: CSR, A: CSR, P: CSR, want: CSR, rel: float = 1e-12) -> np.ndarray:
    """Per-nonzero tolerance for comparing two FP64 evaluations of R*A*P that
    sum the same products in different orders:

        |x - y| <= rel * |y| + 4 * T_r * u * (|R||A||P|)_ij

    u = 2^-53 and T_r = products of row r in both stages (ub of R*A plus ub of
    (RA)*P).  The second term is the first-order bound of recursive summation,
    which any order (the reference's own included) satisfies against the exact
    value; it only matters for entries that cancel far below their scale."""
    scale = abs_product(R, A, P)
    if not scale.pattern_equal(want):
        raise AssertionError("scale pattern differs")
    ub1 = ub(R, A)[:-1]
    ub2 = ub(spmm(R, A), P)[:-1]
    T = (ub1 + ub2)[want.row_ids()].astype(np.float64)
    return rel * np.abs(want.vals) + 4.0 * T * 2.0 ** -53 * scale.vals
