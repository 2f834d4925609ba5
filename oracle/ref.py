"""ctypes binding of the reference library itself (oracle/_ref/libsp3mm_ref.so).

TEST INFRASTRUCTURE ONLY.  The shared object is compiled by ``make -C oracle
ref`` from the reference's C sources where they lie (gcc -O2 -fopenmp, as the
reference's own test Makefile does), is never committed, and travels to the GPU
host as a built file.  It is used to pin the oracle's outputs and as the timed
CPU baseline ("kind": "reference").

Types follow the reference headers: ``spmat`` (code/src/include/sparseMatrix.h:41-53,
64-bit ``idx_t`` from code/src/include/config.h:51) and ``CONFIG``
(code/src/include/config.h:42-49).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from sp3mm4amg_amd.csr import CSR

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libsp3mm_ref.so")
_LIB = None

RBTREE, IDXMAP = 0, 1


class Spmat(C.Structure):
    _fields_ = [("NZ", C.c_ulong), ("M", C.c_ulong), ("N", C.c_ulong),
                ("AS", C.POINTER(C.c_double)), ("JA", C.POINTER(C.c_ulong)),
                ("IRP", C.POINTER(C.c_ulong)), ("MAX_ROW_NZ", C.c_ulong)]


CHUNKS = C.CFUNCTYPE(None, C.c_ulong, C.c_void_p, C.c_void_p)


class Config(C.Structure):
    _fields_ = [("gridRows", C.c_ushort), ("gridCols", C.c_ushort),
                ("symbMMRowImplID", C.c_int), ("threadNum", C.c_uint),
                ("chunkDistrbFunc", C.c_void_p)]


def _noop(rows, mat, cfg):  # CHUNKS_DISTR_INTERF that leaves OMP_SCHEDULE alone
    return None


_NOOP = CHUNKS(_noop)

SPMM = C.CFUNCTYPE(C.POINTER(Spmat), C.POINTER(Spmat), C.POINTER(Spmat), C.POINTER(Config))


def available() -> bool:
    return os.path.exists(PATH)


def lib():
    global _LIB
    if _LIB is None:
        if not available():
            raise RuntimeError("oracle/_ref/libsp3mm_ref.so is not built (make -C oracle ref)")
        L = C.CDLL(PATH)
        P = C.POINTER
        for b in (0, 1):
            for n in ("sp3mmRowByRowPair_", "sp3mmRowByRowMerged_"):
                f = getattr(L, f"{n}{b}")
                f.argtypes = [P(Spmat), P(Spmat), P(Spmat), P(Config), C.c_void_p]
                f.restype = P(Spmat)
            for n in ("spmmSerial_", "spmmRowByRow_", "spmmRowByRow1DBlocks_",
                      "spmmRowByRow2DBlocks_", "spmmRowByRow2DBlocksAllocated_",
                      "spmmRowByRow_SymbNum_", "spmmRowByRow1DBlocks_SymbNum_"):
                f = getattr(L, f"{n}{b}")
                f.argtypes = [P(Spmat), P(Spmat), P(Config)]
                f.restype = P(Spmat)
            f = getattr(L, f"spMMSizeUpperbound_{b}")
            f.argtypes = [P(Spmat), P(Spmat)]
            f.restype = P(C.c_ulong)
            f = getattr(L, f"SpMM_Symb___{b}")
            f.argtypes = [C.c_int, P(Spmat), P(Spmat)]
            f.restype = P(C.c_ulong)
        _LIB = L
    return _LIB


class Held:
    """A reference ``spmat`` whose arrays are owned by numpy (never freed by C)."""

    def __init__(self, m: CSR, base: int = 0):
        self.irp = np.ascontiguousarray(m.irp + base, dtype=np.uint64)
        self.ja = np.ascontiguousarray(m.ja + base, dtype=np.uint64)
        self.vals = np.ascontiguousarray(m.vals, dtype=np.float64)
        self.s = Spmat(m.nnz, m.M, m.N,
                       self.vals.ctypes.data_as(C.POINTER(C.c_double)),
                       self.ja.ctypes.data_as(C.POINTER(C.c_ulong)),
                       self.irp.ctypes.data_as(C.POINTER(C.c_ulong)), 0)


_gomp = None


def set_threads(n: int) -> None:
    """The reference sizes its per-thread accumulators by CONFIG.threadNum and
    indexes them by omp_get_thread_num() (code/test/Sp3MM_test.c:296-297), so
    the OpenMP team size must equal threadNum."""
    global _gomp
    if _gomp is None:
        _gomp = C.CDLL("libgomp.so.1")
        _gomp.omp_set_num_threads.argtypes = [C.c_int]
    _gomp.omp_set_num_threads(int(n))


def chunks_fair_folded():
    """Address of the reference's chunksFairFolded (code/src/include/ompChunksDivide.h:77),
    the chunk distribution its test driver installs for a non-static OpenMP
    schedule (code/test/Sp3MM_test.c:80, :304-308); exported by oracle/ref_sched.c.
    None if the reference build predates that export."""
    L = lib()
    try:
        f = L.sp3mm_ref_chunks_fair_folded
    except AttributeError:
        return None
    f.restype = C.c_void_p
    return f()


def config(threads: int, grid_rows: int = 0, grid_cols: int = 0, mode: int = IDXMAP,
           fair_folded: bool = False) -> Config:
    """CONFIG as the reference test prepares it (code/test/Sp3MM_test.c:82-86,
    :175-207 grid adaptation: 1D blocks use 2*threads row blocks).  With
    fair_folded the chunk distribution function is the driver's
    chunksFairFolded (else one that leaves the OpenMP schedule alone)."""
    lib()
    set_threads(threads)
    gr = grid_rows or max(1, 2 * threads)
    gc = grid_cols or 1
    fn = chunks_fair_folded() if fair_folded else None
    return Config(gr, gc, mode, threads, fn or C.cast(_NOOP, C.c_void_p).value)


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def take(p, base: int = 0) -> CSR:
    """Copy a returned spmat into a 0-based CSR and release it with free()."""
    if not p:
        raise RuntimeError("reference call returned NULL")
    s = p.contents
    M, nz = int(s.M), int(s.NZ)
    irp = np.ctypeslib.as_array(s.IRP, shape=(M + 1,)).astype(np.int64) - base
    ja = (np.ctypeslib.as_array(s.JA, shape=(max(nz, 1),))[:nz].astype(np.int64) - base
          if nz else np.zeros(0, np.int64))
    vals = np.ctypeslib.as_array(s.AS, shape=(max(nz, 1),))[:nz].copy() if nz else np.zeros(0)
    out = CSR(M, int(s.N), irp, ja, vals)
    for a in (s.AS, s.JA, s.IRP):
        _libc.free(C.cast(a, C.c_void_p))
    _libc.free(C.cast(p, C.c_void_p))
    return out


def take_idx(p, n: int) -> np.ndarray:
    if not p:
        raise RuntimeError("reference call returned NULL")
    out = np.ctypeslib.as_array(p, shape=(n,)).astype(np.int64).copy()
    _libc.free(C.cast(p, C.c_void_p))
    return out


def spmm_fn(name: str, base: int = 0):
    return C.cast(getattr(lib(), f"{name}{base}"), C.c_void_p).value


def sp3mm_pair(R, A, P, spmm: str = "spmmRowByRow_SymbNum_", base: int = 0, threads: int = 1,
               mode: int = IDXMAP, grid=(0, 0), fair_folded: bool = False) -> CSR:
    r, a, p = Held(R, base), Held(A, base), Held(P, base)
    cfg = config(threads, grid[0], grid[1], mode, fair_folded)
    f = getattr(lib(), f"sp3mmRowByRowPair_{base}")
    out = f(C.byref(r.s), C.byref(a.s), C.byref(p.s), C.byref(cfg), spmm_fn(spmm, base))
    return take(out, base)


def sp3mm_merged(R, A, P, base: int = 0, threads: int = 1, fair_folded: bool = False) -> CSR:
    r, a, p = Held(R, base), Held(A, base), Held(P, base)
    cfg = config(threads, fair_folded=fair_folded)
    f = getattr(lib(), f"sp3mmRowByRowMerged_{base}")
    return take(f(C.byref(r.s), C.byref(a.s), C.byref(p.s), C.byref(cfg), None), base)


def spmm(A, B, name: str = "spmmRowByRow_SymbNum_", base: int = 0, threads: int = 1,
         mode: int = IDXMAP, grid=(0, 0)) -> CSR:
    a, b = Held(A, base), Held(B, base)
    cfg = config(threads, grid[0], grid[1], mode)
    return take(getattr(lib(), f"{name}{base}")(C.byref(a.s), C.byref(b.s), C.byref(cfg)), base)


def ub(A, B, base: int = 0) -> np.ndarray:
    a, b = Held(A, base), Held(B, base)
    return take_idx(getattr(lib(), f"spMMSizeUpperbound_{base}")(C.byref(a.s), C.byref(b.s)),
                    A.M + 1)


def symb(A, B, base: int = 0, mode: int = IDXMAP) -> np.ndarray:
    a, b = Held(A, base), Held(B, base)
    return take_idx(getattr(lib(), f"SpMM_Symb___{base}")(mode, C.byref(a.s), C.byref(b.s)),
                    A.M + 1)


def elapsed_internal() -> float:
    return C.c_double.in_dll(lib(), "ElapsedInternal").value
