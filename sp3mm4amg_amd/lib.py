"""Python mirror of the Sp3MM interface over the C ABI of libsp3mm_amd.so.

Two layers, as in include/sp3mm_amd.h:

* the reference-compatible entry points (``sp3mmRowByRowPair_0`` ...), called
  with host ``spmat`` structs built from ``CSR`` arrays -- the same calls a C or
  Fortran caller of Sp3MM4AMG makes (code/test/Sp3MM_test.c:121,254);
* the device-resident interface (``DeviceCSR``, ``rap``, ``spgemm``) used by
  the benchmark and by callers that keep matrices in HBM.

There is no fallback: every function raises ``RuntimeError`` when the HIP
library is missing or the GPU call fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .csr import CSR

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsp3mm_amd.so")
_LIB = None

F64, F32 = 0, 1
PAIR, DIRECT = 0, 1
FP64, MIXED, MIXED64 = 0, 1, 2
RBTREE, IDXMAP = 0, 1


class Spmat(C.Structure):
    """code/src/include/sparseMatrix.h:41-53"""
    _fields_ = [("NZ", C.c_ulong), ("M", C.c_ulong), ("N", C.c_ulong),
                ("AS", C.POINTER(C.c_double)), ("JA", C.POINTER(C.c_ulong)),
                ("IRP", C.POINTER(C.c_ulong)), ("MAX_ROW_NZ", C.c_ulong)]


class Config(C.Structure):
    """code/src/include/config.h:42-49"""
    _fields_ = [("gridRows", C.c_ushort), ("gridCols", C.c_ushort),
                ("symbMMRowImplID", C.c_int), ("threadNum", C.c_uint),
                ("chunkDistrbFunc", C.c_void_p)]


class DMat(C.Structure):
    _fields_ = [("M", C.c_int64), ("N", C.c_int64), ("nnz", C.c_int64),
                ("rp", C.c_void_p), ("rp32", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p), ("vtype", C.c_int32), ("owned", C.c_int32)]


class Opts(C.Structure):
    _fields_ = [("path", C.c_int32), ("precision", C.c_int32), ("guard", C.c_double),
                ("timing", C.c_int32), ("reserved", C.c_int32), ("rowsum", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_ub", C.c_double), ("ms_symbolic", C.c_double),
                ("ms_scan", C.c_double), ("ms_numeric", C.c_double), ("ms_main", C.c_double),
                ("mac1", C.c_int64), ("mac2", C.c_int64), ("nnz_mid", C.c_int64),
                ("nnz_out", C.c_int64), ("guarded_rows", C.c_int64),
                ("guarded_entries", C.c_int64), ("main_kernel_launches", C.c_int32),
                ("spilled_rows", C.c_int32), ("reassociated", C.c_int32), ("reserved2", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class Tuning(C.Structure):
    _fields_ = [("u1", C.c_int32), ("u2", C.c_int32), ("subwave_h", C.c_int32),
                ("slog1_cap", C.c_int32), ("slog2_cap", C.c_int32), ("h1_shift", C.c_int32),
                ("h2_div", C.c_int32), ("xcd_map", C.c_int32), ("hash1", C.c_uint32),
                ("hash2", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SPMM_NAMES = ["spmmSerial_", "spmmRowByRow_", "spmmRowByRow1DBlocks_", "spmmRowByRow2DBlocks_",
              "spmmRowByRow2DBlocksAllocated_", "spmmRowByRow_SymbNum_",
              "spmmRowByRow1DBlocks_SymbNum_", "spmmRowByRow2DBlocks_SymbNum_"]


def lib():
    """Load libsp3mm_amd.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run python -m sp3mm4amg_amd.build")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        for b in (0, 1):
            for n in ("sp3mmRowByRowPair_", "sp3mmRowByRowMerged_"):
                f = getattr(L, f"{n}{b}")
                f.argtypes = [P(Spmat), P(Spmat), P(Spmat), P(Config), C.c_void_p]
                f.restype = P(Spmat)
            for n in SPMM_NAMES:
                f = getattr(L, f"{n}{b}")
                f.argtypes = [P(Spmat), P(Spmat), P(Config)]
                f.restype = P(Spmat)
            f = getattr(L, f"spMMSizeUpperbound_{b}")
            f.argtypes = [P(Spmat), P(Spmat)]
            f.restype = P(C.c_ulong)
            f = getattr(L, f"spMMSizeUpperboundColParts_{b}")
            f.argtypes = [P(Spmat), P(Spmat), C.c_ushort, P(C.c_ulong)]
            f.restype = P(C.c_ulong)
            f = getattr(L, f"SpMM_Symb___{b}")
            f.argtypes = [C.c_int, P(Spmat), P(Spmat)]
            f.restype = P(C.c_ulong)
            f = getattr(L, f"SpMM_Symb_OutIdxs__{b}")
            f.argtypes = [C.c_int, P(Spmat), P(Spmat), P(P(P(C.c_ulong)))]
            f.restype = P(C.c_ulong)
            f = getattr(L, f"Sp3MM_Symb_OutIdxs_{b}")
            f.argtypes = [C.c_int, P(Spmat), P(Spmat), P(Spmat), P(P(P(C.c_ulong)))]
            f.restype = P(C.c_ulong)
            f = getattr(L, f"SpMM_Symb__ColParts_{b}")
            f.argtypes = [C.c_int, P(Spmat), P(Spmat), C.c_ushort, P(P(C.c_ulong))]
            f.restype = P(C.c_ulong)
            f = getattr(L, f"SpMM_Symb_OutIdxs_ColParts_{b}")
            f.argtypes = [C.c_int, P(Spmat), P(Spmat), P(P(P(C.c_ulong))), C.c_ushort, P(P(C.c_ulong))]
            f.restype = P(C.c_ulong)
        L.freeSpmat.argtypes = [P(Spmat)]
        L.C_FortranShiftIdxs.argtypes = [P(Spmat)]
        L.Fortran_C_ShiftIdxs.argtypes = [P(Spmat)]
        L.allocSpMatrix.argtypes = [C.c_ulong, C.c_ulong]
        L.allocSpMatrix.restype = P(Spmat)
        L.sp3mm_init.argtypes = [C.c_int]
        L.sp3mm_last_error.restype = C.c_char_p
        L.sp3mm_dmat_upload.argtypes = [P(DMat), C.c_int64, C.c_int64, C.c_int64, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]
        L.sp3mm_dmat_download.argtypes = [P(DMat), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.sp3mm_dmat_convert.argtypes = [P(DMat), C.c_int32]
        L.sp3mm_dmat_import.argtypes = [P(DMat), C.c_int64, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int32, C.c_int32]
        L.sp3mm_dmat_export.argtypes = [P(DMat), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.sp3mm_dmat_free.argtypes = [P(DMat)]
        L.sp3mm_dmat_free.restype = None
        L.sp3mm_rap.argtypes = [P(DMat), P(DMat), P(DMat), P(DMat), P(Opts), P(Stats)]
        L.sp3mm_rap_noalloc.argtypes = [P(DMat), P(DMat), P(DMat), C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int64, P(C.c_int64), P(Stats)]
        L.sp3mm_spgemm.argtypes = [P(DMat), P(DMat), P(DMat), P(Stats)]
        L.sp3mm_ub.argtypes = [P(DMat), P(DMat), C.c_void_p]
        L.sp3mm_symbolic.argtypes = [P(DMat), P(DMat), P(DMat), C.c_void_p]
        L.sp3mm_batch_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, P(Opts), C.c_int32,
                                         P(C.c_void_p)]
        L.sp3mm_batch_run.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
        L.sp3mm_batch_repack.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.sp3mm_batch_result.argtypes = [C.c_void_p, C.c_int32, P(DMat)]
        L.sp3mm_batch_info.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_int64), P(C.c_int32)]
        L.sp3mm_batch_free.argtypes = [C.c_void_p]
        L.sp3mm_batch_autotune.argtypes = [C.c_void_p, C.c_int32, P(Tuning), P(C.c_double), P(C.c_double)]
        L.sp3mm_batch_free.restype = None
        L.sp3mm_transpose.argtypes = [P(DMat), P(DMat)]
        L.sp3mm_sa_prolongator.argtypes = [P(DMat), C.c_void_p, C.c_int64, C.c_double, C.c_int32, P(DMat)]
        L.sp3mm_get_tuning.argtypes = [P(Tuning)]
        L.sp3mm_set_tuning.argtypes = [P(Tuning)]
        L.sp3mm_autotune.argtypes = [P(DMat), P(DMat), P(DMat), P(Opts), C.c_int32, P(Tuning),
                                     P(C.c_double), P(C.c_double)]
        L.sp3mm_device_alloc.argtypes = [C.c_int64]
        L.sp3mm_device_alloc.restype = C.c_void_p
        L.sp3mm_device_free.argtypes = [C.c_void_p]
        L.sp3mm_device_free.restype = None
        L.sp3mm_copy_to_host.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.sp3mm_copy_to_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.sp3mm_spmv.argtypes = [P(DMat), C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        L.sp3mm_residual.argtypes = [P(DMat), C.c_void_p, C.c_void_p, C.c_void_p]
        L.sp3mm_jacobi.argtypes = [P(DMat), C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int32]
        L.sp3mm_axpby.argtypes = [C.c_int64, C.c_double, C.c_void_p, C.c_double, C.c_void_p]
        L.sp3mm_dot.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, P(C.c_double)]
        L.sp3mm_vcycle_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, P(DMat), C.c_int32,
                                          C.c_int32, C.c_double, C.c_int32, C.c_int32, P(C.c_void_p)]
        L.sp3mm_vcycle_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int32,
                                         P(C.c_int32), C.c_void_p, P(C.c_double)]
        L.sp3mm_vcycle_free.argtypes = [C.c_void_p]
        L.sp3mm_vcycle_free.restype = None
        L.sp3mm_hier_create.argtypes = [P(DMat), C.c_int32, C.c_void_p, C.c_void_p, C.c_double, C.c_int32,
                                        C.c_int32, P(C.c_void_p)]
        L.sp3mm_hier_update.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
        L.sp3mm_hier_level.argtypes = [C.c_void_p, C.c_int32, P(DMat), P(DMat), P(DMat)]
        L.sp3mm_hier_info.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_double), P(C.c_int32)]
        L.sp3mm_hier_free.argtypes = [C.c_void_p]
        L.sp3mm_hier_free.restype = None
        L.sp3mm_sync.restype = C.c_int
        L.sp3mm_set_stream.argtypes = [C.c_void_p]
        L.sp3mm_get_stream.restype = C.c_void_p
        L.sp3mm_copy_probe.argtypes = [C.c_int64, C.c_int32]
        L.sp3mm_copy_probe.restype = C.c_double
        L.sp3mm_hash_scores.argtypes = [P(DMat), P(DMat), P(DMat), C.c_int32, C.c_int32, C.c_void_p,
                                        C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.sp3mm_hash_scores.restype = C.c_int32
        _LIB = L
    return _LIB


def last_error() -> str:
    return lib().sp3mm_last_error().decode(errors="replace")


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: {last_error()}")


def init(device: int = 0) -> None:
    _check(lib().sp3mm_init(device), "sp3mm_init")


# ---------------------------------------------------------------------------
# device-resident interface
# ---------------------------------------------------------------------------
class DeviceCSR:
    """A sorted CSR matrix resident in HBM (owned by the library's pool)."""

    def __init__(self, d: DMat):
        self.d = d

    @classmethod
    def upload(cls, m: CSR, vtype: int = F64, base: int = 0) -> "DeviceCSR":
        irp = np.ascontiguousarray(m.irp + base, dtype=np.int64)
        ja = np.ascontiguousarray(m.ja + base, dtype=np.int64)
        vals = np.ascontiguousarray(m.vals, dtype=np.float32 if vtype == F32 else np.float64)
        d = DMat()
        _check(lib().sp3mm_dmat_upload(C.byref(d), m.M, m.N, m.nnz, irp.ctypes.data,
                                       ja.ctypes.data, vals.ctypes.data, vtype, base),
               "sp3mm_dmat_upload")
        return cls(d)

    @property
    def shape(self):
        return (self.d.M, self.d.N)

    @property
    def nnz(self) -> int:
        return self.d.nnz

    def download(self, base: int = 0) -> CSR:
        M, nz = self.d.M, self.d.nnz
        irp = np.empty(M + 1, dtype=np.int64)
        ja = np.empty(max(nz, 1), dtype=np.int64)
        vals = np.empty(max(nz, 1), dtype=np.float32 if self.d.vtype == F32 else np.float64)
        _check(lib().sp3mm_dmat_download(C.byref(self.d), irp.ctypes.data, ja.ctypes.data,
                                         vals.ctypes.data, base), "sp3mm_dmat_download")
        return CSR(M, self.d.N, irp - base, ja[:nz] - base, vals[:nz])

    @classmethod
    def import_device(cls, M: int, N: int, nnz: int, irp: "DeviceArray", ja: "DeviceArray",
                      vals: "DeviceArray", vtype: int = F64, base: int = 0) -> "DeviceCSR":
        """A matrix whose int64 offsets/columns (index base `base`) and values
        are already in device memory; converted on the GPU."""
        d = DMat()
        _check(lib().sp3mm_dmat_import(C.byref(d), M, N, nnz, irp.ptr, ja.ptr if ja else None,
                                       vals.ptr if vals else None, vtype, base), "sp3mm_dmat_import")
        return cls(d)

    def export_device(self, irp: "DeviceArray", ja: "DeviceArray", vals: "DeviceArray", base: int = 0) -> None:
        """Writes int64 offsets/columns (index base `base`) and values into
        caller-owned device arrays of M+1, nnz and nnz entries."""
        if irp.n < self.d.M + 1 or (ja is not None and ja.n < self.d.nnz) or \
                (vals is not None and vals.n < self.d.nnz):
            raise ValueError("export: arrays too short")
        _check(lib().sp3mm_dmat_export(C.byref(self.d), irp.ptr, ja.ptr if ja is not None else None,
                                       vals.ptr if vals is not None else None, base), "sp3mm_dmat_export")

    def convert(self, vtype: int) -> None:
        _check(lib().sp3mm_dmat_convert(C.byref(self.d), vtype), "sp3mm_dmat_convert")

    def free(self) -> None:
        if self.d.owned:
            lib().sp3mm_dmat_free(C.byref(self.d))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rap(R: DeviceCSR, A: DeviceCSR, P: DeviceCSR, path: int = DIRECT, precision: int = FP64,
        guard: float = 0.0, timing: bool = False, rowsum_dev: int | None = None, raw_stats: bool = False):
    """R*A*P on device-resident operands; returns (DeviceCSR, stats dict), or
    the Stats structure itself with raw_stats (no dict built: timing loops)."""
    out = DMat()
    st = Stats()
    o = Opts(path, precision, guard, int(timing), 0, rowsum_dev)
    _check(lib().sp3mm_rap(C.byref(R.d), C.byref(A.d), C.byref(P.d), C.byref(out), C.byref(o),
                           C.byref(st)), "sp3mm_rap")
    return DeviceCSR(out), (st if raw_stats else st.as_dict())


def rap_noalloc(R: DeviceCSR, A: DeviceCSR, P: DeviceCSR, rp: "DeviceArray", col: "DeviceArray",
                val: "DeviceArray", cap: int):
    """R*A*P (FP64, direct path) into caller-owned device arrays (rp: M+1
    int64, col: int32, val: float64, `cap` entries each) with no device
    allocation inside the library.  Returns (nnz, stats dict); on failure
    raises RuntimeError, whose `needed` attribute holds the capacity a repeat
    needs when the output arrays were too short (else -1)."""
    nnz = C.c_int64(-1)
    st = Stats()
    rc = lib().sp3mm_rap_noalloc(C.byref(R.d), C.byref(A.d), C.byref(P.d), rp.ptr, col.ptr, val.ptr, cap,
                                 C.byref(nnz), C.byref(st))
    if rc != 0:
        e = RuntimeError(f"sp3mm_rap_noalloc failed: {last_error()}")
        e.needed = nnz.value
        raise e
    return nnz.value, st.as_dict()


def spgemm(A: DeviceCSR, B: DeviceCSR):
    out = DMat()
    st = Stats()
    _check(lib().sp3mm_spgemm(C.byref(A.d), C.byref(B.d), C.byref(out), C.byref(st)),
           "sp3mm_spgemm")
    return DeviceCSR(out), st.as_dict()


def ub(A: DeviceCSR, B: DeviceCSR) -> np.ndarray:
    out = np.zeros(A.d.M + 1, dtype=np.int64)
    _check(lib().sp3mm_ub(C.byref(A.d), C.byref(B.d), out.ctypes.data), "sp3mm_ub")
    return out


def symbolic(A: DeviceCSR, B: DeviceCSR, Cm: DeviceCSR | None = None) -> np.ndarray:
    out = np.zeros(A.d.M + 1, dtype=np.int64)
    _check(lib().sp3mm_symbolic(C.byref(A.d), C.byref(B.d), C.byref(Cm.d) if Cm else None,
                                out.ctypes.data), "sp3mm_symbolic")
    return out


class DeviceArray:
    """A raw device buffer from the library's pool (e.g. the FP64 row sums)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n, self.dtype = n, np.dtype(dtype)
        self.ptr = lib().sp3mm_device_alloc(self.n * self.dtype.itemsize)
        if not self.ptr:
            raise RuntimeError("sp3mm_device_alloc failed")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        _check(lib().sp3mm_copy_to_device(d.ptr, a.ctypes.data, a.nbytes), "sp3mm_copy_to_device")
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check(lib().sp3mm_copy_to_host(out.ctypes.data, self.ptr, out.nbytes), "sp3mm_copy_to_host")
        return out

    def free(self) -> None:
        if self.ptr:
            lib().sp3mm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many products R[i]*A[i]*P[i] packed into one block-diagonal product
    (optionally captured as one HIP graph)."""

    def __init__(self, Rs, As, Ps, mixed: bool = False, guard: float = 0.0, graph: bool = True,
                 precision: int | None = None):
        n = len(Rs)
        arr = DMat * n
        self._R, self._A, self._P = arr(*[m.d for m in Rs]), arr(*[m.d for m in As]), arr(*[m.d for m in Ps])
        if precision is None:
            precision = MIXED if mixed else FP64
        o = Opts(DIRECT, precision, guard, 0, 0, None)
        h = C.c_void_p()
        _check(lib().sp3mm_batch_create(n, self._R, self._A, self._P, C.byref(o), int(graph), C.byref(h)),
               "sp3mm_batch_create")
        self.h = h
        self.count = n

    def run(self, graph: bool = True) -> float:
        ms = C.c_double()
        _check(lib().sp3mm_batch_run(self.h, int(graph), C.byref(ms)), "sp3mm_batch_run")
        return ms.value

    def repack(self, Rs=None, As=None, Ps=None) -> None:
        """Copy the operands' current entries into the batch (by default the
        matrices given at creation, whose device arrays the caller may have
        rewritten); shapes and nonzero counts must be unchanged."""
        n = self.count
        arr = DMat * n
        R = arr(*[m.d for m in Rs]) if Rs is not None else self._R
        A = arr(*[m.d for m in As]) if As is not None else self._A
        P = arr(*[m.d for m in Ps]) if Ps is not None else self._P
        _check(lib().sp3mm_batch_repack(self.h, n, R, A, P), "sp3mm_batch_repack")

    def result(self, i: int) -> DeviceCSR:
        d = DMat()
        _check(lib().sp3mm_batch_result(self.h, i, C.byref(d)), "sp3mm_batch_result")
        return DeviceCSR(d)

    def autotune(self, reps: int = 3):
        """Tune the schedule on the packed product and rebuild the batch (and
        its graph) under it; returns (schedule, ms best, ms of the starting
        schedule)."""
        t = Tuning()
        mb, md = C.c_double(), C.c_double()
        _check(lib().sp3mm_batch_autotune(self.h, int(reps), C.byref(t), C.byref(mb), C.byref(md)),
               "sp3mm_batch_autotune")
        return t.as_dict(), mb.value, md.value

    def info(self):
        a, b, n, g = C.c_double(), C.c_double(), C.c_int64(), C.c_int32()
        lib().sp3mm_batch_info(self.h, C.byref(a), C.byref(b), C.byref(n), C.byref(g))
        return {"ms_pack": a.value, "ms_build": b.value, "nnz": n.value, "graph": bool(g.value)}

    def free(self) -> None:
        if self.h:
            lib().sp3mm_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rap_mixed_host(R: CSR, A: CSR, P: CSR, guard: float = 0.0, rowsum: bool = False,
                   precision: int = MIXED):
    """Mixed-precision direct product from host matrices (values rounded to FP32
    on upload); returns (CSR with FP64 values, stats, row sums or None)."""
    dR, dA, dP = (DeviceCSR.upload(m, vtype=F32) for m in (R, A, P))
    rs = DeviceArray(R.M) if rowsum else None
    out, st = rap(dR, dA, dP, path=DIRECT, precision=precision, guard=guard,
                  rowsum_dev=rs.ptr if rs else None)
    res = out.download()
    sums = rs.numpy() if rs else None
    for m in (dR, dA, dP, out):
        m.free()
    if rs:
        rs.free()
    return res, st, sums


def transpose(A: DeviceCSR) -> DeviceCSR:
    """A^T on the device (sorted CSR, same value type)."""
    d = DMat()
    _check(lib().sp3mm_transpose(C.byref(A.d), C.byref(d)), "sp3mm_transpose")
    return DeviceCSR(d)


def sa_prolongator(A: DeviceCSR, agg: np.ndarray, nagg: int, omega: float = 2.0 / 3.0,
                   diag_mode: int = 0) -> DeviceCSR:
    """P = (I - omega D^-1 A) P0 on the device; agg[i] is the aggregate of row i,
    D = diag(A) (diag_mode 0) or |diag(A)| + 1 (diag_mode 1)."""
    a = np.ascontiguousarray(agg, dtype=np.int64)
    if a.shape != (A.d.M,):
        raise ValueError("one aggregate per row of A")
    d = DMat()
    _check(lib().sp3mm_sa_prolongator(C.byref(A.d), a.ctypes.data, int(nagg), float(omega), int(diag_mode),
                                      C.byref(d)), "sp3mm_sa_prolongator")
    return DeviceCSR(d)


class Hierarchy:
    """A smoothed-aggregation hierarchy of A built on the device as one
    replayable plan (sp3mm_hier_*): level l forms P = (I - omega D^-1 A_l) P0,
    R = P^T and A_{l+1} = R A_l P.  update() recomputes every level from A's
    current values (same pattern), by default as one HIP graph replay; the
    matrices of level(l) are views owned by the hierarchy."""

    def __init__(self, A: DeviceCSR, aggs, omega: float = 2.0 / 3.0, diag_mode: int = 0, graph: bool = True):
        self._aggs = [np.ascontiguousarray(a, dtype=np.int64) for a, _ in aggs]
        self._nagg = np.array([int(n) for _, n in aggs], dtype=np.int64)
        ptrs = (C.c_void_p * len(self._aggs))(*[a.ctypes.data for a in self._aggs])
        self._A = A
        h = C.c_void_p()
        _check(lib().sp3mm_hier_create(C.byref(A.d), len(self._aggs), ptrs, self._nagg.ctypes.data, float(omega),
                                       int(diag_mode), int(graph), C.byref(h)), "sp3mm_hier_create")
        self.h = h

    def update(self, graph: bool = True) -> float:
        ms = C.c_double()
        _check(lib().sp3mm_hier_update(self.h, int(graph), C.byref(ms)), "sp3mm_hier_update")
        return ms.value

    def level(self, l: int):
        """(R_l, P_l, A_{l+1}) as views."""
        R, P, Ac = DMat(), DMat(), DMat()
        _check(lib().sp3mm_hier_level(self.h, int(l), C.byref(R), C.byref(P), C.byref(Ac)), "sp3mm_hier_level")
        return DeviceCSR(R), DeviceCSR(P), DeviceCSR(Ac)

    def info(self):
        n, ms, g = C.c_int32(), C.c_double(), C.c_int32()
        _check(lib().sp3mm_hier_info(self.h, C.byref(n), C.byref(ms), C.byref(g)), "sp3mm_hier_info")
        return {"levels": n.value, "ms_build": ms.value, "graph": bool(g.value)}

    @property
    def coarse(self) -> DeviceCSR:
        return self.level(len(self._aggs) - 1)[2]

    def free(self) -> None:
        if getattr(self, "h", None):
            lib().sp3mm_hier_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def hierarchy_device(A: DeviceCSR, aggs, omega: float = 2.0 / 3.0, diag_mode: int = 0, path: int = DIRECT):
    """Builds a smoothed-aggregation hierarchy on the device: for every
    (agg, nagg) in `aggs`, P = sa_prolongator(A), R = P^T, A <- R*A*P.
    Returns the list of (R, A, P) levels and the coarsest operator; the caller
    owns (and frees) every returned matrix except the input A."""
    levels = []
    cur = A
    for agg, nagg in aggs:
        P = sa_prolongator(cur, agg, nagg, omega, diag_mode)
        R = transpose(P)
        nxt, _ = rap(R, cur, P, path=path)
        levels.append((R, cur, P))
        cur = nxt
    return levels, cur


# ---------------------------------------------------------------------------
# level consumers (FP64 operators, device vectors as DeviceArray); queued on
# the library's stream: dot(), DeviceArray.numpy() and sync() wait for them
# ---------------------------------------------------------------------------
def spmv(A: DeviceCSR, x: DeviceArray, y: DeviceArray, alpha: float = 1.0, beta: float = 0.0) -> None:
    """y = alpha A x + beta y."""
    if x.n != A.d.N or y.n != A.d.M:
        raise ValueError("spmv: vector lengths do not match A")
    _check(lib().sp3mm_spmv(C.byref(A.d), x.ptr, y.ptr, float(alpha), float(beta)), "sp3mm_spmv")


def residual(A: DeviceCSR, b: DeviceArray, x: DeviceArray, r: DeviceArray) -> None:
    """r = b - A x."""
    if not (b.n == x.n == r.n == A.d.M):
        raise ValueError("residual: vector lengths do not match A")
    _check(lib().sp3mm_residual(C.byref(A.d), b.ptr, x.ptr, r.ptr), "sp3mm_residual")


def jacobi(A: DeviceCSR, b: DeviceArray, x: DeviceArray, work: DeviceArray, omega: float = 2.0 / 3.0,
           sweeps: int = 1) -> None:
    """`sweeps` damped Jacobi sweeps on A x = b, in place on x."""
    if not (b.n == x.n == A.d.M and work.n >= A.d.M):
        raise ValueError("jacobi: vector lengths do not match A")
    _check(lib().sp3mm_jacobi(C.byref(A.d), b.ptr, x.ptr, work.ptr, float(omega), int(sweeps)), "sp3mm_jacobi")


def axpby(alpha: float, x: DeviceArray | None, beta: float, y: DeviceArray) -> None:
    """y = alpha x + beta y (alpha = beta = 0 clears y)."""
    if x is not None and x.n != y.n:
        raise ValueError("axpby: vector lengths differ")
    _check(lib().sp3mm_axpby(y.n, float(alpha), x.ptr if x is not None else None, float(beta), y.ptr),
           "sp3mm_axpby")


def set_stream(handle: int | None) -> None:
    """Queue the library's work on a caller's HIP stream (e.g. a torch stream's
    cuda_stream handle); None restores the library's own stream."""
    _check(lib().sp3mm_set_stream(handle), "sp3mm_set_stream")


def get_stream() -> int | None:
    return lib().sp3mm_get_stream()


def sync() -> None:
    """Waits for the library's stream; raises on an error a queued call left."""
    _check(lib().sp3mm_sync(), "sp3mm_sync")


def dot(x: DeviceArray, y: DeviceArray) -> float:
    if x.n != y.n:
        raise ValueError("dot: vector lengths differ")
    out = C.c_double()
    _check(lib().sp3mm_dot(x.n, x.ptr, y.ptr, C.byref(out)), "sp3mm_dot")
    return out.value


def get_tuning() -> dict:
    t = Tuning()
    _check(lib().sp3mm_get_tuning(C.byref(t)), "sp3mm_get_tuning")
    return t.as_dict()


def set_tuning(t: dict | None) -> None:
    """Install a schedule for the direct kernels (None restores the defaults)."""
    if t is None:
        _check(lib().sp3mm_set_tuning(None), "sp3mm_set_tuning")
        return
    cur = get_tuning()
    cur.update(t)
    _check(lib().sp3mm_set_tuning(C.byref(Tuning(**cur))), "sp3mm_set_tuning")


def autotune(R: DeviceCSR, A: DeviceCSR, P: DeviceCSR, precision: int = FP64, guard: float = 0.0,
             reps: int = 3):
    """Time candidate schedules of the direct product and install the fastest;
    returns (schedule, ms_best, ms_default)."""
    o = Opts(DIRECT, precision, guard, 0, 0, None)
    t = Tuning()
    mb, md = C.c_double(), C.c_double()
    _check(lib().sp3mm_autotune(C.byref(R.d), C.byref(A.d), C.byref(P.d), C.byref(o), int(reps),
                                C.byref(t), C.byref(mb), C.byref(md)), "sp3mm_autotune")
    return t.as_dict(), mb.value, md.value


def hash_scores(R: DeviceCSR, A: DeviceCSR, P: DeviceCSR, H1: int, H2: int, rows):
    """The autotuner's hash-multiplier scores on the given rows of R: returns
    (multipliers, stage-1 scores, stage-2 scores, per-row distinct key counts
    as an (nrows, 2) array), all numpy arrays."""
    import numpy as np
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cap = 4096
    cand = np.zeros(cap, np.uint32)
    s1 = np.zeros(cap, np.uint64)
    s2 = np.zeros(cap, np.uint64)
    nk = np.zeros((len(rows), 2), np.int32)
    n = lib().sp3mm_hash_scores(C.byref(R.d), C.byref(A.d), C.byref(P.d), int(H1), int(H2),
                                rows.ctypes.data_as(C.c_void_p), len(rows), cap,
                                cand.ctypes.data_as(C.c_void_p), s1.ctypes.data_as(C.c_void_p),
                                s2.ctypes.data_as(C.c_void_p), nk.ctypes.data_as(C.c_void_p))
    if n < 0:
        raise RuntimeError(f"sp3mm_hash_scores failed: {last_error()}")
    return cand[:n], s1[:n], s2[:n], nk


def copy_probe(nbytes: int, reps: int = 20) -> float:
    """Milliseconds per device-to-device copy of nbytes (HBM bandwidth probe)."""
    return lib().sp3mm_copy_probe(int(nbytes), int(reps))


def rap_host(R: CSR, A: CSR, P: CSR, path: int = DIRECT, base: int = 0):
    """Upload, multiply, download: convenience for tests."""
    dR, dA, dP = (DeviceCSR.upload(m, base=base) for m in (R, A, P))
    out, st = rap(dR, dA, dP, path=path)
    res = out.download()
    for m in (dR, dA, dP, out):
        m.free()
    return res, st


# ---------------------------------------------------------------------------
# reference-compatible entry points over host spmat
# ---------------------------------------------------------------------------
class HostSpmat:
    """A host ``spmat`` whose arrays are owned by numpy (inputs of the drop-in calls)."""

    def __init__(self, m: CSR, base: int = 0):
        self.irp = np.ascontiguousarray(m.irp + base, dtype=np.uint64)
        self.ja = np.ascontiguousarray(m.ja + base, dtype=np.uint64)
        self.vals = np.ascontiguousarray(m.vals, dtype=np.float64)
        self.s = Spmat(m.nnz, m.M, m.N, self.vals.ctypes.data_as(C.POINTER(C.c_double)),
                       self.ja.ctypes.data_as(C.POINTER(C.c_ulong)),
                       self.irp.ctypes.data_as(C.POINTER(C.c_ulong)), 0)


def take_spmat(p, base: int = 0) -> CSR:
    """Copy a returned spmat into a 0-based CSR and release it with freeSpmat."""
    if not p:
        raise RuntimeError("drop-in call returned NULL (see stderr)")
    s = p.contents
    M, nz = int(s.M), int(s.NZ)
    irp = np.ctypeslib.as_array(s.IRP, shape=(M + 1,)).astype(np.int64) - base
    ja = (np.ctypeslib.as_array(s.JA, shape=(nz,)).astype(np.int64) - base) if nz else np.zeros(0, np.int64)
    vals = np.ctypeslib.as_array(s.AS, shape=(nz,)).copy() if nz else np.zeros(0)
    out = CSR(M, int(s.N), irp, ja, vals)
    lib().freeSpmat(p)
    return out


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def take_idx(p, n: int) -> np.ndarray:
    if not p:
        raise RuntimeError("drop-in call returned NULL (see stderr)")
    out = np.ctypeslib.as_array(p, shape=(n,)).astype(np.int64).copy()
    _libc.free(C.cast(p, C.c_void_p))
    return out


def default_config() -> Config:
    return Config(20, 2, IDXMAP, 1, None)


def sp3mm_row_by_row_pair(R: CSR, AC: CSR, P: CSR, base: int = 0, spmm: str | None = None) -> CSR:
    """sp3mmRowByRowPair_{base} (code/src/Sp3MM_CSR_OMP_UB_Generic.c:456-494)."""
    L = lib()
    r, a, p = HostSpmat(R, base), HostSpmat(AC, base), HostSpmat(P, base)
    cfg = default_config()
    f = C.cast(getattr(L, f"{spmm}{base}"), C.c_void_p).value if spmm else None
    out = getattr(L, f"sp3mmRowByRowPair_{base}")(C.byref(r.s), C.byref(a.s), C.byref(p.s),
                                                  C.byref(cfg), f)
    return take_spmat(out, base)


def sp3mm_row_by_row_merged(R: CSR, AC: CSR, P: CSR, base: int = 0) -> CSR:
    """sp3mmRowByRowMerged_{base} (code/src/Sp3MM_CSR_OMP_UB_Generic.c:498-573)."""
    L = lib()
    r, a, p = HostSpmat(R, base), HostSpmat(AC, base), HostSpmat(P, base)
    cfg = default_config()
    out = getattr(L, f"sp3mmRowByRowMerged_{base}")(C.byref(r.s), C.byref(a.s), C.byref(p.s),
                                                    C.byref(cfg), None)
    return take_spmat(out, base)


def spmm_dropin(name: str, A: CSR, B: CSR, base: int = 0) -> CSR:
    L = lib()
    a, b = HostSpmat(A, base), HostSpmat(B, base)
    cfg = default_config()
    return take_spmat(getattr(L, f"{name}{base}")(C.byref(a.s), C.byref(b.s), C.byref(cfg)), base)


def size_upper_bound(A: CSR, B: CSR, base: int = 0) -> np.ndarray:
    L = lib()
    a, b = HostSpmat(A, base), HostSpmat(B, base)
    return take_idx(getattr(L, f"spMMSizeUpperbound_{base}")(C.byref(a.s), C.byref(b.s)), A.M + 1)


def size_upper_bound_colparts(A: CSR, B: CSR, grid_cols: int, boffsets: np.ndarray,
                              base: int = 0) -> np.ndarray:
    L = lib()
    a, b = HostSpmat(A, base), HostSpmat(B, base)
    off = np.ascontiguousarray(boffsets, dtype=np.uint64)
    p = getattr(L, f"spMMSizeUpperboundColParts_{base}")(
        C.byref(a.s), C.byref(b.s), grid_cols, off.ctypes.data_as(C.POINTER(C.c_ulong)))
    return take_idx(p, A.M * grid_cols + 1)


def symb_row_lengths(A: CSR, B: CSR, base: int = 0, mode: int = IDXMAP) -> np.ndarray:
    L = lib()
    a, b = HostSpmat(A, base), HostSpmat(B, base)
    return take_idx(getattr(L, f"SpMM_Symb___{base}")(mode, C.byref(a.s), C.byref(b.s)), A.M + 1)


def symb_out_idxs(A: CSR, B: CSR, Cm: CSR | None = None, base: int = 0, mode: int = IDXMAP):
    """SpMM_Symb_OutIdxs__ / Sp3MM_Symb_OutIdxs_: (row lengths, list of row index arrays)."""
    L = lib()
    hs = [HostSpmat(m, base) for m in ((A, B) if Cm is None else (A, B, Cm))]
    rows = C.POINTER(C.POINTER(C.c_ulong))()
    if Cm is None:
        p = getattr(L, f"SpMM_Symb_OutIdxs__{base}")(mode, C.byref(hs[0].s), C.byref(hs[1].s),
                                                     C.byref(rows))
    else:
        p = getattr(L, f"Sp3MM_Symb_OutIdxs_{base}")(mode, C.byref(hs[0].s), C.byref(hs[1].s),
                                                     C.byref(hs[2].s), C.byref(rows))
    lens = take_idx(p, A.M + 1)
    out = []
    for r in range(A.M):
        n = int(lens[r])
        out.append(np.ctypeslib.as_array(rows[r], shape=(n,)).astype(np.int64) - base if n
                   else np.zeros(0, np.int64))
    first = C.cast(rows[0], C.c_void_p).value if A.M else None
    _libc.free(first)
    _libc.free(C.cast(rows, C.c_void_p))
    return lens, out


def symb_col_parts(A: CSR, B: CSR, grid_cols: int, base: int = 0, out_idxs: bool = False, mode: int = IDXMAP):
    """SpMM_Symb__ColParts_ / SpMM_Symb_OutIdxs_ColParts_: (row lengths, A.M x
    grid_cols per-share counts[, list of row index arrays])."""
    L = lib()
    hs = [HostSpmat(m, base) for m in (A, B)]
    parts = C.POINTER(C.c_ulong)()
    rows = C.POINTER(C.POINTER(C.c_ulong))()
    if out_idxs:
        p = getattr(L, f"SpMM_Symb_OutIdxs_ColParts_{base}")(mode, C.byref(hs[0].s), C.byref(hs[1].s), C.byref(rows),
                                                            grid_cols, C.byref(parts))
    else:
        p = getattr(L, f"SpMM_Symb__ColParts_{base}")(mode, C.byref(hs[0].s), C.byref(hs[1].s), grid_cols,
                                                     C.byref(parts))
    lens = take_idx(p, A.M + 1)
    counts = take_idx(parts, A.M * grid_cols).reshape(A.M, grid_cols)
    if not out_idxs:
        return lens, counts
    out = []
    for r in range(A.M):
        n = int(lens[r])
        out.append(np.ctypeslib.as_array(rows[r], shape=(n,)).astype(np.int64) - base if n
                   else np.zeros(0, np.int64))
    first = C.cast(rows[0], C.c_void_p).value if A.M else None
    _libc.free(first)
    _libc.free(C.cast(rows, C.c_void_p))
    return lens, counts, out


def elapsed_internal() -> float:
    return C.c_double.in_dll(lib(), "ElapsedInternal").value
