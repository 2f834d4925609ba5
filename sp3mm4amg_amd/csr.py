"""Host-side CSR container used by the Python mirror of the Sp3MM interface.

A ``CSR`` holds a sorted compressed-sparse-row matrix in 0-based numpy arrays:
``irp`` (int64, M+1 row offsets), ``ja`` (int64 column indices, strictly
increasing inside each row) and ``vals`` (float64 or float32).  This mirrors the
reference's ``spmat`` (code/src/include/sparseMatrix.h:41-53), whose index type
``idx_t`` is a 64-bit unsigned integer (code/src/include/config.h:51); the
1-based (Fortran) form is produced only at the C boundary.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class CSR:
    M: int
    N: int
    irp: np.ndarray
    ja: np.ndarray
    vals: np.ndarray

    def __post_init__(self):
        self.M = int(self.M)
        self.N = int(self.N)
        self.irp = np.ascontiguousarray(self.irp, dtype=np.int64)
        self.ja = np.ascontiguousarray(self.ja, dtype=np.int64)
        if self.vals.dtype not in (np.float64, np.float32):
            self.vals = self.vals.astype(np.float64)
        self.vals = np.ascontiguousarray(self.vals)

    @property
    def nnz(self) -> int:
        return int(self.irp[-1]) if self.M > 0 else 0

    @property
    def shape(self):
        return (self.M, self.N)

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.irp)

    def row_ids(self) -> np.ndarray:
        """Row index of every stored entry."""
        return np.repeat(np.arange(self.M, dtype=np.int64), self.row_lengths())

    def check(self) -> None:
        """Raise ValueError unless the matrix is well-formed sorted CSR."""
        if self.irp.shape != (self.M + 1,):
            raise ValueError("irp has wrong length")
        if self.M and self.irp[0] != 0:
            raise ValueError("irp[0] != 0")
        if np.any(np.diff(self.irp) < 0):
            raise ValueError("irp not monotone")
        if self.ja.shape[0] != self.nnz or self.vals.shape[0] != self.nnz:
            raise ValueError("ja/vals length != nnz")
        if self.nnz:
            if self.ja.min() < 0 or self.ja.max() >= self.N:
                raise ValueError("column index out of range")
            rows = self.row_ids()
            same = rows[1:] == rows[:-1]
            if np.any(self.ja[1:][same] <= self.ja[:-1][same]):
                raise ValueError("columns not strictly increasing inside a row")

    def transpose(self) -> "CSR":
        """Sorted CSR of the transpose (stable counting sort by column)."""
        rows = self.row_ids()
        order = np.argsort(self.ja, kind="stable")
        cnt = np.bincount(self.ja, minlength=self.N) if self.nnz else np.zeros(self.N, np.int64)
        irp = np.zeros(self.N + 1, dtype=np.int64)
        np.cumsum(cnt, out=irp[1:])
        return CSR(self.N, self.M, irp, rows[order], self.vals[order])

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.M, self.N), dtype=np.float64)
        d[self.row_ids(), self.ja] = self.vals
        return d

    def pattern_equal(self, other: "CSR") -> bool:
        return (self.M == other.M and self.N == other.N
                and np.array_equal(self.irp, other.irp)
                and np.array_equal(self.ja, other.ja))

    def astype(self, dtype) -> "CSR":
        return CSR(self.M, self.N, self.irp, self.ja, self.vals.astype(dtype))

    def with_vals(self, vals: np.ndarray) -> "CSR":
        return CSR(self.M, self.N, self.irp, self.ja, vals)


def from_coo(M: int, N: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray,
             sum_duplicates: bool = False) -> CSR:
    """Sorted CSR from COO triples.  Duplicates are summed (in input order) or rejected."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals)
    key = rows * np.int64(N) + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    v = vals[order]
    if key.size:
        first = np.ones(key.size, dtype=bool)
        first[1:] = key[1:] != key[:-1]
        if not first.all():
            if not sum_duplicates:
                raise ValueError("duplicate entries in COO input")
            starts = np.flatnonzero(first)
            v = np.add.reduceat(v, starts)
            key = key[starts]
    r = key // N if N else key
    c = key - r * N
    irp = np.zeros(M + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=M), out=irp[1:])
    return CSR(M, N, irp, c, v)


def block_diag(mats) -> CSR:
    """Block-diagonal concatenation of CSR matrices (rows and columns offset)."""
    mats = list(mats)
    M = sum(m.M for m in mats)
    N = sum(m.N for m in mats)
    irp = [np.zeros(1, np.int64)]
    ja, vals = [], []
    roff = coff = 0
    for m in mats:
        irp.append(m.irp[1:] + roff)
        ja.append(m.ja + coff)
        vals.append(m.vals)
        roff += m.nnz
        coff += m.N
    return CSR(M, N, np.concatenate(irp), np.concatenate(ja) if ja else np.zeros(0, np.int64),
               np.concatenate(vals) if vals else np.zeros(0))
