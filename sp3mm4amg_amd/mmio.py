"""Matrix Market (coordinate, real/integer/pattern) reader and writer for CSR.

Semantics follow the reference's loader (code/src/lib/parser.c:55-118 MMtoCOO,
:171-229 COOtoCSR): 1-based COO triples, a ``pattern`` entry counts as 1.0, a
``symmetric`` file's off-diagonal entry is followed by its mirror, and each
row's entries must appear in strictly increasing column order (the
reference's default CONSISTENCY_CHECKS build rejects the file otherwise,
parser.c:205-211).  ``sorted_rows="sort"`` accepts any order instead (and
rejects duplicate coordinates).  Compressed files are read directly for the
extensions the reference's test driver decompresses (.gz, .xz, .bz2, .zip;
code/src/commons/utils.c:472).
"""
from __future__ import annotations

import bz2
import gzip
import io
import lzma
import zipfile

import numpy as np

from .csr import CSR, from_coo


def _open_text(path: str):
    if path.endswith(".gz"):
        return gzip.open(path, "rt")
    if path.endswith(".xz"):
        return lzma.open(path, "rt")
    if path.endswith(".bz2"):
        return bz2.open(path, "rt")
    if path.endswith(".zip"):
        z = zipfile.ZipFile(path)
        names = [n for n in z.namelist() if not n.endswith("/")]
        if len(names) != 1:
            raise ValueError(f"{path}: expected one file in the archive, found {len(names)}")
        return io.TextIOWrapper(z.open(names[0]))
    return open(path, "rt")


def read(path: str, sorted_rows: str = "require") -> CSR:
    if sorted_rows not in ("require", "sort"):
        raise ValueError("sorted_rows is 'require' or 'sort'")
    with _open_text(path) as f:
        header = f.readline().split()
        if len(header) < 5 or header[0] != "%%MatrixMarket" or header[1].lower() != "matrix":
            raise ValueError(f"{path}: not a Matrix Market matrix file")
        if header[2].lower() != "coordinate":
            raise ValueError(f"{path}: only coordinate format is supported")
        field, symm = header[3].lower(), header[4].lower()
        if field not in ("real", "integer", "pattern"):
            raise ValueError(f"{path}: unsupported field {field}")
        if symm not in ("general", "symmetric"):
            raise ValueError(f"{path}: unsupported symmetry {symm}")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        M, N, nz = (int(x) for x in line.split())
        ncol = 2 if field == "pattern" else 3
        data = np.loadtxt(f, dtype=np.float64, ndmin=2, max_rows=nz) if nz else np.zeros((0, ncol))
    if data.shape[0] != nz or (nz and data.shape[1] < ncol):
        raise ValueError(f"{path}: expected {nz} entries of {ncol} fields")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = data[:, 2].copy() if ncol == 3 else np.ones(nz)
    if nz and (rows.min() < 0 or rows.max() >= M or cols.min() < 0 or cols.max() >= N):
        raise ValueError(f"{path}: an entry lies outside the {M} x {N} matrix")
    if symm == "symmetric":
        # each entry, then its mirror (off the diagonal), in file order
        mirror = rows != cols
        r2 = np.stack([rows, cols], axis=1).ravel()
        c2 = np.stack([cols, rows], axis=1).ravel()
        v2 = np.repeat(vals, 2)
        keep = np.stack([np.ones(nz, bool), mirror], axis=1).ravel()
        rows, cols, vals = r2[keep], c2[keep], v2[keep]
        if rows.size and (rows.max() >= M or cols.max() >= N):
            raise ValueError(f"{path}: a mirrored entry lies outside the {M} x {N} matrix")
    if sorted_rows == "sort":
        return from_coo(M, N, rows, cols, vals)
    order = np.argsort(rows, kind="stable")
    r, c = rows[order], cols[order]
    bad = np.flatnonzero((r[1:] == r[:-1]) & (c[1:] <= c[:-1]))
    if bad.size:
        i = order[bad[0] + 1]
        raise ValueError(f"{path}: entry ({rows[i] + 1}, {cols[i] + 1}) is not after the previous "
                         f"entry of its row in column order (unsorted row or duplicate)")
    irp = np.zeros(M + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=M), out=irp[1:])
    return CSR(M, N, irp, c, vals[order])


def write(path: str, m: CSR, symmetric: bool = False) -> None:
    """Writes `m` (rows in column order, 1-based); with `symmetric`, only the
    lower triangle of a symmetric matrix is written, in column-major order
    (so that reading it back, mirrors included, keeps every row sorted)."""
    if symmetric:
        if m.M != m.N:
            raise ValueError("a symmetric file needs a square matrix")
        rows = m.row_ids()
        low = rows >= m.ja
        order = np.lexsort((rows[low], m.ja[low]))
        r, c, v = rows[low][order], m.ja[low][order], m.vals[low][order]
        kind = "symmetric"
    else:
        r, c, v = m.row_ids(), m.ja, m.vals
        kind = "general"
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real {kind}\n")
        f.write(f"{m.M} {m.N} {len(r)}\n")
        for a, b, x in zip(r + 1, c + 1, v):
            f.write(f"{a} {b} {float(x)!r}\n")
