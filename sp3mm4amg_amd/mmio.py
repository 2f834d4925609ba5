"""Matrix Market (coordinate, real) reader and writer for CSR.

Semantics follow the reference's loader (code/src/lib/parser.c:55-118,171-229):
1-based COO triples, ``symmetric`` files expanded to both triangles, output
rows sorted by column.  Duplicate coordinates are rejected.
"""
from __future__ import annotations

import gzip

import numpy as np

from .csr import CSR, from_coo


def read(path: str) -> CSR:
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as f:
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
    if data.shape[0] != nz:
        raise ValueError(f"{path}: expected {nz} entries, found {data.shape[0]}")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = data[:, 2] if ncol == 3 else np.ones(nz)
    if symm == "symmetric":
        off = rows != cols
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, vals[off]]))
    return from_coo(M, N, rows, cols, vals)


def write(path: str, m: CSR) -> None:
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{m.M} {m.N} {m.nnz}\n")
        rows = m.row_ids() + 1
        for r, c, v in zip(rows, m.ja + 1, m.vals):
            f.write(f"{r} {c} {v!r}\n")
