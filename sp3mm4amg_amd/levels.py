"""AMG level groups on disk.

A level group is the reference's unit of test data: the restriction R, the
fine operator AC, the prolongator P and the expected coarse operator
AC_next = R*AC*P, each a Matrix Market file (AMG4PSBLAS level dumps).  The
reference lists them one group per line, four paths then an optional
``# comment`` (code/test/matrixGroups.list, read by
code/test/testAllMatrixGroups.sh and passed to code/test/Sp3MM_test.c's
main as argv[1..4]).

Binary level dumps store the same matrices as CSR arrays in one ``.npz``
(0-based int64 offsets, int32 columns, float64 values; no pickled objects),
so a large level reloads without parsing text.
"""
from __future__ import annotations

import lzma
import os
import time
from dataclasses import dataclass

import numpy as np

from . import mmio
from .csr import CSR

LEVEL_NAMES = ("r", "ac", "p", "ac_next")


@dataclass
class GroupEntry:
    r: str
    ac: str
    p: str
    ac_next: str | None
    label: str


def read_group_list(path: str) -> list[GroupEntry]:
    """Parses a group list (plain text or .xz)."""
    opener = lzma.open if path.endswith(".xz") else open
    out = []
    with opener(path, "rt") as f:
        for n, line in enumerate(f, 1):
            body, _, comment = line.partition("#")
            fields = body.split()
            if not fields:
                continue
            if len(fields) not in (3, 4):
                raise ValueError(f"{path}:{n}: expected R, AC, P and optionally AC_next paths")
            out.append(GroupEntry(*fields[:3], fields[3] if len(fields) == 4 else None, comment.strip()))
    return out


def load_group(entry: GroupEntry, root: str = ".", sorted_rows: str = "require") -> dict:
    """Reads a group's Matrix Market files (paths relative to `root`)."""
    out = {}
    for name in LEVEL_NAMES:
        rel = getattr(entry, name)
        if rel is not None:
            out[name] = mmio.read(os.path.join(root, rel), sorted_rows=sorted_rows)
    R, AC, P = out["r"], out["ac"], out["p"]
    if not (AC.M == AC.N == R.N == P.M and R.M == P.N):
        raise ValueError(f"group {entry.label!r}: shapes do not chain: R {R.shape}, AC {AC.shape}, P {P.shape}")
    return out


def save_level(path: str, **mats: CSR) -> None:
    """Binary level dump: each keyword's matrix as <name>_shape/_irp/_ja/_vals."""
    arrays = {}
    for name, m in mats.items():
        m.check()
        arrays[name + "_shape"] = np.array([m.M, m.N], np.int64)
        arrays[name + "_irp"] = np.ascontiguousarray(m.irp, np.int64)
        arrays[name + "_ja"] = np.ascontiguousarray(m.ja, np.int32)
        arrays[name + "_vals"] = np.ascontiguousarray(m.vals, np.float64)
    np.savez(path, **arrays)


def load_level(path: str) -> dict:
    """Reads a binary level dump (validating each matrix)."""
    with np.load(path, allow_pickle=False) as z:
        names = sorted({k.rsplit("_", 1)[0] for k in z.files if k.endswith("_shape")})
        out = {}
        for name in names:
            M, N = (int(x) for x in z[name + "_shape"])
            m = CSR(M, N, z[name + "_irp"].astype(np.int64), z[name + "_ja"].astype(np.int64),
                    z[name + "_vals"].astype(np.float64))
            m.check()
            out[name] = m
    return out


def check_level(R: CSR, AC: CSR, P: CSR, AC_next: CSR | None = None, path: int | None = None) -> dict:
    """R*AC*P on the GPU (as the reference's test driver forms the next level),
    compared with the expected AC_next when given: pattern equality, and the
    largest value difference relative to the largest |AC_next| entry."""
    from . import lib
    t0 = time.perf_counter()
    got, st = lib.rap_host(R, AC, P, path=lib.DIRECT if path is None else path)
    rec = {"shape": got.shape, "nnz": got.nnz, "ms_host_call": (time.perf_counter() - t0) * 1e3,
           "ms_device": st["ms_total"]}
    if AC_next is not None:
        same = got.shape == AC_next.shape and np.array_equal(got.irp, AC_next.irp) and \
            np.array_equal(got.ja, AC_next.ja)
        rec["pattern_equal"] = bool(same)
        if same and got.nnz:
            scale = float(np.max(np.abs(AC_next.vals))) or 1.0
            rec["max_diff_rel_to_max"] = float(np.max(np.abs(got.vals - AC_next.vals))) / scale
    return rec
