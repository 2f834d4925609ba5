"""The autotuner's hash-multiplier scores (k_hash_score through
sp3mm_hash_scores) against a CPU model written here from the definition.

For a row r of R the stage-1 keys are the distinct columns of the rows of A
that row r of R selects, the stage-2 keys the distinct columns of the rows of
P those keys select.  A candidate multiplier C sends key k to the home slot
((k * C) mod 2^32) >> (32 - log2 H) of a table of H slots; inserting a row's
distinct keys by linear probing, the sum over keys of (final slot - home
slot) mod H does not depend on the insert order.  The kernel sums it over
rows per candidate; a row with more than 1024 distinct keys at a stage, or
more than 7/8 of the stage's table, adds nothing to that stage.  The scores
are integers, so the comparison is exact.
"""
import numpy as np
import pytest

from sp3mm4amg_amd import lib as L
from sp3mm4amg_amd import problems

pytestmark = pytest.mark.gpu
BIG = np.iinfo(np.int32).max


@pytest.fixture(scope="module", autouse=True)
def gpu_lib():
    L.init(0)
    yield


def distinct_keys(R, A, P, r):
    left = R.ja[R.irp[r]:R.irp[r + 1]]
    k1 = np.unique(np.concatenate([A.ja[A.irp[k]:A.irp[k + 1]] for k in left])) if len(left) else \
        np.zeros(0, np.int64)
    if len(k1) > 1024:
        return k1, None
    k2 = np.unique(np.concatenate([P.ja[P.irp[k]:P.irp[k + 1]] for k in k1])) if len(k1) else \
        np.zeros(0, np.int64)
    return k1, k2


def displacement(keys, mul, H):
    shift = 32 - (H.bit_length() - 1)
    mask = H - 1
    home = (((keys.astype(np.uint64) * np.uint64(mul)) & np.uint64(0xFFFFFFFF)) >> np.uint64(shift)).astype(np.int64)
    occupied = np.zeros(H, bool)
    total = 0
    for h in home.tolist():
        s = h
        while occupied[s]:
            s = (s + 1) & mask
        occupied[s] = True
        total += (s - h) & mask
    return total


def model(R, A, P, rows, H1, H2, cands):
    s1 = np.zeros(len(cands), np.uint64)
    s2 = np.zeros(len(cands), np.uint64)
    nk = np.zeros((len(rows), 2), np.int64)
    for i, r in enumerate(rows):
        k1, k2 = distinct_keys(R, A, P, r)
        n1 = len(k1)
        n2 = len(k2) if k2 is not None else BIG
        nk[i] = (n1 if n1 <= 1024 else BIG, n2 if n2 <= 1024 else BIG)
        use1 = n1 <= 1024 and n1 <= H1 - H1 // 8
        use2 = k2 is not None and n2 <= 1024 and n2 <= H2 - H2 // 8
        for c, mul in enumerate(cands):
            if use1:
                s1[c] += displacement(k1, int(mul), H1)
            if use2:
                s2[c] += displacement(k2, int(mul), H2)
    return s1, s2, nk


CASES = [
    ("poisson3d_16", lambda: problems.poisson3d_sa(16), 256, 128),
    ("poisson3d_16_small", lambda: problems.poisson3d_sa(16), 128, 64),
    ("stencil27_12", lambda: problems.stencil27_sa(12, seed=3), 1024, 256),
    ("irregular_3000", lambda: problems.irregular_sa(3000, seed=4), 4096, 4096),
]


@pytest.mark.parametrize("name,make,H1,H2", CASES, ids=[c[0] for c in CASES])
def test_hash_scores_match_model(name, make, H1, H2):
    p = make()
    R, A, P = p["R"], p["A"], p["P"]
    rows = np.unique(np.linspace(0, R.M - 1, 12).astype(np.int32))
    print(f"{name}: {R.M} rows of R, scoring rows {rows.tolist()} on tables of {H1} and {H2}", flush=True)
    d = [L.DeviceCSR.upload(m) for m in (R, A, P)]
    try:
        cand, s1, s2, nk = L.hash_scores(*d, H1, H2, rows)
    finally:
        for m in d:
            m.free()
    assert len(cand) == 256 and cand[0] == 0x9E3779B1 and np.all(cand % 2 == 1)
    # a few candidates in full (the model is plain Python), and every row's key counts
    pick = np.unique(np.r_[0, 1, 2, np.argmin(s1), np.argmin(s2), np.argmax(s1), np.argmax(s2), 255])
    m1, m2, mk = model(R, A, P, rows, H1, H2, cand[pick])
    assert np.array_equal(nk, mk), f"key counts {nk.tolist()} vs {mk.tolist()}"
    assert np.array_equal(s1[pick], m1), f"stage-1 scores {s1[pick]} vs {m1}"
    assert np.array_equal(s2[pick], m2), f"stage-2 scores {s2[pick]} vs {m2}"


def test_hash_scores_reject_bad_tables():
    p = problems.poisson3d_sa(8)
    d = [L.DeviceCSR.upload(m) for m in (p["R"], p["A"], p["P"])]
    try:
        for H1, H2 in ((100, 64), (64, 32), (8192, 64)):
            with pytest.raises(RuntimeError):
                L.hash_scores(*d, H1, H2, [0])
        with pytest.raises(RuntimeError):
            L.hash_scores(*d, 64, 64, [p["R"].M])  # row out of range
    finally:
        for m in d:
            m.free()
