"""Pins the CPU oracle (oracle/sp3mm_oracle.c) before it is trusted as the checker.

What fixes each part, independently of the oracle's own formulas:
  * the reference's known-answer group code/test/{r,ac,p,ac_next}: exact pattern,
    values to 1e-12 relative (ac_next is printed to 19 digits);
  * outputs of the reference library itself, committed as golden vectors
    (tests/golden/ref_outputs_*.npz, written by tests/golden/make_golden.py):
    bit-exact pattern and values for the pair, direct and 1-based entry points;
  * the live reference build (oracle/_ref) on random matrices, when present;
  * integer-valued random matrices, whose products are exact in binary64 so any
    summation order must reproduce numpy's dense integer product bit for bit,
    with the pattern fixed by the boolean-semiring product (explicit zeros kept);
  * an x87 long-double recomputation bounding the rounding error;
  * closed forms: ub[r] = sum of right row lengths; 1-based = 0-based + 1.
Planted mistakes (dropped term, transposed operand, shifted index, perturbed
value) must be caught by the same comparisons.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref
from sp3mm4amg_amd.csr import CSR, from_coo

TOL = 1e-12


def rand_csr(rng, M, N, density, ints=False, zero_frac=0.0):
    mask = rng.random((M, N)) < density
    r, c = np.nonzero(mask)
    if ints:
        v = rng.integers(-3, 4, size=r.size).astype(np.float64)
    else:
        v = rng.uniform(-1, 1, size=r.size)
    if zero_frac:
        v[rng.random(r.size) < zero_frac] = 0.0
    return from_coo(M, N, r, c, v)


def boolean_pattern(R, A, P):
    pr = (R.to_dense() != 0) | _pattern_dense(R)
    pa = _pattern_dense(A)
    pp = _pattern_dense(P)
    ra = (pr.astype(np.int64) @ pa.astype(np.int64)) > 0
    return (ra.astype(np.int64) @ pp.astype(np.int64)) > 0


def _pattern_dense(m):
    d = np.zeros((m.M, m.N), dtype=bool)
    d[m.row_ids(), m.ja] = True
    return d


def assert_close(got, want, scale=None, tol=TOL):
    assert got.pattern_equal(want)
    assert O.rel_errors(got, want, scale).max() <= tol


# ---------------------------------------------------------------------------
def test_fixture_known_answer(fixture_2744):
    R, A, P, AN = (fixture_2744[k] for k in ("r", "ac", "p", "ac_next"))
    assert (R.M, R.N, A.M, A.N, P.M, P.N, AN.M, AN.N) == (343, 2744, 2744, 2744, 2744, 343, 343, 343)
    assert (R.nnz, A.nnz, P.nnz, AN.nnz) == (12642, 18032, 12642, 11189)
    scale = O.abs_product(R, A, P)
    for fn in (O.sp3mm_pair, O.sp3mm_merged):
        C = fn(R, A, P)
        assert_close(C, AN, scale)


def test_golden_reference_outputs_bit_exact(fixture_2744, poisson2d64, ref_outputs):
    cases = {"fixture": (fixture_2744["r"], fixture_2744["ac"], fixture_2744["p"]),
             "poisson2d64": poisson2d64}
    for case, (R, A, P) in cases.items():
        g = ref_outputs[case]
        pair = O.sp3mm_pair(R, A, P)
        merged = O.sp3mm_merged(R, A, P)
        for key, got in (("pair_symbnum", pair), ("pair_2d", pair), ("merged", merged),
                         ("pair_symbnum_1", O.sp3mm_pair(R, A, P, base=1)),
                         ("merged_1", O.sp3mm_merged(R, A, P, base=1))):
            want = g[key]
            assert got.pattern_equal(want), (case, key)
            assert np.array_equal(got.vals, want.vals), (case, key)
        assert np.array_equal(O.ub(R, A), g["ub_ra"])
        assert np.array_equal(O.symb(R, A), g["symb_ra"])


def test_faithful_reset_policy_is_bit_identical(poisson2d64):
    R, A, P = poisson2d64
    for fn in (O.sp3mm_pair, O.sp3mm_merged):
        a, b = fn(R, A, P, faithful=False), fn(R, A, P, faithful=True)
        assert a.pattern_equal(b) and np.array_equal(a.vals, b.vals)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_integer_products_exact_and_boolean_pattern(seed):
    rng = np.random.default_rng(seed)
    R = rand_csr(rng, 17, 41, 0.15, ints=True, zero_frac=0.1)
    A = rand_csr(rng, 41, 37, 0.12, ints=True, zero_frac=0.1)
    P = rand_csr(rng, 37, 13, 0.2, ints=True, zero_frac=0.1)
    dense = (R.to_dense() @ A.to_dense()) @ P.to_dense()
    pat = boolean_pattern(R, A, P)
    for fn in (O.sp3mm_pair, O.sp3mm_merged):
        C = fn(R, A, P)
        C.check()
        assert np.array_equal(_pattern_dense(C), pat)
        assert np.array_equal(C.to_dense(), dense)
    # explicit zeros survive: some stored entries of the product are exactly 0
    C = O.sp3mm_pair(R, A, P)
    assert C.nnz == int(pat.sum())


@pytest.mark.parametrize("seed", [3, 4])
def test_spmm_matches_dense_and_symbolic(seed):
    rng = np.random.default_rng(seed)
    A = rand_csr(rng, 50, 60, 0.1)
    B = rand_csr(rng, 60, 45, 0.1)
    C = O.spmm(A, B)
    np.testing.assert_allclose(C.to_dense(), A.to_dense() @ B.to_dense(), rtol=0, atol=1e-14)
    lens = O.symb(A, B)
    assert np.array_equal(lens[:-1], np.diff(C.irp)) and lens[-1] == C.nnz


def test_upper_bound_closed_form(poisson2d64, fixture_2744):
    for (A, B) in (poisson2d64[:2], (fixture_2744["r"], fixture_2744["ac"])):
        ub = O.ub(A, B)
        blen = np.diff(B.irp)
        want = np.add.reduceat(blen[A.ja], A.irp[:-1]) if A.nnz else np.zeros(A.M, np.int64)
        want[np.diff(A.irp) == 0] = 0
        assert np.array_equal(ub[:-1], want) and ub[-1] == blen[A.ja].sum()


def test_upper_bound_colparts_sums_to_bound(poisson2d64):
    R, A, _ = poisson2d64
    for g in (1, 3, 7):
        parts = O.ub_colparts(R, A, g)
        ub = O.ub(R, A)
        assert np.array_equal(parts[:-1].reshape(R.M, g).sum(1), ub[:-1])
        assert parts[-1] == ub[-1]
    # brute force per range for g=3
    g = 3
    cb, cr = A.N // g, A.N % g
    starts = [i * cb + min(i, cr) for i in range(g)] + [A.N]
    parts = O.ub_colparts(R, A, g)[:-1].reshape(R.M, g)
    for r in range(0, R.M, 97):
        ks = R.ja[R.irp[r]:R.irp[r + 1]]
        for i in range(g):
            n = sum(int(((A.ja[A.irp[k]:A.irp[k + 1]] >= starts[i]) &
                         (A.ja[A.irp[k]:A.irp[k + 1]] < starts[i + 1])).sum()) for k in ks)
            assert parts[r, i] == n


def test_direct_symbolic_equals_pair_pattern(fixture_2744, poisson2d64):
    for R, A, P in ((fixture_2744["r"], fixture_2744["ac"], fixture_2744["p"]), poisson2d64):
        lens = O.symb3(R, A, P)
        C = O.sp3mm_pair(R, A, P)
        assert np.array_equal(lens[:-1], np.diff(C.irp)) and lens[-1] == C.nnz


def test_long_double_bound(fixture_2744, poisson2d64):
    """FP64 results sit within the summation error bound of an extended-precision
    recomputation: |x - x_ld| <= 64 eps (|R||A||P|)_ij, and within the 1e-12 contract."""
    eps = np.finfo(np.float64).eps
    for R, A, P in ((fixture_2744["r"], fixture_2744["ac"], fixture_2744["p"]), poisson2d64):
        L = O.sp3mm_pair_ld(R, A, P)
        scale = O.abs_product(R, A, P)
        for fn in (O.sp3mm_pair, O.sp3mm_merged):
            C = fn(R, A, P)
            assert_close(C, L, scale)
            assert np.all(np.abs(C.vals - L.vals) <= 64 * eps * scale.vals)


def test_error_bound_holds_for_reference_orders():
    """The tolerance model used against the GPU holds for the oracle's two
    summation orders (pair: sorted k; direct: first touch) against long double,
    on random values with heavy cancellation."""
    from sp3mm4amg_amd import problems
    p = problems.irregular_sa(1500, seed=9)
    R, A, P = p["R"], p["A"], p["P"]
    L = O.sp3mm_pair_ld(R, A, P)
    for fn in (O.sp3mm_pair, O.sp3mm_merged):
        C = fn(R, A, P)
        b = O.error_bound(R, A, P, L)
        assert np.all(np.abs(C.vals - L.vals) <= b / 2)
    # and it is not vacuous: the plain 1e-12 relative check fails here
    assert O.rel_errors(O.sp3mm_merged(R, A, P), L).max() > 1e-12


def test_one_based_is_shift(poisson2d64):
    R, A, P = poisson2d64
    a, b = O.sp3mm_merged(R, A, P), O.sp3mm_merged(R, A, P, base=1)
    assert a.pattern_equal(b) and np.array_equal(a.vals, b.vals)


@pytest.mark.skipif(not ref.available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("seed", [5, 6])
def test_live_reference_random(seed):
    rng = np.random.default_rng(seed)
    R = rand_csr(rng, 40, 90, 0.08)
    A = rand_csr(rng, 90, 90, 0.06)
    P = rand_csr(rng, 90, 30, 0.1)
    for base in (0, 1):
        got = O.sp3mm_pair(R, A, P, base=base)
        want = ref.sp3mm_pair(R, A, P, base=base, threads=2)
        assert got.pattern_equal(want) and np.array_equal(got.vals, want.vals)
        got = O.sp3mm_merged(R, A, P, base=base)
        want = ref.sp3mm_merged(R, A, P, base=base, threads=2)
        assert got.pattern_equal(want) and np.array_equal(got.vals, want.vals)
        assert np.array_equal(O.ub(R, A, base), ref.ub(R, A, base))
        assert np.array_equal(O.symb(R, A, base), ref.symb(R, A, base))


# ---------------------------------------------------------------------------
# planted mistakes: the comparisons used against the GPU path must catch them
def test_planted_mistakes_are_caught(fixture_2744):
    R, A, P, AN = (fixture_2744[k] for k in ("r", "ac", "p", "ac_next"))
    scale = O.abs_product(R, A, P)
    good = O.sp3mm_pair(R, A, P)
    assert_close(good, AN, scale)
    # a dropped term: one A entry removed
    keep = np.ones(A.nnz, bool)
    keep[A.nnz // 2] = False
    rows = A.row_ids()[keep]
    A2 = from_coo(A.M, A.N, rows, A.ja[keep], A.vals[keep])
    bad = O.sp3mm_pair(R, A2, P)
    assert not (bad.pattern_equal(AN) and O.rel_errors(bad, AN, scale).max() <= TOL)
    # a transposed operand (the fixture's AC is not symmetric)
    bad = O.sp3mm_pair(R, A.transpose(), P)
    assert not (bad.pattern_equal(AN) and O.rel_errors(bad, AN, scale).max() <= TOL)
    # a shifted column index
    ja = good.ja.copy()
    ja[10] += 1
    assert not CSR(good.M, good.N, good.irp, ja, good.vals).pattern_equal(AN)
    # a value off by 1e-10 relative
    v = good.vals.copy()
    v[123] *= 1 + 1e-10
    assert O.rel_errors(CSR(good.M, good.N, good.irp, good.ja, v), AN, scale).max() > TOL
