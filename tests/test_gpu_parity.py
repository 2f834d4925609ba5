"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Bar: row offsets and sorted column indices bit-exact.  FP64 values: within
1e-12 relative per nonzero on the stencil workloads (whose entries do not
cancel below 1e-2 of their magnitude scale), and within the rounding-error
model |x - y| <= 1e-12 |y| + 4 T_r u (|R||A||P|)_ij on the random-valued
workloads (oracle.error_bound), which the reference's own summation order
also satisfies against the exact value.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sp3mm4amg_amd import lib as L
from sp3mm4amg_amd import problems
from sp3mm4amg_amd.csr import CSR, from_coo

pytestmark = pytest.mark.gpu
TOL = 1e-12


def check(got: CSR, want: CSR, scale: CSR, tol=TOL, bound=None):
    got.check()
    assert got.shape == want.shape
    assert np.array_equal(got.irp, want.irp), "row offsets differ"
    assert np.array_equal(got.ja, want.ja), "column indices differ"
    if bound is None:
        err = O.rel_errors(got, want, scale)
        assert err.max(initial=0.0) <= tol, f"max rel err {err.max()}"
    else:
        diff = np.abs(got.vals - want.vals)
        assert np.all(diff <= bound), f"max excess {np.max(diff / bound)}"


def triple(name):
    if name == "poisson2d64":
        p = problems.poisson2d_sa(64)
    elif name == "poisson3d32":
        p = problems.poisson3d_sa(32)
    elif name == "stencil27_15":
        p = problems.stencil27_sa(15, seed=3)
    elif name == "irregular3000":
        p = problems.irregular_sa(3000, seed=4)
    else:
        raise KeyError(name)
    return p["R"], p["A"], p["P"]


@pytest.fixture(scope="module", autouse=True)
def gpu_lib():
    L.init(0)
    yield


# ---------------------------------------------------------------------------
# drop-in entry points on the reference's own known-answer group
@pytest.mark.parametrize("base", [0, 1])
def test_fixture_dropin(fixture_2744, base):
    R, A, P, AN = (fixture_2744[k] for k in ("r", "ac", "p", "ac_next"))
    print(f"fixture_2744 base={base}", flush=True)
    scale = O.abs_product(R, A, P)
    oracle_pair = O.sp3mm_pair(R, A, P)
    for got in (L.sp3mm_row_by_row_pair(R, A, P, base=base),
                L.sp3mm_row_by_row_pair(R, A, P, base=base, spmm="spmmRowByRow_SymbNum_"),
                L.sp3mm_row_by_row_merged(R, A, P, base=base)):
        check(got, AN, scale)
        check(got, oracle_pair, scale)
    assert L.elapsed_internal() > 0


@pytest.mark.parametrize("base", [0, 1])
def test_every_spmm_entry_point(poisson2d64, base):
    R, A, P = poisson2d64
    print(f"poisson2d64 spmm entry points base={base}", flush=True)
    want_ra = O.spmm(R, A)
    scale_ra = O.spmm(R.with_vals(np.abs(R.vals)), A.with_vals(np.abs(A.vals)))
    want = O.sp3mm_pair(R, A, P)
    scale = O.abs_product(R, A, P)
    for name in L.SPMM_NAMES:
        check(L.spmm_dropin(name, R, A, base=base), want_ra, scale_ra)
        check(L.sp3mm_row_by_row_pair(R, A, P, base=base, spmm=name), want, scale)


def test_caller_supplied_spmm_is_composed_on_host(poisson2d64):
    """A foreign SPMM_INTERF is called twice on the host, as the reference does."""
    R, A, P = poisson2d64
    lib = L.lib()
    calls = []
    inner = lib.spmmRowByRow_0

    @C.CFUNCTYPE(C.c_void_p, C.POINTER(L.Spmat), C.POINTER(L.Spmat), C.POINTER(L.Config))
    def my_spmm(a, b, cfg):
        calls.append(1)
        return C.cast(inner(a, b, cfg), C.c_void_p).value

    r, a, p = L.HostSpmat(R), L.HostSpmat(A), L.HostSpmat(P)
    cfg = L.default_config()
    out = lib.sp3mmRowByRowPair_0(C.byref(r.s), C.byref(a.s), C.byref(p.s), C.byref(cfg),
                                  C.cast(my_spmm, C.c_void_p).value)
    got = L.take_spmat(out)
    assert len(calls) == 2
    check(got, O.sp3mm_pair(R, A, P), O.abs_product(R, A, P))


@pytest.mark.parametrize("base", [0, 1])
def test_sizing_and_symbolic_entry_points(poisson2d64, fixture_2744, base):
    for R, A, P in (poisson2d64, (fixture_2744["r"], fixture_2744["ac"], fixture_2744["p"])):
        print(f"sizing M={R.M} base={base}", flush=True)
        assert np.array_equal(L.size_upper_bound(R, A, base=base), O.ub(R, A))
        assert np.array_equal(L.symb_row_lengths(R, A, base=base), O.symb(R, A))
        for g in (1, 3):
            # column-range offsets of A's rows, as colsOffsetsPartitioningUnifRanges_ builds them
            cb, cr = A.N // g, A.N % g
            starts = np.array([i * cb + min(i, cr) for i in range(g)], np.int64)
            off = np.empty(A.M * g + 1, np.int64)
            for i in range(g):
                off[i:-1:g] = [A.irp[k] + np.searchsorted(A.ja[A.irp[k]:A.irp[k + 1]], starts[i])
                               for k in range(A.M)]
            off[-1] = A.nnz
            assert np.array_equal(L.size_upper_bound_colparts(R, A, g, off, base=base),
                                  O.ub_colparts(R, A, g))
        lens, rows = L.symb_out_idxs(R, A, base=base)
        RA = O.spmm(R, A)
        assert np.array_equal(lens, O.symb(R, A))
        assert all(np.array_equal(rows[r], RA.ja[RA.irp[r]:RA.irp[r + 1]]) for r in range(R.M))
        lens, rows = L.symb_out_idxs(R, A, P, base=base)
        C3 = O.sp3mm_pair(R, A, P)
        assert np.array_equal(lens, O.symb3(R, A, P))
        assert all(np.array_equal(rows[r], C3.ja[C3.irp[r]:C3.irp[r + 1]]) for r in range(R.M))


@pytest.mark.parametrize("base", [0, 1])
def test_symbolic_column_parts(poisson2d64, fixture_2744, base):
    """SpMM_Symb__ColParts_ / _OutIdxs_ColParts_: per row, the product's
    nonzeros in each uniform share of B's columns (the first N % g shares one
    column wider), against the oracle's pattern of R*A."""
    for R, A, _ in (poisson2d64, (fixture_2744["r"], fixture_2744["ac"], fixture_2744["p"])):
        RA = O.spmm(R, A)
        for g in (1, 3, 7, 64):
            sizes = [A.N // g + (1 if i < A.N % g else 0) for i in range(g)]
            edges = np.concatenate([[0], np.cumsum(sizes)])
            want = np.stack([np.histogram(RA.ja[RA.irp[r]:RA.irp[r + 1]], bins=edges)[0] for r in range(R.M)])
            lens, counts = L.symb_col_parts(R, A, g, base=base)
            assert np.array_equal(lens, O.symb(R, A))
            assert np.array_equal(counts, want), f"g={g}"
            assert np.array_equal(counts.sum(axis=1), lens[:-1])
            lens2, counts2, rows = L.symb_col_parts(R, A, g, base=base, out_idxs=True)
            assert np.array_equal(lens2, lens) and np.array_equal(counts2, want)
            assert all(np.array_equal(rows[r], RA.ja[RA.irp[r]:RA.irp[r + 1]]) for r in range(R.M))


# ---------------------------------------------------------------------------
# device-resident path on scaled versions of the benchmark workloads
@pytest.mark.parametrize("name", ["poisson2d64", "poisson3d32", "stencil27_15", "irregular3000"])
@pytest.mark.parametrize("path", [L.PAIR, L.DIRECT])
def test_workloads_device_api(name, path):
    R, A, P = triple(name)
    print(f"{name} path={path}", flush=True)
    got, st = L.rap_host(R, A, P, path=path)
    want = O.sp3mm_pair(R, A, P) if path == L.PAIR else O.sp3mm_merged(R, A, P)
    if name.startswith("irregular"):
        check(got, want, None, bound=O.error_bound(R, A, P, want))
    else:
        check(got, want, O.abs_product(R, A, P))
    assert st["nnz_out"] == want.nnz
    # the direct path skips measuring rows when they all fit one bin by the
    # longest rows of R and A (mac1 -1 then)
    assert st["mac1"] == O.ub(R, A)[-1] or (path == L.DIRECT and st["mac1"] == -1)
    if path == L.PAIR:
        assert st["nnz_mid"] == O.symb(R, A)[-1]


def test_device_ub_and_symbolic_arrays():
    R, A, P = triple("poisson3d32")
    dR, dA, dP = (L.DeviceCSR.upload(m) for m in (R, A, P))
    assert np.array_equal(L.ub(dR, dA), O.ub(R, A))
    assert np.array_equal(L.symbolic(dR, dA), O.symb(R, A))
    assert np.array_equal(L.symbolic(dR, dA, dP), O.symb3(R, A, P))
    for m in (dR, dA, dP):
        m.free()


# ---------------------------------------------------------------------------
# mixed precision: FP32 storage, FP32-pair accumulation, FP64 where it cancels
def _rounded32(m: CSR) -> CSR:
    return m.with_vals(m.vals.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("name,guard", [("stencil27_15", 0.0), ("stencil27_15", 1e-3),
                                        ("poisson3d32", 0.0)])
def test_mixed_precision_against_fp64_oracle(name, guard):
    """Bar: max relative error <= 1e-10 against the FP64 oracle evaluated on the
    same FP32-rounded inputs; pattern bit-exact; FP64 row sums within the
    rounding-error model of an exact row sum."""
    R, A, P = triple(name)
    print(f"mixed {name} guard={guard}", flush=True)
    got, st, rs = L.rap_mixed_host(R, A, P, guard=guard, rowsum=True)
    R32, A32, P32 = _rounded32(R), _rounded32(A), _rounded32(P)
    want = O.sp3mm_pair_ld(R32, A32, P32)
    assert np.array_equal(got.irp, want.irp) and np.array_equal(got.ja, want.ja)
    err = O.rel_errors(got, want)
    print(f"  max rel {err.max():.3e} median {np.median(err):.3e} fp64 rows {st['guarded_rows']}")
    assert err.max() <= 1e-10
    # FP64 row sums: against the exactly summed long-double row entries
    import math
    exact = np.array([math.fsum(want.vals[want.irp[i]:want.irp[i + 1]]) for i in range(want.M)])
    scale = O.abs_product(R32, A32, P32)
    srow = np.add.reduceat(scale.vals, scale.irp[:-1])
    T = (O.ub(R32, A32)[:-1] + O.ub(O.spmm(R32, A32), P32)[:-1]).astype(np.float64)
    assert np.all(np.abs(rs - exact) <= 4 * T * 2.0 ** -53 * srow + 1e-15 * np.abs(exact))


@pytest.mark.parametrize("name", ["stencil27_15", "poisson3d32", "irregular3000"])
def test_mixed64_against_fp64_oracle(name):
    """FP32 storage with FP64 accumulation: every FP32 x FP32 product is exact in
    FP64, so the result obeys the FP64 rounding-error model of the oracle on the
    same FP32-rounded inputs; pattern bit-exact."""
    R, A, P = (_rounded32(m) for m in triple(name))
    print(f"mixed64 {name}", flush=True)
    got, st, _ = L.rap_mixed_host(R, A, P, precision=L.MIXED64)
    want = O.sp3mm_pair(R, A, P)
    check(got, want, None, bound=O.error_bound(R, A, P, want))
    err = O.rel_errors(got, O.sp3mm_pair_ld(R, A, P), O.abs_product(R, A, P))
    print(f"  max rel vs long double {err.max():.3e}")
    assert err.max() <= 1e-10


def test_mixed64_rejects_rowsum():
    R, A, P = triple("poisson2d64")
    with pytest.raises(RuntimeError):
        L.rap_mixed_host(R, A, P, rowsum=True, precision=L.MIXED64)


def test_mixed_precision_requires_fp32_operands():
    R, A, P = triple("poisson2d64")
    dR, dA, dP = (L.DeviceCSR.upload(m) for m in (R, A, P))
    with pytest.raises(RuntimeError):
        L.rap(dR, dA, dP, path=L.DIRECT, precision=L.MIXED)
    with pytest.raises(RuntimeError):
        L.rap(dR, dA, dP, path=L.PAIR, precision=L.MIXED)
    for m in (dR, dA, dP):
        m.free()


# ---------------------------------------------------------------------------
# setup operations on the device: transpose, smoothed-aggregation prolongator
def _rect_with_empties():
    rng = np.random.default_rng(3)
    M, N = 300, 170
    rows = rng.integers(0, M, 2000)
    rows = rows[rows % 7 != 0]          # empty rows
    cols = rng.integers(0, N, rows.size)
    cols = cols[: rows.size]
    keep = cols % 11 != 0               # empty columns
    return from_coo(M, N, rows[keep], cols[keep], rng.standard_normal(int(keep.sum())), sum_duplicates=True)


@pytest.mark.parametrize("which", ["P_poisson2d64", "P_irregular3000", "A_irregular3000", "rect"])
def test_transpose_exact(which):
    if which == "rect":
        m = _rect_with_empties()
    else:
        kind, name = which.split("_", 1)
        R, A, P = triple(name)
        m = {"P": P, "A": A}[kind]
    print(f"transpose {which}: {m.shape} nnz {m.nnz}", flush=True)
    d = L.DeviceCSR.upload(m)
    t = L.transpose(d)
    got = t.download()
    t.free()
    d.free()
    want = m.transpose()
    got.check()
    assert got.shape == want.shape
    assert np.array_equal(got.irp, want.irp) and np.array_equal(got.ja, want.ja)
    assert np.array_equal(got.vals, want.vals)


def _check_prolongator(A, agg, nagg, mode):
    dA = L.DeviceCSR.upload(A)
    dP = L.sa_prolongator(dA, agg, nagg, diag_mode=mode)
    got = dP.download()
    dP.free()
    dA.free()
    name = "diag" if mode == 0 else "absplus1"
    want = problems.smoothed_prolongator(A, agg, nagg, diag_mode=name)
    scale = problems.smoothed_prolongator(A, agg, nagg, diag_mode=name, absolute=True)
    got.check()
    assert np.array_equal(got.irp, want.irp) and np.array_equal(got.ja, want.ja)
    n = np.diff(A.irp)[want.row_ids()] + 1.0
    bound = 1e-14 * np.abs(want.vals) + 4 * n * 2.0 ** -53 * scale.vals
    assert np.all(np.abs(got.vals - want.vals) <= bound)


def test_sa_prolongator_structured():
    for dims, block in (((64, 64), (2, 2)), ((15, 15, 15), (3, 3, 3))):
        A = problems.poisson_operator(dims) if len(dims) == 2 else problems.stencil27_random(dims[0], seed=3)
        agg, cdims = problems.structured_aggregation(dims, block)
        print(f"prolongator {dims}", flush=True)
        _check_prolongator(A, agg, int(np.prod(cdims)), 0)


def test_sa_prolongator_irregular_and_missing_diagonal():
    A = problems.lognormal_matrix(3000, 4)
    agg, nagg = problems.random_aggregation(3000, 5)
    _check_prolongator(A, agg, nagg, 1)
    # rows without a stored diagonal get the identity term inserted
    rows = A.row_ids()
    keep = ~((A.ja == rows) & (rows % 5 == 0))
    B = from_coo(A.M, A.N, rows[keep], A.ja[keep], A.vals[keep])
    _check_prolongator(B, agg, nagg, 1)
    with pytest.raises(RuntimeError):  # zero diagonal with diag_mode 0
        _check_prolongator(B, agg, nagg, 0)


def test_hierarchy_on_device():
    """Three smoothed-aggregation levels from the 16^3 7-point operator built on
    the GPU (prolongator, transpose, R*A*P) against the host construction."""
    dims = (16, 16, 16)
    A = problems.poisson_operator(dims)
    aggs, host = [], []
    cur = A
    for lvl in range(3):
        agg, cdims = problems.structured_aggregation(dims, (2, 2, 2))
        nagg = int(np.prod(cdims))
        aggs.append((agg, nagg))
        R, _, P = problems.galerkin_inputs(cur, agg, nagg)
        host.append((R, cur, P))
        cur = O.sp3mm_pair(R, cur, P)
        dims = cdims
    dA = L.DeviceCSR.upload(A)
    levels, coarse = L.hierarchy_device(dA, aggs)
    for i, (dR, dAl, dP) in enumerate(levels):
        R, Al, P = host[i]
        got = dP.download()
        assert np.array_equal(got.irp, P.irp) and np.array_equal(got.ja, P.ja)
        assert np.max(np.abs(got.vals - P.vals)) <= 1e-13
        got = dR.download()
        assert np.array_equal(got.irp, R.irp) and np.array_equal(got.ja, R.ja)
        dR.free()
        dP.free()
        if i > 0:
            dAl.free()
    got = coarse.download()
    coarse.free()
    dA.free()
    print(f"coarsest {got.shape} nnz {got.nnz}")
    check(got, cur, O.abs_product(*host[-1]), tol=1e-11)


# ---------------------------------------------------------------------------
# schedules: every tuning gives the oracle's pattern and values
SCHEDULES = [
    {"subwave_h": 0},                                   # wavefront groups only
    {"subwave_h": 0, "h1_shift": 1},
    {"subwave_h": 128, "h2_div": 2},
    {"subwave_h": 512, "u1": 2, "u2": 2, "slog1_cap": 1, "slog2_cap": 0},
    {"h1_shift": 1, "h2_div": 8},                       # tables that fill -> retry path
    {"h1_shift": 2, "h2_div": 1, "u1": 2, "u2": 4},
    {"xcd_map": 0},                                     # rows interleaved over all workgroups
    {"xcd_map": 0, "subwave_h": 0, "u2": 3},
    {"hash1": 0xD0B79761, "hash2": 0xCB06E8FB},        # other table hash multipliers
    {"hash1": 1, "hash2": 1},                           # every key of a row on one home slot
    {"hash1": 0xFFFFFFFF, "hash2": 0x00010001, "h1_shift": 1},
    {"h1_shift": 3, "h2_div": 8, "u1": 2, "u2": 3},     # tables at an eighth of the bound: many rows retried
]


@pytest.mark.parametrize("sched", range(len(SCHEDULES)))
@pytest.mark.parametrize("name", ["poisson3d32", "irregular3000", "stencil27_15"])
def test_schedules_keep_parity(name, sched):
    R, A, P = triple(name)
    print(f"schedule {SCHEDULES[sched]} on {name}", flush=True)
    want = O.sp3mm_pair(R, A, P)
    bound = O.error_bound(R, A, P, want)
    L.set_tuning(SCHEDULES[sched])
    try:
        # the pair path's SpGEMMs hash with the schedule's multipliers too
        for path in (L.DIRECT, L.PAIR) if "hash1" in SCHEDULES[sched] else (L.DIRECT,):
            got, st = L.rap_host(R, A, P, path=path)
            check(got, want, None, bound=bound)
        R32, A32, P32 = _rounded32(R), _rounded32(A), _rounded32(P)
        want32 = O.sp3mm_pair(R32, A32, P32)
        got, st, _ = L.rap_mixed_host(R32, A32, P32, precision=L.MIXED64)
        check(got, want32, None, bound=O.error_bound(R32, A32, P32, want32))
        for guard in (0.0, 1e-2):
            print(f"  mixed guard={guard}", flush=True)
            got, st, _ = L.rap_mixed_host(R32, A32, P32, guard=guard, precision=L.MIXED)
            assert got.pattern_equal(want32)
            err = O.rel_errors(got, want32, O.abs_product(R32, A32, P32))
            assert err.max() <= 1e-10 or guard == 0.0
    finally:
        L.set_tuning(None)


def test_long_rows_window_and_spill():
    """Rows beyond the hash tables: the LDS window kernel takes rows whose
    column windows fit in LDS, the global dense kernel the rest (spilled)."""
    p = problems.irregular_sa(40000, seed=4)
    R, A, P = p["R"], p["A"], p["P"]
    print(f"irregular40000: {R.M} coarse rows", flush=True)
    want = O.sp3mm_pair(R, A, P)
    bound = O.error_bound(R, A, P, want)
    got, st = L.rap_host(R, A, P, path=L.DIRECT)
    print(f"  off-table rows {st['guarded_entries']} spilled {st['spilled_rows']}")
    assert st["guarded_entries"] > 0 and st["spilled_rows"] > 0
    check(got, want, None, bound=bound)
    # windows that fit: a structured problem whose tables are forced to fill
    R, A, P = triple("poisson3d32")
    want = O.sp3mm_pair(R, A, P)
    L.set_tuning({"h1_shift": 2, "h2_div": 8})
    try:
        got, st = L.rap_host(R, A, P, path=L.DIRECT)
    finally:
        L.set_tuning(None)
    print(f"  poisson3d32 forced retries: off-table rows {st['guarded_entries']} spilled {st['spilled_rows']}")
    assert st["guarded_entries"] > 0 and st["spilled_rows"] == 0
    check(got, want, O.abs_product(R, A, P))


def test_pair_path_long_rows():
    """The pair form on the 40K irregular workload: R*A has rows beyond the
    hash tables whose columns (A's 40K) exceed LDS (global dense kernel), and
    (R*A)*P has long rows over P's ~5K columns (LDS dense kernel)."""
    p = problems.irregular_sa(40000, seed=4)
    R, A, P = p["R"], p["A"], p["P"]
    want = O.sp3mm_pair(R, A, P)
    got, st = L.rap_host(R, A, P, path=L.PAIR)
    print(f"  pair: nnz RA {st['nnz_mid']} RAP {st['nnz_out']}")
    check(got, want, None, bound=O.error_bound(R, A, P, want))


def _sampled_rows(out: L.DeviceCSR, rows: np.ndarray) -> CSR:
    """Rows `rows` of a device product, copied back one segment at a time."""
    d = out.d
    rp = np.empty(d.M + 1, dtype=np.int64)
    L.lib().sp3mm_copy_to_host(rp.ctypes.data, d.rp, (d.M + 1) * 8)
    irp, cols, vals = [0], [], []
    for r in rows:
        a, b = int(rp[r]), int(rp[r + 1])
        cj = np.empty(b - a, dtype=np.int32)
        cv = np.empty(b - a, dtype=np.float64)
        L.lib().sp3mm_copy_to_host(cj.ctypes.data, d.col + 4 * a, (b - a) * 4)
        L.lib().sp3mm_copy_to_host(cv.ctypes.data, d.val + 8 * a, (b - a) * 8)
        cols.append(cj.astype(np.int64))
        vals.append(cv)
        irp.append(irp[-1] + b - a)
    return CSR(len(rows), d.N, np.array(irp, dtype=np.int64), np.concatenate(cols), np.concatenate(vals))


def _row_subset(R: CSR, rows: np.ndarray) -> CSR:
    irp, ja, va = [0], [], []
    for r in rows:
        a, b = R.irp[r], R.irp[r + 1]
        ja.append(R.ja[a:b])
        va.append(R.vals[a:b])
        irp.append(irp[-1] + b - a)
    return CSR(len(rows), R.N, np.array(irp, dtype=np.int64), np.concatenate(ja), np.concatenate(va))


def _irregular_sampled(n, reassociated=None):
    """Lognormal irregular operator with random aggregation (the irregular
    workload): full product on the GPU, 40 rows (first, last, random) against
    the oracle's R[rows]*A*P, patterns exact, values within the error model."""
    p = problems.irregular_sa(n, seed=4)
    R, A, P = p["R"], p["A"], p["P"]
    print(f"irregular {n}: R {R.shape} nnz {R.nnz}", flush=True)
    d = [L.DeviceCSR.upload(m) for m in (R, A, P)]
    out, st = L.rap(*d, path=L.DIRECT)
    print(f"  nnz {st['nnz_out']} off-table rows {st['guarded_entries']} spilled {st['spilled_rows']} "
          f"reassociated {st['reassociated']} ({st['ms_total']:.1f} ms)", flush=True)
    if reassociated is not None:
        assert st["reassociated"] == reassociated
    rows = np.unique(np.concatenate([[0, R.M - 1], np.random.default_rng(7).choice(R.M, 38, replace=False)]))
    got = _sampled_rows(out, rows)
    out.free()
    for m in d:
        m.free()
    Rs = _row_subset(R, rows)
    want = O.sp3mm_pair(Rs, A, P)
    check(got, want, None, bound=O.error_bound(Rs, A, P, want))


def test_irregular_100k_sampled():
    """P's ~12.5K columns fit in LDS: the dense rows' second stage runs there
    (direct forwarding, no reassociation)."""
    _irregular_sampled(100_000, reassociated=0)


def test_irregular_200k_sampled():
    """P's ~25K columns exceed an LDS accumulator and nearly every row is
    beyond the one-wavefront tables: the direct product runs as R*(A*P), its
    long rows in LDS column windows (k_spgemm_window)."""
    _irregular_sampled(200_000, reassociated=1)


def _random_csr(rng, M, N, lens):
    rows = np.repeat(np.arange(M), lens)
    cols = np.concatenate([np.sort(rng.choice(N, int(l), replace=False)) for l in lens])
    return from_coo(M, N, rows, cols, rng.uniform(-1.0, 1.0, rows.size))


def test_spgemm_window_long_wide_rows():
    """SpGEMM rows beyond the hash tables whose output is wider than one LDS
    accumulator: the numeric pass runs in LDS column windows of 16K columns
    (k_spgemm_window; 60,000 columns = 4 windows), including a row of 5,000
    left entries, whose cursors beyond the 4,096 kept in LDS are found by
    binary search.  Pattern exact, values within the rounding-error model."""
    rng = np.random.default_rng(21)
    M, K, N = 48, 8000, 60000
    la = rng.integers(150, 400, M)
    la[3] = 5000
    A = _random_csr(rng, M, K, la)
    B = _random_csr(rng, K, N, rng.integers(20, 60, K))
    dA, dB = L.DeviceCSR.upload(A), L.DeviceCSR.upload(B)
    try:
        out, st = L.spgemm(dA, dB)
        got = out.download()
        out.free()
    finally:
        dA.free()
        dB.free()
    want = O.spmm(A, B)
    scale = O.spmm(A.with_vals(np.abs(A.vals)), B.with_vals(np.abs(B.vals)))
    T = O.ub(A, B)[:-1][want.row_ids()].astype(np.float64)
    bound = 1e-12 * np.abs(want.vals) + 4.0 * T * 2.0 ** -53 * scale.vals
    check(got, want, None, bound=bound)


@pytest.mark.slow
def test_irregular_1m_sampled():
    """The full-size irregular workload: 1,000,000 fine rows, ~1.1e10 output
    nonzeros (about 140 GB of HBM for the output and workspaces), run as
    R*(A*P) with LDS column windows."""
    _irregular_sampled(1_000_000, reassociated=1)


def test_autotune_mixed():
    R, A, P = (_rounded32(m) for m in triple("stencil27_15"))
    d = [L.DeviceCSR.upload(m, vtype=L.F32) for m in (R, A, P)]
    try:
        best, ms_best, ms_default = L.autotune(*d, precision=L.MIXED, reps=1)
        print(f"autotune mixed: {best} {ms_best:.4f} ms (start {ms_default:.4f} ms)")
    finally:
        L.set_tuning(None)
        for m in d:
            m.free()


def test_autotune_installs_a_valid_schedule():
    R, A, P = triple("poisson3d32")
    d = [L.DeviceCSR.upload(m) for m in (R, A, P)]
    default = L.get_tuning()
    best, ms_best, ms_default = L.autotune(*d, reps=2)
    print(f"autotune: {best} {ms_best:.4f} ms (start {ms_default:.4f} ms)")
    assert ms_best <= ms_default
    assert L.get_tuning() == best
    assert best["hash1"] % 2 == 1 and best["hash2"] % 2 == 1
    out, _ = L.rap(*d, path=L.DIRECT)
    got = out.download()
    out.free()
    want = O.sp3mm_pair(R, A, P)
    check(got, want, O.abs_product(R, A, P))
    L.set_tuning(None)
    assert L.get_tuning() == default
    with pytest.raises(RuntimeError):
        L.set_tuning({"u1": 3})
    with pytest.raises(RuntimeError):
        L.set_tuning({"hash1": 2})  # multipliers are odd
    for m in d:
        m.free()


# ---------------------------------------------------------------------------
# batches: hierarchy levels + uneven independent problems in one product
def _batch_problems():
    probs = []
    levels = problems.hierarchy_3d(16, levels=3, rap=lambda R, A, P: O.sp3mm_pair(R, A, P))
    probs += [(lv["R"], lv["A"], lv["P"]) for lv in levels]
    for p in problems.poisson2d_batch(9, seed=11, lo=8, hi=40):
        probs.append((p["R"], p["A"], p["P"]))
    return probs


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed64"])
def test_batch_graph_replays(prec):
    probs = _batch_problems()
    print(f"batch of {len(probs)} precision={prec}", flush=True)
    mixed = prec != "fp64"
    vt = L.F32 if mixed else L.F64
    rnd = (lambda m: m.with_vals(m.vals.astype(np.float32).astype(np.float64))) if mixed else (lambda m: m)
    dev = [[L.DeviceCSR.upload(rnd(m), vtype=vt) for m in t] for t in probs]
    wants = [O.sp3mm_pair(*[rnd(m) for m in t]) for t in probs]
    scales = [O.abs_product(*[rnd(m) for m in t]) for t in probs]
    code = {"fp64": L.FP64, "mixed": L.MIXED, "mixed64": L.MIXED64}[prec]
    b = L.Batch([d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev], graph=True, precision=code)
    assert b.info()["graph"]
    tol = 1e-10 if prec == "mixed" else 1e-12

    def check_all():
        for i, want in enumerate(wants):
            out = b.result(i)
            got = out.download()
            out.free()
            check(got, want, scales[i], tol=tol)

    check_all()  # the eager run inside create()
    for rep in range(3):
        b.run(graph=True)
        check_all()
        # eager work and a sync between replays
        L.rap_host(*probs[0], path=L.DIRECT)
        L.lib().sp3mm_sync()
    b.run(graph=False)
    check_all()
    b.free()
    for t in dev:
        for m in t:
            m.free()


@pytest.mark.parametrize("prec", ["fp64", "mixed64"])
def test_batch_autotune_rebuilds(prec):
    """Tuning a batch installs a valid schedule and rebuilds its plan and
    graph; replays under it match the oracle."""
    probs = _batch_problems()
    mixed = prec != "fp64"
    vt = L.F32 if mixed else L.F64
    rnd = (lambda m: m.with_vals(m.vals.astype(np.float32).astype(np.float64))) if mixed else (lambda m: m)
    dev = [[L.DeviceCSR.upload(rnd(m), vtype=vt) for m in t] for t in probs]
    wants = [O.sp3mm_pair(*[rnd(m) for m in t]) for t in probs]
    scales = [O.abs_product(*[rnd(m) for m in t]) for t in probs]
    b = L.Batch([d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev], graph=True,
                precision=L.FP64 if not mixed else L.MIXED64)
    try:
        best, ms_best, ms_default = b.autotune(reps=2)
        print(f"batch autotune {prec}: {best} {ms_best:.3f} ms (start {ms_default:.3f})")
        assert best == L.get_tuning()
        assert b.info()["graph"]
        for rep in range(3):
            b.run(graph=True)
            for i, want in enumerate(wants):
                out = b.result(i)
                check(out.download(), want, scales[i])
                out.free()
    finally:
        b.free()
        L.set_tuning(None)
        for t in dev:
            for m in t:
                m.free()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_repack_values_and_changed_pattern(graph):
    """repack() refreshes the batch's packed operands from the caller's
    matrices: new values give the new product; a pattern change that alters
    the product's size is reported by run(), not returned as a wrong result."""
    R, A, P = triple("poisson2d64")
    R2, A2, P2 = triple("poisson3d32")
    d1 = [L.DeviceCSR.upload(m) for m in (R, A, P)]
    d2 = [L.DeviceCSR.upload(m) for m in (R2, A2, P2)]
    b = L.Batch([d1[0], d2[0]], [d1[1], d2[1]], [d1[2], d2[2]], graph=graph)
    try:
        b.run(graph=graph)
        # new values of A in the first problem (same pattern)
        rng = np.random.default_rng(11)
        A_new = A.with_vals(A.vals * rng.uniform(0.5, 1.5, A.nnz))
        dA_new = L.DeviceCSR.upload(A_new)
        b.repack(As=[dA_new, d2[1]])
        b.run(graph=graph)
        for i, (r, a, p) in enumerate(((R, A_new, P), (R2, A2, P2))):
            out = b.result(i)
            check(out.download(), O.sp3mm_merged(r, a, p), O.abs_product(r, a, p))
            out.free()
        # a changed pattern: the last entry of row 65 of A (an interior point
        # of the 64x64 grid, columns 1, 64, 65, 66, 129) moves to column 4000,
        # far across the grid, keeping the row sorted and the nonzero count;
        # R*A*P gains entries (12,676 -> 12,685 nonzeros)
        e1 = int(A.irp[66])
        ja = A.ja.copy()
        assert ja[e1 - 1] == 129
        ja[e1 - 1] = 4000
        A_moved = CSR(A.M, A.N, A.irp, ja, A.vals)
        A_moved.check()
        want = O.sp3mm_merged(R, A_moved, P)
        assert want.nnz != O.sp3mm_merged(R, A, P).nnz
        dA_moved = L.DeviceCSR.upload(A_moved)
        b.repack(As=[dA_moved, d2[1]])
        with pytest.raises(RuntimeError, match="replayed product"):
            b.run(graph=graph)
        # shapes or nonzero counts that differ from the batch's are refused
        with pytest.raises(RuntimeError, match="differ"):
            b.repack(As=[d2[1], d1[1]])
        dA_new.free()
        dA_moved.free()
    finally:
        b.free()
        for m in d1 + d2:
            m.free()


# ---------------------------------------------------------------------------
# edge cases
def test_edge_cases_shapes_and_zeros():
    rng = np.random.default_rng(7)
    # R with empty rows, A with an empty row, explicit zero values that cancel
    R = from_coo(5, 6, [0, 0, 2, 4, 4], [1, 3, 5, 0, 2], [1.0, 2.0, 0.0, 1.0, -1.0])
    A = from_coo(6, 6, [0, 1, 1, 3, 5, 2], [0, 2, 4, 3, 5, 2], [1.0, 1.0, -1.0, 2.0, 0.5, 1.0])
    P = from_coo(6, 3, [0, 1, 2, 3, 4, 5], [0, 1, 1, 2, 0, 2], [1.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    for path in (L.PAIR, L.DIRECT):
        got, _ = L.rap_host(R, A, P, path=path)
        want = O.sp3mm_pair(R, A, P)
        check(got, want, O.abs_product(R, A, P))
        assert np.any(got.vals == 0.0)  # a structural entry that sums to zero is kept
    # all-empty product
    Z = CSR(4, 6, np.zeros(5, np.int64), np.zeros(0, np.int64), np.zeros(0))
    got, _ = L.rap_host(Z, A, P, path=L.DIRECT)
    assert got.nnz == 0 and got.shape == (4, 3)
    got = L.sp3mm_row_by_row_pair(Z, A, P)
    assert got.nnz == 0
    # 1x1
    one = CSR(1, 1, np.array([0, 1]), np.array([0]), np.array([3.0]))
    got = L.sp3mm_row_by_row_merged(one, one, one)
    assert got.nnz == 1 and got.vals[0] == 27.0
    # non-square, random
    Rr = from_coo(30, 70, *_rand_coo(rng, 30, 70, 0.1))
    Ar = from_coo(70, 50, *_rand_coo(rng, 70, 50, 0.08))
    Pr = from_coo(50, 20, *_rand_coo(rng, 50, 20, 0.15))
    for path in (L.PAIR, L.DIRECT):
        got, _ = L.rap_host(Rr, Ar, Pr, path=path)
        want = O.sp3mm_pair(Rr, Ar, Pr)
        check(got, want, None, bound=O.error_bound(Rr, Ar, Pr, want))


def _rand_coo(rng, M, N, d):
    r, c = np.nonzero(rng.random((M, N)) < d)
    return r, c, rng.uniform(-1, 1, r.size)


def test_error_paths_return_null(poisson2d64, capfd):
    R, A, P = poisson2d64
    lib = L.lib()
    cfg = L.default_config()
    r, a = L.HostSpmat(R), L.HostSpmat(A)
    # dimension mismatch R*R*P
    out = lib.sp3mmRowByRowPair_0(C.byref(r.s), C.byref(r.s), C.byref(a.s), C.byref(cfg), None)
    assert not out
    # unsorted columns in a row (every host struct stays alive across the call)
    bad = CSR(A.M, A.N, A.irp, A.ja.copy(), A.vals)
    bad.ja[[0, 1]] = bad.ja[[1, 0]]
    b, pp = L.HostSpmat(bad), L.HostSpmat(P)
    out = lib.sp3mmRowByRowMerged_0(C.byref(r.s), C.byref(b.s), C.byref(pp.s), C.byref(cfg), None)
    assert not out
    # inconsistent row offsets are rejected on the host, before any device work
    broken = CSR(A.M, A.N, A.irp.copy(), A.ja, A.vals)
    broken.irp[5] = broken.irp[6] + 3
    bb = L.HostSpmat(broken)
    out = lib.sp3mmRowByRowMerged_0(C.byref(r.s), C.byref(bb.s), C.byref(pp.s), C.byref(cfg), None)
    assert not out
    err = capfd.readouterr().err
    assert "sp3mmRowByRowPair_0" in err and "sp3mmRowByRowMerged_0" in err
    assert "row offsets decrease" in err
    with pytest.raises(RuntimeError):
        L.DeviceCSR.upload(bad)
    # the device is still healthy after the rejected inputs
    got, _ = L.rap_host(R, A, P, path=L.DIRECT)
    check(got, O.sp3mm_merged(R, A, P), O.abs_product(R, A, P))


# ---------------------------------------------------------------------------
@pytest.mark.slow
def test_poisson3d128_full_size():
    """The benchmark workload at full size: both paths against the oracle."""
    p = problems.poisson3d_sa(128)
    R, A, P = p["R"], p["A"], p["P"]
    scale = O.abs_product(R, A, P)
    dR, dA, dP = (L.DeviceCSR.upload(m) for m in (R, A, P))
    for path, want in ((L.DIRECT, O.sp3mm_merged(R, A, P)), (L.PAIR, O.sp3mm_pair(R, A, P))):
        print(f"poisson3d128 path={path}", flush=True)
        out, st = L.rap(dR, dA, dP, path=path)
        got = out.download()
        out.free()
        check(got, want, scale)
        assert st["nnz_out"] == 8382712
        # R rows of at most 32 entries, A rows of at most 7: every bound fits
        # the 256-slot bin, so the direct path bins without measuring
        assert st["mac1"] == (-1 if path == L.DIRECT else 57740288)
    for m in (dR, dA, dP):
        m.free()


# ---------------------------------------------------------------------------
# device-pointer import/export (CSR already in HBM, reference spmat layout)
@pytest.mark.parametrize("base", [0, 1])
def test_device_import_export_round_trip(poisson2d64, base):
    R, A, P = poisson2d64
    dev = []
    for m in (R, A, P):
        irp = L.DeviceArray.from_numpy(m.irp.astype(np.int64) + base)
        ja = L.DeviceArray.from_numpy(m.ja.astype(np.int64) + base)
        va = L.DeviceArray.from_numpy(m.vals.astype(np.float64))
        dev.append(L.DeviceCSR.import_device(m.M, m.N, m.nnz, irp, ja, va, base=base))
        for v in (irp, ja, va):
            v.free()
    out, _ = L.rap(*dev, path=L.DIRECT)
    want = out.download()
    oirp, oja, ova = L.DeviceArray(out.d.M + 1, np.int64), L.DeviceArray(out.d.nnz, np.int64), L.DeviceArray(out.d.nnz)
    out.export_device(oirp, oja, ova, base=base)
    assert np.array_equal(oirp.numpy() - base, want.irp)
    assert np.array_equal(oja.numpy() - base, want.ja)
    assert np.array_equal(ova.numpy(), want.vals)
    check(want, O.sp3mm_merged(R, A, P), O.abs_product(R, A, P))
    for v in (oirp, oja, ova):
        v.free()
    for m in dev + [out]:
        m.free()


def test_device_import_rejects_bad_csr():
    n = 4
    good_irp = np.array([0, 1, 2, 3, 4], dtype=np.int64)
    cases = {
        "decreasing offsets": (np.array([0, 2, 1, 3, 4], dtype=np.int64), np.array([0, 1, 2, 3])),
        "offsets past nnz": (np.array([0, 1, 2, 3, 9], dtype=np.int64), np.array([0, 1, 2, 3])),
        "column out of range": (good_irp, np.array([0, 1, 2, 4])),
    }
    for what, (irp, ja) in cases.items():
        di = L.DeviceArray.from_numpy(irp)
        dj = L.DeviceArray.from_numpy(ja.astype(np.int64))
        dv = L.DeviceArray.from_numpy(np.ones(4))
        with pytest.raises(RuntimeError, match="import"):
            L.DeviceCSR.import_device(n, n, 4, di, dj, dv)
        for v in (di, dj, dv):
            v.free()


# ---------------------------------------------------------------------------
# the caller's stream
def test_caller_stream():
    """Products on both paths and a batch graph queued on a caller's HIP
    stream (created, default priority, through the HIP runtime the library
    links), then the library's own stream restored."""
    hip = C.CDLL("libamdhip64.so.7")
    handle = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(handle)) == 0
    R, A, P = triple("poisson3d32")
    want = O.sp3mm_merged(R, A, P)
    scale = O.abs_product(R, A, P)
    own = L.get_stream()
    L.set_stream(handle.value)
    try:
        assert L.get_stream() == handle.value
        for path in (L.DIRECT, L.PAIR):
            got, _ = L.rap_host(R, A, P, path=path)
            check(got, want, scale)
        d = [L.DeviceCSR.upload(m) for m in (R, A, P)]
        b = L.Batch([d[0]], [d[1]], [d[2]], graph=True)
        for _ in range(3):
            b.run(True)
            out = b.result(0)
            check(out.download(), want, scale)
            out.free()
        b.free()
        for m in d:
            m.free()
    finally:
        L.set_stream(None)
        hip.hipStreamDestroy(handle)
    assert L.get_stream() == own
