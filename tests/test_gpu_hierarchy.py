"""GPU: a smoothed-aggregation hierarchy as one replayable plan (sp3mm_hier_*).

Every level (P = (I - omega D^-1 A_l) P0, R = P^T, A_{l+1} = R A_l P) against
the host construction and the oracle's triple product; replays of the
captured graph after A's values change, each against a fresh host build, with
eager library work between the replays; and the guard that refuses a replay
once A's pattern has changed.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sp3mm4amg_amd import lib as L
from sp3mm4amg_amd import problems
from sp3mm4amg_amd.csr import CSR

pytestmark = pytest.mark.gpu


def _aggregations(dims, nlev):
    aggs = []
    for _ in range(nlev):
        agg, cdims = problems.structured_aggregation(dims, (2, 2, 2))
        aggs.append((agg, int(np.prod(cdims))))
        dims = cdims
    return aggs


def _host_levels(A: CSR, aggs):
    """(R, P, A_next) per level, built on the host with the oracle's product."""
    out, cur = [], A
    for agg, nagg in aggs:
        R, _, P = problems.galerkin_inputs(cur, agg, nagg)
        nxt = O.sp3mm_pair(R, cur, P)
        out.append((R, P, nxt, O.abs_product(R, cur, P)))
        cur = nxt
    return out


def _check_levels(H, host, tol=1e-11):
    for lvl, (R, P, Ac, scale) in enumerate(host):
        dR, dP, dAc = H.level(lvl)
        for got, want in ((dP.download(), P), (dR.download(), R)):
            assert np.array_equal(got.irp, want.irp) and np.array_equal(got.ja, want.ja), f"level {lvl} pattern"
            assert np.max(np.abs(got.vals - want.vals), initial=0.0) <= 1e-13, f"level {lvl} values"
        got = dAc.download()
        got.check()
        assert np.array_equal(got.irp, Ac.irp) and np.array_equal(got.ja, Ac.ja), f"level {lvl} RAP pattern"
        err = O.rel_errors(got, Ac, scale)
        assert err.max(initial=0.0) <= tol, f"level {lvl}: max rel err {err.max()}"


@pytest.mark.parametrize("graph", [True, False])
def test_hierarchy_plan_levels(graph):
    """Three levels from the 16^3 7-point operator, against the host build."""
    L.init(0)
    dims = (16, 16, 16)
    A = problems.poisson_operator(dims)
    aggs = _aggregations(dims, 3)
    dA = L.DeviceCSR.upload(A)
    H = L.Hierarchy(dA, aggs, graph=graph)
    info = H.info()
    print(f"graph={graph} levels={info['levels']} build {info['ms_build']:.2f} ms")
    assert info["levels"] == 3 and info["graph"] == graph
    host = _host_levels(A, aggs)
    _check_levels(H, host)
    assert H.coarse.shape == host[-1][2].shape
    # a replay over unchanged values reproduces the same levels
    ms = H.update(graph=graph)
    print(f"replay {ms:.3f} ms")
    _check_levels(H, host)
    H.free()
    dA.free()


def test_hierarchy_plan_replays_new_values():
    """A's values change in place between replays (diagonal scaled, off-diagonal
    entries perturbed); each graph replay matches a fresh host build, with eager
    library work (a device hierarchy built call by call) between replays."""
    L.init(0)
    dims = (16, 16, 12)
    A = problems.poisson_operator(dims)
    aggs = _aggregations(dims, 2)
    dA = L.DeviceCSR.upload(A)
    H = L.Hierarchy(dA, aggs, graph=True)
    rng = np.random.default_rng(11)
    rows = A.row_ids()
    diag = rows == A.ja
    for rep in range(3):
        vals = A.vals.copy()
        vals[diag] *= 1.0 + 0.25 * (rep + 1)
        vals[~diag] *= 1.0 + 0.1 * rng.random(int((~diag).sum()))
        A2 = CSR(A.M, A.N, A.irp, A.ja, vals)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        assert L.lib().sp3mm_copy_to_device(dA.d.val, vals.ctypes.data, vals.nbytes) == 0
        ms = H.update(graph=True)
        print(f"replay {rep}: {ms:.3f} ms")
        _check_levels(H, _host_levels(A2, aggs))
        # eager work on the same stream between replays
        levels, coarse = L.hierarchy_device(dA, aggs)
        got, want = coarse.download(), H.coarse.download()
        assert np.array_equal(got.ja, want.ja)
        assert np.max(np.abs(got.vals - want.vals)) <= 1e-12 * np.max(np.abs(want.vals))
        for dR, _, dP in levels:
            dR.free()
            dP.free()
        for dR_, dAl, _ in levels[1:]:
            dAl.free()
        coarse.free()
    H.free()
    dA.free()


def test_hierarchy_plan_refuses_changed_pattern():
    """A column index of A changed on the device (the CSR stays valid): the
    replay is refused before any kernel of the plan runs; restoring it makes
    replays work again."""
    L.init(0)
    dims = (8, 8, 8)
    A = problems.poisson_operator(dims)
    aggs = _aggregations(dims, 2)
    dA = L.DeviceCSR.upload(A)
    H = L.Hierarchy(dA, aggs, graph=True)
    # row 0 holds columns 0, 1, 8, 64: move the second to column 2 (still sorted)
    assert list(A.ja[:4]) == [0, 1, 8, 64]
    addr = C.cast(dA.d.col, C.c_void_p).value + 4
    moved, orig = np.array([2], np.int32), np.array([1], np.int32)
    assert L.lib().sp3mm_copy_to_device(addr, moved.ctypes.data, 4) == 0
    with pytest.raises(RuntimeError, match="pattern"):
        H.update()
    assert L.lib().sp3mm_copy_to_device(addr, orig.ctypes.data, 4) == 0
    H.update()
    _check_levels(H, _host_levels(A, aggs))
    H.free()
    dA.free()
