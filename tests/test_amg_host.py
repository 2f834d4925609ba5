"""CPU: the V-cycle driver's control flow (amg.VCycle) with the library's
vector operations replaced by numpy stand-ins, against a plain scipy V-cycle
on the oracle's hierarchy.  This checks the order of smoothing, restriction,
coarse solve and correction, and that every correction starts from zero; the
GPU operations themselves are checked in test_gpu_consumer.py.
"""
import types

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from sp3mm4amg_amd import amg
from sp3mm4amg_amd import problems


class _Vec:
    def __init__(self, n, dtype=np.float64):
        # poisoned, like fresh device memory that nothing initialised
        self.a = np.full(n, np.nan)
        self.n = n

    @classmethod
    def from_numpy(cls, a):
        v = cls(a.size)
        v.a[:] = a
        return v

    def numpy(self):
        return self.a.copy()

    def free(self):
        pass


class _Mat:
    def __init__(self, S):
        self.S = sp.csr_matrix(S)
        self.d = types.SimpleNamespace(M=S.shape[0], N=S.shape[1])


def _fake_lib():
    def spmv(A, x, y, alpha=1.0, beta=0.0):
        y.a[:] = alpha * (A.S @ x.a) + (beta * y.a if beta != 0.0 else 0.0)

    def residual(A, b, x, r):
        r.a[:] = b.a - A.S @ x.a

    def jacobi(A, b, x, w, omega, sweeps):
        d = A.S.diagonal()
        for _ in range(sweeps):
            w.a[:] = x.a + omega * (b.a - A.S @ x.a) / d
            x.a[:] = w.a

    def axpby(alpha, x, beta, y):
        y.a[:] = (alpha * x.a if alpha != 0.0 else 0.0) + (beta * y.a if beta != 0.0 else 0.0)

    def dot(x, y):
        return float(np.dot(x.a, y.a))

    return types.SimpleNamespace(DeviceArray=_Vec, spmv=spmv, residual=residual, jacobi=jacobi, axpby=axpby,
                                 dot=dot)


def _scipy_vcycle(levels, Ac, b, nu=2, omega=2.0 / 3.0, cs=50, cycles=25):
    def jac(A, rhs, x, k):
        d = A.diagonal()
        for _ in range(k):
            x = x + omega * (rhs - A @ x) / d
        return x

    def cyc(lvl, rhs, x):
        if lvl == len(levels):
            return jac(Ac, rhs, np.zeros_like(rhs), cs)
        R, A, P = levels[lvl]
        x = jac(A, rhs, x, nu)
        rc = R @ (rhs - A @ x)
        x = x + P @ cyc(lvl + 1, rc, np.zeros_like(rc))
        return jac(A, rhs, x, nu)

    A0 = levels[0][1]
    x = np.zeros_like(b)
    hist = [1.0]
    for _ in range(cycles):
        if hist[-1] <= 1e-10:
            break
        x = cyc(0, b, x)
        hist.append(np.linalg.norm(b - A0 @ x) / np.linalg.norm(b))
    return x, hist


def _hierarchy(n):
    dims = (n, n, n)
    A = problems.poisson_operator(dims)
    levels, cur = [], A
    while dims[0] > 4:
        agg, cdims = problems.structured_aggregation(dims, (2, 2, 2))
        nagg = int(np.prod(cdims))
        R, _, P = problems.galerkin_inputs(cur, agg, nagg)
        levels.append(tuple(sp.csr_matrix((m.vals, m.ja, m.irp), shape=m.shape) for m in (R, cur, P)))
        cur = O.sp3mm_pair(R, cur, P)
        dims = cdims
    return levels, sp.csr_matrix((cur.vals, cur.ja, cur.irp), shape=cur.shape)


@pytest.mark.parametrize("n", [12, 16])
def test_vcycle_driver_matches_scipy_cycle(monkeypatch, n):
    monkeypatch.setattr(amg, "L", _fake_lib())
    levels, Ac = _hierarchy(n)
    b = np.random.default_rng(n).standard_normal(levels[0][1].shape[0])
    vc = amg.VCycle([tuple(_Mat(m) for m in lv) for lv in levels], _Mat(Ac))
    x, hist = vc.solve(b, rtol=1e-10, maxit=25)
    xw, hw = _scipy_vcycle(levels, Ac, b)
    assert len(hist) == len(hw)
    assert np.allclose(hist, hw, rtol=1e-9, atol=1e-14)
    assert np.allclose(x, xw, rtol=1e-9, atol=1e-12)
    assert hist[-1] <= 1e-10
    # a second solve reuses the work vectors and must not see the first one
    x2, hist2 = vc.solve(b, rtol=1e-10, maxit=25)
    assert hist2 == hist


def test_vcycle_rejects_mismatched_levels():
    levels, Ac = _hierarchy(8)
    R, A, P = levels[0]
    with pytest.raises(ValueError, match="do not chain"):
        amg.VCycle([(_Mat(R), _Mat(A), _Mat(P))], _Mat(A))
