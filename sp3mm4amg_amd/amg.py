"""V-cycles over a device hierarchy: the consumers of the coarse operators.

The reference library stops at the product R*A*P (its AMG4PSBLAS caller owns
the solver).  A V-cycle that runs on the operators and transfer matrices
this library built is the check that they are usable as an AMG hierarchy:
damped Jacobi smoothing, restriction with R, prolongation with P, and a
coarsest level smoothed to convergence.  Every vector operation is one call
of the library's C ABI (SpMV, residual, Jacobi, axpby, dot); the vectors stay
in HBM between calls.  GraphVCycle runs the same cycle natively
(sp3mm_vcycle_*), captured once as a HIP graph and replayed per cycle.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import lib as L


class VCycle:
    """V(nu1, nu2) cycle with damped Jacobi smoothing.

    `levels` is a list of (R, A, P) device matrices from the finest level
    down (as returned by lib.hierarchy_device), `coarse` the coarsest
    operator, on which `coarse_sweeps` Jacobi sweeps from zero stand in for a
    direct solve.
    """

    def __init__(self, levels, coarse: L.DeviceCSR, nu1: int = 2, nu2: int = 2, omega: float = 2.0 / 3.0,
                 coarse_sweeps: int = 50):
        if not levels:
            raise ValueError("a V-cycle needs at least one level above the coarsest")
        for i, (R, A, P) in enumerate(levels):
            nxt = levels[i + 1][1] if i + 1 < len(levels) else coarse
            m, n = A.d.M, A.d.N
            if m != n or P.d.M != m or R.d.N != m or R.d.M != P.d.N or nxt.d.M != R.d.M or nxt.d.N != R.d.M:
                raise ValueError(f"level {i}: operator and transfer shapes do not chain")
        self.levels, self.coarse = levels, coarse
        self.nu1, self.nu2, self.omega, self.coarse_sweeps = nu1, nu2, omega, coarse_sweeps
        ops = [A for _, A, _ in levels] + [coarse]
        # per level: solution, right-hand side, residual, Jacobi work vector
        self.x = [L.DeviceArray(A.d.M) for A in ops]
        self.b = [L.DeviceArray(A.d.M) for A in ops]
        self.r = [L.DeviceArray(A.d.M) for A in ops]
        self.w = [L.DeviceArray(A.d.M) for A in ops]

    def cycle(self, lvl: int = 0) -> None:
        """One V-cycle on level `lvl` for b[lvl], updating x[lvl] in place."""
        x, b = self.x[lvl], self.b[lvl]
        if lvl > 0:
            L.axpby(0.0, None, 0.0, x)                    # a correction starts at zero
        if lvl == len(self.levels):
            L.jacobi(self.coarse, b, x, self.w[lvl], self.omega, self.coarse_sweeps)
            return
        R, A, P = self.levels[lvl]
        L.jacobi(A, b, x, self.w[lvl], self.omega, self.nu1)
        L.residual(A, b, x, self.r[lvl])
        L.spmv(R, self.r[lvl], self.b[lvl + 1])          # restrict
        self.cycle(lvl + 1)
        L.spmv(P, self.x[lvl + 1], x, 1.0, 1.0)           # prolongate and correct
        L.jacobi(A, b, x, self.w[lvl], self.omega, self.nu2)

    def residual_norm(self) -> float:
        A = self.levels[0][1]
        L.residual(A, self.b[0], self.x[0], self.r[0])
        return math.sqrt(L.dot(self.r[0], self.r[0]))

    def solve(self, b: np.ndarray, x0: np.ndarray | None = None, rtol: float = 1e-10, maxit: int = 50):
        """Cycles until ||b - A x|| <= rtol ||b||; returns (x, history of
        relative residual norms, one per cycle, starting with the initial one)."""
        n = self.levels[0][1].d.M
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (n,):
            raise ValueError("right-hand side length does not match the finest operator")
        self.b[0].free()
        self.b[0] = L.DeviceArray.from_numpy(b)
        if x0 is None:
            L.axpby(0.0, None, 0.0, self.x[0])
        else:
            self.x[0].free()
            self.x[0] = L.DeviceArray.from_numpy(np.asarray(x0, dtype=np.float64))
        bn = math.sqrt(L.dot(self.b[0], self.b[0])) or 1.0
        hist = [self.residual_norm() / bn]
        for _ in range(maxit):
            if hist[-1] <= rtol:
                break
            self.cycle(0)
            hist.append(self.residual_norm() / bn)
        return self.x[0].numpy(), hist

    def free(self) -> None:
        for vs in (self.x, self.b, self.r, self.w):
            for v in vs:
                v.free()


class GraphVCycle:
    """The V-cycle of VCycle, run by the library: one cycle is a HIP graph
    (use_graph) replayed until the residual target; the same kernels in the
    same order as VCycle, so both give the same bits."""

    def __init__(self, levels, coarse: L.DeviceCSR, nu1: int = 2, nu2: int = 2, omega: float = 2.0 / 3.0,
                 coarse_sweeps: int = 50, use_graph: bool = True):
        n = len(levels)
        arr = L.DMat * max(n, 1)
        self._R = arr(*[R.d for R, _, _ in levels])
        self._A = arr(*[A.d for _, A, _ in levels])
        self._P = arr(*[P.d for _, _, P in levels])
        self._coarse = coarse.d
        self.n0 = levels[0][1].d.M if levels else 0
        h = C.c_void_p()
        L._check(L.lib().sp3mm_vcycle_create(n, self._R, self._A, self._P, C.byref(self._coarse), nu1, nu2,
                                             float(omega), coarse_sweeps, int(use_graph), C.byref(h)),
                 "sp3mm_vcycle_create")
        self.h = h

    def solve(self, b, x0=None, rtol: float = 1e-10, maxit: int = 50):
        """b, x0: numpy vectors or DeviceArrays.  Returns (x as numpy, history
        of relative residual norms, device ms of the cycles)."""
        own = []
        if not isinstance(b, L.DeviceArray):
            b = L.DeviceArray.from_numpy(np.asarray(b, dtype=np.float64))
            own.append(b)
        if x0 is None:
            x = L.DeviceArray(self.n0)
            L.axpby(0.0, None, 0.0, x)
            own.append(x)
        elif not isinstance(x0, L.DeviceArray):
            x = L.DeviceArray.from_numpy(np.asarray(x0, dtype=np.float64))
            own.append(x)
        else:
            x = x0
        if b.n != self.n0 or x.n != self.n0:
            raise ValueError("vector lengths do not match the finest operator")
        hist = np.zeros(maxit + 1)
        k, ms = C.c_int32(), C.c_double()
        try:
            L._check(L.lib().sp3mm_vcycle_solve(self.h, b.ptr, x.ptr, float(rtol), int(maxit), C.byref(k),
                                                hist.ctypes.data, C.byref(ms)), "sp3mm_vcycle_solve")
            out = x.numpy()
        finally:
            for v in own:
                v.free()
        return out, list(hist[:k.value + 1]), ms.value

    def free(self) -> None:
        if self.h:
            L.lib().sp3mm_vcycle_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
