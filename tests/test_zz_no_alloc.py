"""The allocation-free entry point sp3mm_rap_noalloc through the C ABI.

It computes the direct FP64 product of sp3mm_rap into caller-owned device
arrays, with every scratch buffer carved from the library's fixed device
scratch array.  Checked against sp3mm_rap (the entry point it is a variant of)
and the CPU oracle on two calls in a row, the second with a smaller input
(the array keeps the first call's contents), then its two refusals: output
arrays shorter than the product (the needed capacity is reported) and a
product whose scratch exceeds the fixed array.

Bar: row offsets and sorted columns bit-exact; FP64 values within 1e-12
relative per nonzero (stencil workloads, no cancellation below 1e-2 of the
magnitude scale).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sp3mm4amg_amd import lib as L
from sp3mm4amg_amd import problems
from sp3mm4amg_amd.csr import CSR

pytestmark = pytest.mark.gpu
TOL = 1e-12
SCRATCH_BYTES = 256 << 20


@pytest.fixture(scope="module", autouse=True)
def gpu_lib():
    L.init(0)
    yield


def _hip_free_bytes() -> int:
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _noalloc(dR, dA, dP, cap):
    """Caller-owned output arrays, the call, and the result as a host CSR."""
    M = dR.d.M
    rp = L.DeviceArray(M + 1, np.int64)
    col = L.DeviceArray(max(cap, 1), np.int32)
    val = L.DeviceArray(max(cap, 1), np.float64)
    try:
        free0 = _hip_free_bytes()
        nnz, st = L.rap_noalloc(dR, dA, dP, rp, col, val, cap)
        free1 = _hip_free_bytes()
        irp = rp.numpy()
        out = CSR(M, dP.d.N, irp, col.numpy()[:nnz].astype(np.int64), val.numpy()[:nnz])
    finally:
        for a in (rp, col, val):
            a.free()
    return out, nnz, st, free0 - free1


def _check(got: CSR, want: CSR, scale: CSR):
    got.check()
    assert got.shape == want.shape
    assert np.array_equal(got.irp, want.irp), "row offsets differ"
    assert np.array_equal(got.ja, want.ja), "column indices differ"
    err = O.rel_errors(got, want, scale)
    assert err.max(initial=0.0) <= TOL, f"max rel err {err.max()}"


def test_noalloc_matches_rap_twice_then_smaller():
    # first a 3D product (4,096 coarse rows, scratch slots of several MB),
    # then a smaller 2D one on the same (now dirty) scratch array
    cases = [("poisson3d_sa(32)", problems.poisson3d_sa(32)), ("poisson2d_sa(64)", problems.poisson2d_sa(64))]
    sizes = []
    for name, p in cases:
        R, A, P = p["R"], p["A"], p["P"]
        print(f"rap_noalloc on {name}: R {R.shape}", flush=True)
        dR, dA, dP = (L.DeviceCSR.upload(m) for m in (R, A, P))
        try:
            ref_out, _ = L.rap(dR, dA, dP, path=L.DIRECT)
            ref = ref_out.download()
            ref_out.free()
            got, nnz, st, dropped = _noalloc(dR, dA, dP, ref.nnz + 7)
        finally:
            for m in (dR, dA, dP):
                m.free()
        assert nnz == ref.nnz
        assert st["nnz_out"] == nnz
        scale = O.abs_product(R, A, P)
        _check(got, ref, scale)                       # the existing entry point
        _check(got, O.sp3mm_merged(R, A, P), scale)   # and the oracle
        # the library allocated nothing during the call (a pool or hipMalloc
        # buffer for the scratch slots alone would be several MB here)
        assert dropped < (2 << 20), f"free device memory dropped by {dropped} bytes"
        sizes.append(R.M)
    assert sizes[1] < sizes[0]


def test_noalloc_reports_short_output():
    p = problems.poisson2d_sa(32)
    R, A, P = p["R"], p["A"], p["P"]
    want = O.sp3mm_merged(R, A, P)
    dR, dA, dP = (L.DeviceCSR.upload(m) for m in (R, A, P))
    try:
        with pytest.raises(RuntimeError, match="output capacity") as ei:
            _noalloc(dR, dA, dP, want.nnz - 1)
        assert ei.value.needed == want.nnz
        got, nnz, _, _ = _noalloc(dR, dA, dP, ei.value.needed)
    finally:
        for m in (dR, dA, dP):
            m.free()
    assert nnz == want.nnz
    _check(got, want, O.abs_product(R, A, P))


def test_noalloc_refuses_products_beyond_the_fixed_scratch():
    # 128^3 7-point SA under a schedule with stage-2 tables as large as the
    # stage-1 ones (256 slots): 262,144 coarse rows then need about 0.8 GB of
    # scratch slots, more than the fixed 256 MiB array
    p = problems.poisson3d_sa(128)
    R, A, P = p["R"], p["A"], p["P"]
    print(f"rap_noalloc on poisson3d_sa(128), h2_div=1: R {R.shape}", flush=True)
    dR, dA, dP = (L.DeviceCSR.upload(m) for m in (R, A, P))
    try:
        L.set_tuning({"h1_shift": 0, "h2_div": 1})
        try:
            with pytest.raises(RuntimeError, match="fixed device scratch"):
                _noalloc(dR, dA, dP, 9_000_000)
        finally:
            L.set_tuning(None)
        # the library is still usable afterwards, through both entry points
        q = problems.poisson2d_sa(16)
        small = [L.DeviceCSR.upload(m) for m in (q["R"], q["A"], q["P"])]
        try:
            want = O.sp3mm_merged(q["R"], q["A"], q["P"])
            scale = O.abs_product(q["R"], q["A"], q["P"])
            got, nnz, _, _ = _noalloc(*small, want.nnz)
            _check(got, want, scale)
            out, _ = L.rap(*small, path=L.DIRECT)
            _check(out.download(), want, scale)
            out.free()
        finally:
            for m in small:
                m.free()
    finally:
        for m in (dR, dA, dP):
            m.free()
    assert SCRATCH_BYTES == 256 << 20
