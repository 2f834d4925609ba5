"""CPU baseline driver: the reference library compiled from its own sources
(oracle/_ref/libsp3mm_ref.so), or the oracle's restatement where that build
is absent, timed on row samples of a workload.

TEST INFRASTRUCTURE ONLY: run by bench.py's cpu_baseline leg, in a process of
its own so that libgomp reads the OpenMP environment the reference's timing
protocol asks for (OMP_NUM_THREADS, OMP_PROC_BIND=close, OMP_PLACES=cores;
OMP_SCHEDULE left at libgomp's default, adapted per call by the driver's
chunksFairFolded as code/test/Sp3MM_test.c:80, :304-308 does).

  python -m oracle.cpu_bench INPUTS.npz BUDGET_S [ITERS]

INPUTS.npz holds products 0..n-1 as R{i}_*, A{i}_*, P{i}_* (shape, irp, ja,
vals) and a `name`.  Each form gets BUDGET_S / (number of forms) seconds: a
probe on 1/256 of every product's rows sizes a row fraction f, then the form
runs 1 untimed warm-up and ITERS timed runs (default 3) of R[rows 0..f*M]*A*P
for every product, timed with the wall clock around each call (the
reference's own omp_get_wtime bracket, Sp3MM_test.c:120-125).  Prints one
JSON object: per form the median, minimum and maximum time and GFLOP/s at the
median, FLOP = 2 * (multiply-adds of R*A + of (R*A)*P) of the sample.
"""
from __future__ import annotations

import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import ref  # noqa: E402
from sp3mm4amg_amd.csr import CSR  # noqa: E402


def load(path):
    z = np.load(path)
    n = int(z["n"])
    out = []
    for i in range(n):
        ms = []
        for k in ("R", "A", "P"):
            M, N = (int(x) for x in z[f"{k}{i}_shape"])
            ms.append(CSR(M, N, z[f"{k}{i}_irp"], z[f"{k}{i}_ja"].astype(np.int64), z[f"{k}{i}_vals"]))
        out.append(tuple(ms))
    return str(z["name"]), out


def rows(R: CSR, frac: float) -> CSR:
    k = max(1, min(R.M, int(round(R.M * frac))))
    return CSR(k, R.N, R.irp[:k + 1], R.ja[:R.irp[k]], R.vals[:R.irp[k]])


def flop(Rs, A, P):
    mac1 = int(O.ub(Rs, A)[-1])
    mac2 = int(O.ub(O.spmm(Rs, A), P)[-1])
    return 2 * (mac1 + mac2)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    path, budget = sys.argv[1], float(sys.argv[2])
    iters = int(sys.argv[3]) if len(sys.argv) > 3 else 3
    name, products = load(path)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    kind = "reference" if ref.available() else "port"
    if kind == "reference":
        folded = ref.chunks_fair_folded() is not None
        forms = [("sp3mmRowByRowPair_0 + spmmRowByRow2DBlocks_0 (20 x 2 grid)",
                  lambda R, A, P: ref.sp3mm_pair(R, A, P, spmm="spmmRowByRow2DBlocks_", threads=threads,
                                                 grid=(20, 2), fair_folded=folded)),
                 ("sp3mmRowByRowPair_0 + spmmRowByRow_SymbNum_0 (IDXMAP)",
                  lambda R, A, P: ref.sp3mm_pair(R, A, P, spmm="spmmRowByRow_SymbNum_", threads=threads,
                                                 fair_folded=folded)),
                 ("sp3mmRowByRowMerged_0",
                  lambda R, A, P: ref.sp3mm_merged(R, A, P, threads=threads, fair_folded=folded))]
    else:
        folded = False

        def port(R, A, P):
            O.set_threads(threads)
            O.sp3mm_pair(R, A, P, faithful=True)
        forms = [("oracle restatement of sp3mmRowByRowPair_0 + spmmRowByRow_SymbNum_0", port)]
    each = budget / len(forms)
    results = []
    for fname, fn in forms:
        def run(frac):
            t = 0.0
            for (R, A, P) in products:
                S = rows(R, frac)
                t0 = time.perf_counter()
                fn(S, A, P)
                t += time.perf_counter() - t0
            return t
        frac = 1.0
        if budget > 0:
            probe = 1 / 256
            t = run(probe)
            frac = min(1.0, probe * (each / (iters + 1)) / max(t, 1e-4))
        run(frac)  # warm-up
        ts = [run(frac) for _ in range(iters)]
        f = sum(flop(rows(R, frac), A, P) for (R, A, P) in products)
        med = float(np.median(ts))
        results.append({"form": fname, "value": round(f / med / 1e9, 4), "unit": "GFLOP/s",
                        "median_s": round(med, 4), "min_s": round(min(ts), 4), "max_s": round(max(ts), 4),
                        "runs": iters, "flop": f,
                        "sample": f"rows 0..{max(1, int(round(products[0][0].M * frac))) - 1} of R "
                                  f"({frac:.3%} of every product's rows, {len(products)} product(s))"})
    best = max(results, key=lambda r: r["value"])
    print(json.dumps({"workload": name, "kind": kind, "cores": threads, "cpu_model": cpu_model(),
                      "nproc": os.cpu_count(),
                      "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES",
                                                             "OMP_SCHEDULE")},
                      "chunk_distribution": "chunksFairFolded" if folded else "none (schedule left alone)",
                      "value": best["value"], "unit": "GFLOP/s", "best_form": best["form"],
                      "forms": results}), flush=True)


if __name__ == "__main__":
    main()
