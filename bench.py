"""Benchmark: Galerkin triple product R*A*P on one MI355X.

Workload (default): 3D 7-point Poisson operator on a 128^3 grid (2,097,152
rows), 2x2x2 smoothed aggregation to 262,144 coarse rows, FP64 values, int32
column indices; synthetic inputs from sp3mm4amg_amd.problems (deterministic).
A step is one complete product R*A*P on device-resident inputs: upper bounds,
binning, symbolic, scan, output allocation from the pool and numeric phase.

Prints ONE JSON line (rank 0): throughput in GFLOP/s (FLOP = 2 x multiply-adds
of the pair form, the reference's flop count), ms per product, the HBM
roofline of the dominant kernel and the reference CPU path timed on a bounded
sample on this host's cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--path direct|pair]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def committed_traffic(path=os.path.join(ROOT, "profiles", "pmc_direct_poisson3d128.md")):
    """HBM bytes per product of the main direct kernel(s), from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated with a copy kernel)."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def alg_bytes(R, A, P, mac1, mac2, nnz_ra, nnz_out, vbytes=8):
    """Row-gather traffic of the product (stream R, gather A rows by R, gather P
    rows by RA, write RAP), 4-byte column indices; row offsets of 4 bytes, or
    8 for a matrix with 2^31 or more entries."""
    i = 4
    po = lambda nnz: 8 if nnz >= 2 ** 31 else 4
    m = R.M
    return (R.nnz * (vbytes + i) + (m + 1) * po(R.nnz)
            + 2 * po(A.nnz) * R.nnz + (vbytes + i) * mac1
            + 2 * po(P.nnz) * nnz_ra + (vbytes + i) * mac2
            + nnz_out * (vbytes + i) + (m + 1) * po(nnz_out))


def compulsory_bytes(R, A, P, nnz_out, vbytes=8):
    """Each operand read once and the product written once (values, 4-byte
    columns, row offsets)."""
    po = lambda nnz: 8 if nnz >= 2 ** 31 else 4
    return sum(m.nnz * (vbytes + 4) + (m.M + 1) * po(m.nnz) for m in (R, A, P)) + \
        nnz_out * (vbytes + 4) + (R.M + 1) * po(nnz_out)


def _save_products(path, name, products):
    import numpy as np
    arrs = {"n": len(products), "name": name}
    for i, trio in enumerate(products):
        for k, m in zip(("R", "A", "P"), trio):
            arrs[f"{k}{i}_shape"] = np.array([m.M, m.N])
            arrs[f"{k}{i}_irp"] = m.irp
            arrs[f"{k}{i}_ja"] = m.ja.astype(np.int32)
            arrs[f"{k}{i}_vals"] = m.vals
    np.savez(path, **arrs)


def cpu_baseline(workloads):
    """The reference CPU path timed on this host's cores, one workload at a
    time, by oracle/cpu_bench.py in a process of its own (the reference
    library compiled from its sources, oracle/_ref, in its three forms; the
    oracle's restatement where that build is absent), under the reference's
    timing protocol: OMP_PROC_BIND=close, OMP_PLACES=cores, OMP_NUM_THREADS
    threads, the default OpenMP schedule adapted by chunksFairFolded, one
    untimed warm-up and `iters` timed runs per form on a bounded row sample.
    `workloads` is a list of (key, name, products, budget_s, iters); the
    first is the headline workload, whose fastest form gives `value`."""
    import subprocess
    import tempfile
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for key, name, products, budget, iters in workloads:
            path = os.path.join(d, f"{key}.npz")
            _save_products(path, name, products)
            log(f"cpu baseline: {name} ({len(products)} product(s), budget {budget:.0f} s, {iters} timed runs)")
            r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", path, str(budget), str(iters)],
                               cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"cpu baseline on {name} failed: {r.stderr[-2000:]}")
            out[key] = json.loads(r.stdout.strip().splitlines()[-1])
            os.remove(path)
    head = out[workloads[0][0]]
    best = max(head["forms"], key=lambda f: f["value"])
    return {"value": head["value"], "unit": "GFLOP/s", "cores": head["cores"], "kind": head["kind"],
            "sample": f"{best['form']}: {best['sample']}, median of {best['runs']} runs "
                      f"({best['min_s']}-{best['max_s']} s)",
            "cpu_model": head["cpu_model"], "nproc": head["nproc"], "omp": head["omp"],
            "chunk_distribution": head["chunk_distribution"], "forms": head["forms"],
            "configs": {k: {"workload": v["workload"], "value": v["value"], "best_form": v["best_form"],
                            "forms": v["forms"]} for k, v in out.items() if k != workloads[0][0]}}


def time_form(lib, dR, dA, dP, steps, warmup, **kw):
    """warmup + steps products; returns (wall ms per product, mean main-kernel ms,
    median device ms, last stats, device ms of every product)."""
    import numpy as np
    for _ in range(warmup):
        out, st = lib.rap(dR, dA, dP, **kw)
        out.free()
    lib.lib().sp3mm_sync()
    total_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        out, st = lib.rap(dR, dA, dP, raw_stats=True, **kw)
        total_ms.append(st.ms_total)
        out.free()
    lib.lib().sp3mm_sync()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    st = st.as_dict()
    main_ms = main_kernel_ms(lib, dR, dA, dP, steps, **kw)
    return wall, float(np.mean(main_ms)), float(np.median(total_ms)), st, total_ms


def main_kernel_ms(lib, dR, dA, dP, steps, **kw):
    """Device time of the main kernel(s) of a product, from per-phase HIP
    events (sp3mm_opts.timing); measured on separate products, since the events
    add gaps to the launch sequence."""
    ms = []
    for _ in range(steps):
        out, st = lib.rap(dR, dA, dP, timing=True, **kw)
        ms.append(st["ms_main"])
        out.free()
    return ms


def spread(xs):
    """median, minimum, 90th percentile and maximum of a list of timings"""
    import numpy as np
    a = np.asarray(xs, dtype=float)
    return {"median": round(float(np.median(a)), 4), "min": round(float(a.min()), 4),
            "p90": round(float(np.percentile(a, 90)), 4), "max": round(float(a.max()), 4), "n": int(a.size)}


def mixed_leg(lib, problems, steps, warmup, tune, n=100):
    """Mixed precision on the 27-point random stencil (100^3, 3x3x3 aggregates):
    FP32 operands with FP32-pair accumulation (MIXED) and with FP64 accumulation
    (MIXED64), and FP64 operands, each against the CPU oracle of the same
    FP32-rounded operands accumulated in long double (oracle.sp3mm_pair_ld)."""
    import numpy as np
    from oracle import oracle as O
    prob = problems.stencil27_sa(n, seed=3)
    R, A, P = prob["R"], prob["A"], prob["P"]
    r32 = [m.with_vals(m.vals.astype(np.float32).astype(np.float64)) for m in (R, A, P)]
    d64 = [lib.DeviceCSR.upload(m) for m in r32]
    d32 = [lib.DeviceCSR.upload(m, vtype=lib.F32) for m in r32]
    out, st_pair = lib.rap(*d64, path=lib.PAIR)
    out.free()
    mac1, mac2, nnz_ra = st_pair["mac1"], st_pair["mac2"], st_pair["nnz_mid"]
    flop = 2 * (mac1 + mac2)
    a = O.sp3mm_pair_ld(*r32)
    rec = {"workload": f"{prob['name']} Galerkin R*A*P, direct path", "flop": flop, "mac1": mac1,
           "mac2": mac2, "nnz_RAP": a.nnz,
           "error_reference": "CPU oracle on the same FP32-rounded operands, long-double accumulation"}
    for name, d, prec, vb in (("fp64", d64, lib.FP64, 8), ("mixed", d32, lib.MIXED, 4),
                              ("mixed64", d32, lib.MIXED64, 4)):
        lib.set_tuning(None)
        sched = None
        if tune:
            sched, _, _ = lib.autotune(*d, precision=prec, reps=7)
        wall, main, dev, st, devs = time_form(lib, *d, steps, warmup, path=lib.DIRECT, precision=prec)
        got, _ = lib.rap(*d, path=lib.DIRECT, precision=prec)
        b = got.download()
        got.free()
        same = np.array_equal(a.irp, b.irp) and np.array_equal(a.ja, b.ja)
        if not same:
            raise RuntimeError(f"mixed leg: the {name} pattern differs from the oracle's")
        rel = np.abs(b.vals - a.vals) / np.maximum(np.abs(a.vals), 1e-300)
        rec[name] = {"ms_per_product": round(wall, 4), "device_ms_median": round(dev, 4),
                     "device_ms": spread(devs),
                     "main_kernel_ms": round(main, 4), "gflops": round(flop / wall / 1e6, 2),
                     "bytes_alg": alg_bytes(R, A, P, mac1, mac2, nnz_ra, a.nnz, vb), "schedule": sched,
                     "pattern_equal_oracle": bool(same),
                     "max_rel_err_vs_oracle": float(rel.max()) if rel.size else 0.0,
                     "median_rel_err_vs_oracle": float(np.median(rel)) if rel.size else 0.0}
        if name == "mixed":
            rec[name]["fp64_rows"] = int(st["guarded_rows"])
    lib.set_tuning(None)
    for m in d64 + d32:
        m.free()
    rec["speedup_mixed_vs_fp64"] = round(rec["fp64"]["ms_per_product"] / rec["mixed"]["ms_per_product"], 3)
    rec["speedup_mixed64_vs_fp64"] = round(rec["fp64"]["ms_per_product"] / rec["mixed64"]["ms_per_product"], 3)
    return rec, (R, A, P)


def batch_leg(lib, problems, reps, tune=True):
    """Config of many uneven products: a 5-level smoothed-aggregation hierarchy
    from the 96^3 7-point operator (coarse operators made by this library) plus
    256 independent 2D 5-point problems with sides U{16..256}.  The batched form
    (one block-diagonal product replayed as one HIP graph) is timed against a
    loop of per-product calls, in FP64 and in mixed precision."""
    import numpy as np
    from oracle import oracle as O
    t0 = time.perf_counter()
    hier = problems.hierarchy_3d(96, 5, rap=lambda R, A, P: lib.rap_host(R, A, P)[0])
    b2d = problems.poisson2d_batch(256, seed=5)
    items = [(h["R"], h["A"], h["P"]) for h in hier] + [(p["R"], p["A"], p["P"]) for p in b2d]
    log(f"batch inputs: {len(items)} products generated in {time.perf_counter() - t0:.1f}s")
    rec = {"workload": "5-level SA hierarchy from 96^3 7-point (2x2x2) + 256 2D 5-point problems "
                       "(sides U{16..256}, PCG64 seed 5, 2x2 SA), direct path",
           "products": len(items)}
    flop = 0
    b_alg = b_comp = 0
    lib.set_tuning(None)
    for form, vt, prec in (("fp64", lib.F64, lib.FP64), ("mixed", lib.F32, lib.MIXED),
                           ("mixed64", lib.F32, lib.MIXED64)):
        mats = []
        for (R, A, P) in items:
            if vt == lib.F32:
                R, A, P = (m.with_vals(m.vals.astype(np.float32).astype(np.float64)) for m in (R, A, P))
            mats.append(tuple(lib.DeviceCSR.upload(m, vtype=vt) for m in (R, A, P)))
        kw = {"path": lib.DIRECT, "precision": prec}
        if form == "fp64":
            for (R, A, P), (dR, dA, dP) in zip(items, mats):
                out, st = lib.rap(dR, dA, dP, path=lib.PAIR)
                out.free()
                flop += 2 * (st["mac1"] + st["mac2"])
                b_alg += alg_bytes(R, A, P, st["mac1"], st["mac2"], st["nnz_mid"], st["nnz_out"])
                b_comp += compulsory_bytes(R, A, P, st["nnz_out"])
        # batched: one packed product captured as a graph; with `tune`, the
        # schedule is tuned on the packed product (untimed) and installed, so
        # the per-product loop below runs under the same schedule
        lib.set_tuning(None)
        b = lib.Batch([m[0] for m in mats], [m[1] for m in mats], [m[2] for m in mats], graph=True,
                      precision=prec)
        schedule = b.autotune(reps=5)[0] if tune else None
        # per-product loop
        for (dR, dA, dP) in mats:
            out, _ = lib.rap(dR, dA, dP, **kw)
            out.free()
        lib.lib().sp3mm_sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            for (dR, dA, dP) in mats:
                out, _ = lib.rap(dR, dA, dP, **kw)
                out.free()
        lib.lib().sp3mm_sync()
        loop_ms = (time.perf_counter() - t0) * 1e3 / reps
        b.run(True)
        dev = []
        t0 = time.perf_counter()
        for _ in range(reps):
            dev.append(b.run(True))
        graph_ms = (time.perf_counter() - t0) * 1e3 / reps
        # every batched product against the CPU oracle of its operands (FP64
        # sums for the FP64 form; long-double sums of the FP32-rounded
        # operands for the mixed forms)
        worst = 0.0
        for i, (R, A, P) in enumerate(items):
            if vt == lib.F32:
                R, A, P = (m.with_vals(m.vals.astype(np.float32).astype(np.float64)) for m in (R, A, P))
                want = O.sp3mm_pair_ld(R, A, P)
            else:
                want = O.sp3mm_pair(R, A, P)
            got = b.result(i)
            g = got.download()
            got.free()
            if not (np.array_equal(want.irp, g.irp) and np.array_equal(want.ja, g.ja)):
                raise RuntimeError(f"batched result {i} ({form}) differs in pattern from the oracle")
            if want.nnz:
                worst = max(worst, float(np.max(np.abs(want.vals - g.vals)
                                                / np.maximum(np.abs(want.vals), 1e-300))))
        info = b.info()
        b.free()
        for m in mats:
            for d in m:
                d.free()
        rec[form] = {"schedule": schedule,
                     "loop_ms_per_batch": round(loop_ms, 4), "graph_ms_per_batch": round(graph_ms, 4),
                     "graph_device_ms_median": round(float(np.median(dev)), 4),
                     "graph_device_ms": spread(dev),
                     "products_checked_against_oracle": len(items),
                     "max_rel_err_vs_oracle": worst,
                     "speedup_graph_vs_loop": round(loop_ms / graph_ms, 3),
                     "gflops_graph": round(flop / graph_ms / 1e6, 2),
                     "graph_build_ms": round(info["ms_build"], 3)}
    rec["flop"] = flop
    rec["bytes_alg_fp64"] = b_alg
    rec["bytes_compulsory_fp64"] = b_comp
    g = rec["fp64"]["graph_device_ms_median"] * 1e-3
    rec["fp64"]["roofline_frac"] = round(b_alg / g / 1e9 / HBM_PEAK_GBS, 4)
    rec["fp64"]["roofline_frac_compulsory"] = round(b_comp / g / 1e9 / HBM_PEAK_GBS, 4)
    # the same 5-level hierarchy built entirely on the device: prolongator,
    # transpose and R*A*P per level (the operators above were built on the host)
    dims = (96, 96, 96)
    aggs = []
    for _ in range(5):
        agg, cdims = problems.structured_aggregation(dims, (2, 2, 2))
        aggs.append((agg, int(np.prod(cdims))))
        dims = cdims
    dA0 = lib.DeviceCSR.upload(hier[0]["A"])
    ms = []
    for rep in range(3):
        t0 = time.perf_counter()
        levels, coarse = lib.hierarchy_device(dA0, aggs)
        lib.lib().sp3mm_sync()
        ms.append((time.perf_counter() - t0) * 1e3)
        for i, (dR, dAl, dP) in enumerate(levels):
            dR.free()
            dP.free()
            if i > 0:
                dAl.free()
        coarse.free()
    rec["hierarchy_setup_on_device_ms"] = round(float(np.median(ms)), 3)
    # the same hierarchy as one plan: built once, then every level recomputed
    # from A's values by one HIP graph replay (numeric re-setup, same pattern)
    t0 = time.perf_counter()
    H = lib.Hierarchy(dA0, aggs, graph=True)
    rec["hierarchy_plan_create_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    rep_ms = [H.update(graph=True) for _ in range(5)]
    rec["hierarchy_setup_graph_replay_ms"] = round(float(np.median(rep_ms)), 3)
    rec["speedup_hierarchy_replay_vs_call_by_call"] = round(rec["hierarchy_setup_on_device_ms"] /
                                                            rec["hierarchy_setup_graph_replay_ms"], 2)
    H.free()
    dA0.free()
    for form in ("fp64", "mixed", "mixed64"):
        rec[f"speedup_{form}_graph_vs_fp64_loop"] = round(rec["fp64"]["loop_ms_per_batch"]
                                                          / rec[form]["graph_ms_per_batch"], 3)
        rec[f"speedup_{form}_graph_vs_fp64_graph"] = round(rec["fp64"]["graph_ms_per_batch"]
                                                           / rec[form]["graph_ms_per_batch"], 3)
    lib.set_tuning(None)
    return rec, items


def consumer_leg(lib, problems, dA, n, reps=50):
    """The step after the product: a smoothed-aggregation hierarchy built on the
    device from the n^3 7-point operator (2x2x2 aggregates down to 4^3), then
    V(2,2) cycles with damped Jacobi until the residual drops by 1e-8, and the
    SpMV of the fine operator against the HBM roofline."""
    import numpy as np
    from sp3mm4amg_amd import amg
    dims, aggs = (n, n, n), []
    while dims[0] > 4:
        agg, cdims = problems.structured_aggregation(dims, (2, 2, 2))
        aggs.append((agg, int(np.prod(cdims))))
        dims = cdims
    t0 = time.perf_counter()
    levels, coarse = lib.hierarchy_device(dA, aggs)
    lib.sync()
    setup_ms = (time.perf_counter() - t0) * 1e3
    M, nnz, ncoarse, nlev = dA.d.M, dA.d.nnz, coarse.d.M, len(levels)
    b = np.random.default_rng(1).standard_normal(M)
    # call-by-call cycle (every operation one ABI call) ...
    vc = amg.VCycle(levels, coarse)
    vc.solve(b, rtol=1e-2, maxit=2)          # warm-up
    t0 = time.perf_counter()
    x_calls, hist = vc.solve(b, rtol=1e-8, maxit=40)
    solve_ms = (time.perf_counter() - t0) * 1e3
    cycles = len(hist) - 1
    vc.free()
    # ... and the same cycle captured as one HIP graph
    gv = amg.GraphVCycle(levels, coarse)
    gv.solve(b, rtol=1e-2, maxit=2)
    t0 = time.perf_counter()
    x_graph, hist_g, dev_ms = gv.solve(b, rtol=1e-8, maxit=40)
    graph_ms = (time.perf_counter() - t0) * 1e3
    gv.free()
    if hist_g != hist or not np.array_equal(x_graph, x_calls):
        raise RuntimeError("graph V-cycle differs from the call-by-call cycle")
    # SpMV y = A x on the fine operator, back to back on the library stream
    x = lib.DeviceArray.from_numpy(b)
    y = lib.DeviceArray(M)
    lib.spmv(dA, x, y)
    lib.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.spmv(dA, x, y)
    lib.sync()
    spmv_ms = (time.perf_counter() - t0) * 1e3 / reps
    spmv_bytes = nnz * 12 + (M + 1) * 8 + 2 * M * 8   # values, columns, offsets, x once, y
    x.free()
    y.free()
    for i, (dR, dAl, dP) in enumerate(levels):
        dR.free()
        dP.free()
        if i > 0:
            dAl.free()
    coarse.free()
    return {"workload": f"SA hierarchy on the device from the {n}^3 7-point operator (2x2x2 aggregates, "
                        f"{nlev} levels + {ncoarse}-row coarsest), V(2,2) damped Jacobi (2/3), "
                        f"50 coarse sweeps, rhs PCG64 seed 1",
            "hierarchy_setup_ms": round(setup_ms, 3),
            "cycles_to_1e-8": cycles,
            "convergence_factor": round((hist[-1] / hist[0]) ** (1.0 / max(cycles, 1)), 4),
            "ms_per_cycle_calls": round(solve_ms / max(cycles, 1), 4),
            "ms_per_cycle_graph": round(graph_ms / max(cycles, 1), 4),
            "device_ms_per_cycle_graph": round(dev_ms / max(cycles, 1), 4),
            "graph_equals_calls": True,
            "spmv_fine": {"ms": round(spmv_ms, 4), "bytes_alg": spmv_bytes,
                          "GBs": round(spmv_bytes / (spmv_ms * 1e-3) / 1e9, 1),
                          "roofline_frac": round(spmv_bytes / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


def irregular_leg(lib, problems, reps=2, n=1_000_000, tune=True):
    """Irregular rows at full size: lognormal row lengths on 1M x 1M, random
    aggregation (PCG64 seed 4), FP64, direct path, under the schedule the
    autotuner picks for it (untimed; `tune` False keeps the defaults).  MAC2
    (the products of the second stage) comes from the row upper bounds of
    (R*A)*P, with R*A formed once outside the timed region."""
    import numpy as np
    t0 = time.perf_counter()
    p = problems.irregular_sa(n, seed=4)
    R, A, P = p["R"], p["A"], p["P"]
    log(f"irregular inputs generated in {time.perf_counter() - t0:.1f}s: nnz R/A/P {R.nnz}/{A.nnz}/{P.nnz}")
    dR, dA, dP = (lib.DeviceCSR.upload(m) for m in (R, A, P))
    RA, st_ra = lib.spgemm(dR, dA)
    nnz_ra = RA.nnz
    mac1 = int(st_ra["mac1"])
    mac2 = int(lib.ub(RA, dP)[-1])
    RA.free()
    lib.set_tuning(None)
    schedule, tune_s = None, 0.0
    if tune:
        t0 = time.perf_counter()
        schedule, ms_best, ms_default = lib.autotune(dR, dA, dP, reps=1)
        tune_s = time.perf_counter() - t0
        log(f"irregular autotune: {schedule} {ms_best:.1f} ms (default {ms_default:.1f} ms) in {tune_s:.1f}s")
    out, st = lib.rap(dR, dA, dP, path=lib.DIRECT)
    out.free()
    lib.sync()
    dev, wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        out, st = lib.rap(dR, dA, dP, path=lib.DIRECT)
        lib.sync()
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(st["ms_total"])
        nnz_out = st["nnz_out"]
        out.free()
    for m in (dR, dA, dP):
        m.free()
    ms = float(np.median(dev))
    ba = alg_bytes(R, A, P, mac1, mac2, nnz_ra, nnz_out)
    bc = compulsory_bytes(R, A, P, nnz_out)
    lib.set_tuning(None)
    return {"workload": f"irregular_lognormal_{n}_seed4 Galerkin R*A*P, FP64, direct path",
            "schedule": schedule, "autotune_s": round(tune_s, 1), "reassociated": st["reassociated"],
            "rows_coarse": R.M, "nnz_A": A.nnz, "nnz_RA": nnz_ra, "nnz_RAP": nnz_out, "mac1": mac1, "mac2": mac2,
            "ms_per_product": round(ms, 1), "wall_ms": round(float(np.median(wall)), 1),
            "gflops": round(2 * (mac1 + mac2) / (ms * 1e-3) / 1e9, 1),
            "rows_off_table": st["guarded_entries"], "rows_spilled": st["spilled_rows"],
            "bytes_alg": ba, "roofline_frac": round(ba / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_compulsory": bc, "roofline_frac_compulsory": round(bc / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def host_dropin_leg(lib, R, A, P, flop, reps=3):
    """The drop-in entry point on host buffers, as the reference's callers use
    it: sp3mmRowByRowMerged_0 takes host spmat operands (u64 indices), uploads
    them, computes on the GPU and returns a malloc'd host spmat.  Times the C
    call only (the numpy-owned operands are prepared before); this is the
    PCIe-inclusive rate, never the headline value."""
    import ctypes as C

    import numpy as np
    L = lib.lib()
    hs = [lib.HostSpmat(m, 0) for m in (R, A, P)]
    cfg = lib.default_config()
    ms, nnz = [], 0
    for k in range(reps + 1):
        t0 = time.perf_counter()
        out = L.sp3mmRowByRowMerged_0(C.byref(hs[0].s), C.byref(hs[1].s), C.byref(hs[2].s), C.byref(cfg), None)
        t = (time.perf_counter() - t0) * 1e3
        if not out:
            raise RuntimeError("sp3mmRowByRowMerged_0 returned NULL")
        nnz = int(out.contents.NZ)
        L.freeSpmat(out)
        if k:
            ms.append(t)
    med = float(np.median(ms))
    moved = sum(m.nnz * 16 + (m.M + 1) * 8 for m in (R, A, P)) + nnz * 16 + (R.M + 1) * 8
    return {"call": "sp3mmRowByRowMerged_0 (host spmat in, malloc'd host spmat out)", "ms_per_call": round(med, 3),
            "gflops": round(flop / (med * 1e-3) / 1e9, 2), "device_ms_of_call": round(lib.elapsed_internal() * 1e3, 4),
            "host_bytes_moved": moved, "host_GBs": round(moved / (med * 1e-3) / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", choices=["direct", "pair"], default="direct")
    ap.add_argument("--n", type=int, default=128, help="grid side of the 3D Poisson workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-consumer", action="store_true", help="skip the V-cycle / SpMV leg")
    ap.add_argument("--no-irregular", action="store_true", help="skip the 1M irregular-rows leg")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer drop-in leg")
    ap.add_argument("--schedule", default=None,
                    help="install this schedule instead of autotuning, e.g. u1=4,u2=3,subwave_h=0 "
                         "(fields of sp3mm_tuning; the others keep their defaults)")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the default schedule instead of autotuning each workload (untimed setup)")
    ap.add_argument("--only-leg", choices=["mixed", "batch", "irregular", "consumer"], default=None,
                    help="run only this leg and print its record (for profiling; not the benchmark line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the product is sharded by rows "
                                   "of R across GPUs only in design (--gpus 1)"}))
        sys.exit(2)

    import numpy as np
    from sp3mm4amg_amd import lib, problems

    lib.init(0)
    if args.only_leg:
        steps = max(3, args.steps // 2)
        if args.only_leg == "mixed":
            leg = mixed_leg(lib, problems, steps, args.warmup, not args.no_tune)[0]
        elif args.only_leg == "batch":
            leg = batch_leg(lib, problems, max(3, args.steps // 4), tune=not args.no_tune)[0]
        elif args.only_leg == "irregular":
            leg = irregular_leg(lib, problems, reps=max(1, args.steps), tune=not args.no_tune)
        else:
            p = problems.poisson3d_sa(args.n)
            dA = lib.DeviceCSR.upload(p["A"])
            leg = consumer_leg(lib, problems, dA, args.n)
            dA.free()
        print(json.dumps({"leg": args.only_leg, "record": leg}), flush=True)
        return
    t0 = time.perf_counter()
    prob = problems.poisson3d_sa(args.n)
    R, A, P = prob["R"], prob["A"], prob["P"]
    log(f"generated {prob['name']} in {time.perf_counter() - t0:.1f}s: nnz R/A/P "
        f"{R.nnz}/{A.nnz}/{P.nnz}")
    dR, dA, dP = (lib.DeviceCSR.upload(m) for m in (R, A, P))
    path = lib.DIRECT if args.path == "direct" else lib.PAIR

    # MAC counts and RA size from the pair form (one untimed product)
    out, st_pair = lib.rap(dR, dA, dP, path=lib.PAIR)
    out.free()
    mac1, mac2, nnz_ra = st_pair["mac1"], st_pair["mac2"], st_pair["nnz_mid"]
    flop = 2 * (mac1 + mac2)
    schedule, tune_s = None, 0.0
    if args.schedule:
        lib.set_tuning({k: int(v, 0) for k, v in (f.split("=") for f in args.schedule.split(","))})
        schedule = lib.get_tuning()
        log(f"schedule installed: {schedule}")
    elif not args.no_tune and path == lib.DIRECT:
        t0 = time.perf_counter()
        schedule, ms_best, ms_default = lib.autotune(dR, dA, dP, reps=9)
        tune_s = time.perf_counter() - t0
        log(f"autotune: {schedule} {ms_best:.4f} ms (default schedule {ms_default:.4f} ms) in {tune_s:.1f}s")

    for _ in range(args.warmup):
        out, st = lib.rap(dR, dA, dP, path=path)
        out.free()
    lib.lib().sp3mm_sync()
    total_ms, step_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        out, st = lib.rap(dR, dA, dP, path=path)
        total_ms.append(st["ms_total"])
        nnz_out = st["nnz_out"]
        out.free()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    lib.lib().sp3mm_sync()
    wall = time.perf_counter() - t0
    main_ms = main_kernel_ms(lib, dR, dA, dP, args.steps, path=path)
    ms_per_step = wall * 1e3 / args.steps
    value = flop * args.steps / wall / 1e9

    bytes_alg = alg_bytes(R, A, P, mac1, mac2, nnz_ra, nnz_out)
    main_avg = float(np.mean(main_ms))
    achieved = bytes_alg / (main_avg * 1e-3) / 1e9
    dev_med = float(np.median(total_ms))
    copy_ms = lib.copy_probe(1 << 30, 20)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9 if copy_ms > 0 else None

    rec = {
        "metric": "Galerkin RAP ms & GFLOP/s: FP64 vs mixed-precision vs batched, 1 MI355X",
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{prob['name']} Galerkin R*A*P, FP64, {args.path} path",
                   "rows_fine": A.M, "rows_coarse": R.M, "nnz_R": R.nnz, "nnz_A": A.nnz,
                   "nnz_P": P.nnz, "nnz_RA": nnz_ra, "nnz_RAP": nnz_out, "mac1": mac1,
                   "mac2": mac2, "flop": flop, "schedule": schedule, "autotune_s": round(tune_s, 2)},
        # the dominant kernel's roofline (algorithmic bytes of the product over
        # the main kernel's mean duration), and beside it the target figure:
        # the same bytes over the whole product's device time
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": committed_traffic() if args.path == "direct" and args.n == 128 else None,
                     "traffic_source": "profiles/pmc_direct_poisson3d128.md",
                     "kernel": f"numeric {args.path} kernel(s)", "kernel_ms": round(main_avg, 4),
                     "kernel_ms_spread": spread(main_ms),
                     "bytes_alg": bytes_alg,
                     "target_frac_whole_product": round(bytes_alg / (dev_med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "target_definition": "bytes_alg / (device time of the complete product: binning, "
                                          "main kernel, scan, output allocation, gather) / 8 TB/s"},
        "product": {"device_ms_median": round(dev_med, 4),
                    "device_ms": spread(total_ms),
                    "wall_ms_per_step": spread(step_ms),
                    "roofline_frac": round(bytes_alg / (dev_med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "gflops_device": round(flop / (dev_med * 1e-3) / 1e9, 2),
                    "bytes_compulsory": compulsory_bytes(R, A, P, nnz_out),
                    "roofline_frac_compulsory": round(compulsory_bytes(R, A, P, nnz_out) / (dev_med * 1e-3) / 1e9
                                                      / HBM_PEAK_GBS, 4)},
        "hbm_copy_probe_GBs": round(copy_gbs, 1) if copy_gbs else None,
    }
    if path == lib.DIRECT:
        # the same workload on the pair path (R*A)*P, both SpGEMMs one-pass
        wp, _, devp, _, devsp = time_form(lib, dR, dA, dP, max(5, args.steps // 2), 1, path=lib.PAIR)
        rec["pair_path"] = {"ms_per_product": round(wp, 4), "device_ms_median": round(devp, 4),
                            "device_ms": spread(devsp),
                            "gflops": round(flop / wp / 1e6, 2),
                            "roofline_frac_product": round(bytes_alg / (devp * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    stencil27 = batch_items = None
    if not args.no_mixed:
        log("mixed-precision leg (27-point stencil) ...")
        rec["mixed_precision"], stencil27 = mixed_leg(lib, problems, max(3, args.steps // 2), 1, not args.no_tune)
    if not args.no_batch:
        log("batched leg (hierarchy + 256 problems) ...")
        rec["batched"], batch_items = batch_leg(lib, problems, max(3, args.steps // 4), tune=not args.no_tune)
    if not args.no_consumer:
        log("consumer leg (device hierarchy + V-cycle + SpMV) ...")
        rec["consumer"] = consumer_leg(lib, problems, dA, args.n)
    if not args.no_host:
        log("host-buffer drop-in leg")
        rec["host_dropin"] = host_dropin_leg(lib, R, A, P, flop)
    if not args.no_irregular:
        log("irregular leg (1M lognormal rows) ...")
        rec["irregular"] = irregular_leg(lib, problems, tune=not args.no_tune)
    if not args.no_cpu:
        log("timing the reference CPU path on samples ...")
        p1 = problems.poisson2d_sa(64)
        work = [("poisson3d128", f"{prob['name']} (FP64)", [(R, A, P)], 18.0, 3),
                ("poisson2d64", f"{p1['name']} (FP64, whole product, 40 timed runs)", [(p1["R"], p1["A"], p1["P"])],
                 0.0, 40)]
        if stencil27 is not None:
            work.append(("stencil27_100", "stencil27_100^3_sa3x3x3_seed3 (FP64 values)", [stencil27], 8.0, 3))
        if batch_items is not None:
            work.append(("batch261", "5-level 96^3 hierarchy + 256 2D problems (FP64, every product sampled)",
                         batch_items, 8.0, 3))
        rec["cpu_baseline"] = cpu_baseline(work)
    for m in (dR, dA, dP):
        m.free()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
