#!/usr/bin/env python3
"""Benchmark: Krylov-Schur on one MI355X.

Default workload, ``--workload fp64`` (one "step" = one full EPSSolve): 3-D 7-point variable-coefficient
Laplacian on a 128^3 grid (n = 2,097,152, 14,581,760 nnz, CSR FP64/int32),
lognormal(0,1) face coefficients from seed 1, diagonal = sum of the cell's face
coefficients; nev = 16 largest magnitude, ncv = 48, tol = 1e-10, start vector
U(-1,1) from seed 0.  Inputs are resident in HBM before the timed region.

value = converged eigenpairs/s, counting only pairs among the nev requested
whose relative residual ||Ax - lambda x|| / |lambda|, recomputed on the device
by EPSComputeError after the timed region, is <= 1e-10.

Other forms, each printing its own JSON line:

* ``--workload mixed`` -- nonsymmetric 2-D upwind convection-diffusion on a
  2048^2 grid (Peclet 10, diffusion uniform(0.5,2) per cell, seed 2), nev 8
  largest real part, ncv 40, tol 1e-10: the mixed-precision form (FP32 basis to
  1e-5, then FP64 restarts) timed beside the plain FP64 solve of the same problem.
* ``--workload serve`` -- a stream of weighted Erdos-Renyi graph Laplacians
  (average degree 8, n log-uniform in [1e3, 2e5], nev uniform in {4..32},
  ncv = 2 nev + 16, tol 1e-10) arriving by a Poisson process at ``--qps``,
  served concurrently under an HBM budget, timed beside the same requests
  solved one after another.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
METRIC = "converged eigenpairs/s at rel-residual<=1e-10; SpMV+BV HBM GB/s; p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(csr, n, nev, ncv, tol, gpu_its, budget_s):
    """The CPU restatement of the reference path (oracle/) on the host cores,
    timed on a bounded number of restarts of the same solve and scaled by the
    restart count the full solve needs."""
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("MKL_NUM_THREADS", str(cores))
    v0 = O.random_uniform(0, n)
    t0 = time.time()
    probe = O.eps_solve(csr, nev=nev, ncv=ncv, tol=tol, v0=v0, nthreads=cores, restart_cap=1,
                        want_vectors=False)
    t_probe = time.time() - t0
    cap = max(2, min(gpu_its, int(budget_s / max(t_probe, 1e-3))))
    log(f"[cpu] 1 restart took {t_probe:.2f}s; timing {cap} restarts on {cores} threads")
    t0 = time.time()
    r = O.eps_solve(csr, nev=nev, ncv=ncv, tol=tol, v0=v0, nthreads=cores, restart_cap=cap,
                    want_vectors=False)
    t = time.time() - t0
    per_restart = t / max(r["its"], 1)
    full = per_restart * gpu_its
    return {"value": nev / full, "unit": "eigenpairs/s", "cores": cores, "kind": "port",
            "sample": (f"first {r['its']} restarts of the same solve (same matrix, start vector, "
                       f"nev/ncv/tol) with the oracle's C restatement (OpenMP SpMV, MKL BLAS/LAPACK); "
                       f"{per_restart:.3f} s/restart x {gpu_its} restarts the full solve takes"),
            "seconds_timed": t, "projected_solve_s": full}


def _measured_traffic(kernel, grid):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC profile of this workload (FETCH_SIZE / WRITE_SIZE, calibrated on a
    streaming kernel of known size), or None when there is none for it."""
    path = os.path.join(ROOT, "profiles", "ks_fp64_varcoef128_hbm_traffic.json")
    if grid != 128 or not os.path.exists(path):
        return None
    k = {"gs_update_first": "gs_fused_sweep", "gs_dot_first": "gs_first_sweep", "spmv": "spmv",
         "multinplace": "multinplace"}.get(kernel)
    with open(path) as f:
        prof = json.load(f)
    return prof["kernels"][k]["traffic_bytes_per_launch"] if k in prof["kernels"] else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["fp64", "mixed", "serve", "copycal"], default="fp64")
    ap.add_argument("--requests", type=int, default=400, help="serve: number of requests")
    ap.add_argument("--qps", type=float, default=200.0, help="serve: Poisson arrival rate")
    ap.add_argument("--slots", type=int, default=8, help="serve: concurrent solves")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="serve: HBM admission budget")
    ap.add_argument("--mixed-grid", type=int, default=2048)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--nev", type=int, default=16)
    ap.add_argument("--ncv", type=int, default=48)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-prof", action="store_true", help="time without per-kernel HIP events")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the Krylov-Schur path here is built for a single MI355X "
              f"(--gpus {args.gpus} requested).  Multi-GPU sharding is a design note, not a run mode.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "mixed":
        return run_mixed(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "copycal":
        return run_copycal(args)

    from slepc_amd import eps as E
    from slepc_amd import ks as K
    from slepc_amd import matgen as M

    N = args.grid
    t0 = time.time()
    csr = M.varcoef_laplacian_3d(N, seed=1)
    n = N ** 3
    nnz = int(csr[0][-1])
    log(f"[bench] matrix {N}^3 n={n} nnz={nnz} built in {time.time() - t0:.1f}s")

    E.initialize(0)
    ctx = E.context_handle()
    A = E.Mat(csr)
    info = A.info()
    log(f"[bench] device CSR: {info}")
    eps = E.EPS()
    eps.setOperators(A)
    eps.setProblemType(E.EPS.HEP)
    eps.setDimensions(args.nev, args.ncv)
    eps.setTolerances(args.tol)
    eps.setWhichEigenpairs(E.EPS.LARGEST_MAGNITUDE)
    v0 = E.Vec.random(n, 0)

    def one_solve():
        eps.setInitialSpace([v0])
        ts = time.perf_counter()
        eps.solve()
        return time.perf_counter() - ts

    for w in range(args.warmup):
        dt = one_solve()
        st = eps.stats()
        log(f"[bench] warmup {w}: {dt:.3f}s its={st['its']} nconv={eps.getConverged()} "
            f"gs_passes={st['gs_passes']}")

    L = K.lib()
    if not args.no_prof:
        K.check(L.ks_prof_reset(ctx), "ks_prof_reset")
        K.check(L.ks_prof_enable(ctx, 1), "ks_prof_enable")
    K.check(L.ks_ctx_sync(ctx), "sync")
    times, its_list, passes = [], [], []
    t_begin = time.perf_counter()
    for s in range(args.steps):
        dt = one_solve()
        times.append(dt)
        st = eps.stats()
        its_list.append(st["its"])
        passes.append(st["gs_passes"])
        log(f"[bench] step {s}: {dt:.3f}s its={st['its']} nconv={eps.getConverged()} "
            f"gs_passes={st['gs_passes']} t_ds={st['t_ds']:.3f}s")
    K.check(L.ks_ctx_sync(ctx), "sync")
    t_total = time.perf_counter() - t_begin

    prof = {}
    if not args.no_prof:
        K.check(L.ks_prof_collect(ctx), "ks_prof_collect")
        for i, name in enumerate(K.PROF_CLASSES):
            e = K.ProfEntry()
            K.check(L.ks_prof_get(ctx, i, C_byref(e)), "ks_prof_get")
            prof[name] = dict(launches=e.launches, skipped=e.skipped, ms=e.ms, bytes=e.bytes, flops=e.flops)
        K.check(L.ks_prof_enable(ctx, 0), "ks_prof_enable")

    # verification outside the timed region (last solve)
    nconv = eps.getConverged()
    errs = [eps.computeError(i) for i in range(min(nconv, args.nev))]
    verified = sum(1 for e in errs if e <= args.tol)
    vals = [float(np.real(eps.getEigenvalue(i))) for i in range(min(nconv, args.nev))]
    log(f"[bench] verified {verified}/{args.nev} pairs, max rel residual {max(errs) if errs else float('nan'):.3e}")

    value = verified * args.steps / t_total
    lat = np.array(times) * 1e3

    roofline = None
    kernels = {}
    if prof:
        for name, e in prof.items():
            if e["launches"] and e["ms"] > 0:
                kernels[name] = dict(launches=e["launches"], skipped=e["skipped"],
                                     avg_us=1e3 * e["ms"] / e["launches"],
                                     GBs=e["bytes"] / (e["ms"] * 1e-3) / 1e9,
                                     TFLOPs=e["flops"] / (e["ms"] * 1e-3) / 1e12,
                                     total_ms=e["ms"], bytes_per_launch=e["bytes"] / e["launches"])
        # dominant kernel among those that run on every step (never gated off)
        cands = [k for k in ("gs_update_first", "gs_dot_first", "spmv", "multinplace") if k in kernels]
        dom = max(cands, key=lambda k: kernels[k]["total_ms"])
        dk = kernels[dom]
        roofline = {"bound": "hbm", "achieved": round(dk["GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(dk["GBs"] / HBM_PEAK_GBS, 4), "traffic": _measured_traffic(dom, N), "kernel": dom,
                    "avg_us": round(dk["avg_us"], 2), "bytes_per_launch": dk["bytes_per_launch"]}
        tot_b = sum(e["bytes"] for e in prof.values())
        tot_ms = sum(e["ms"] for e in prof.values())
        roofline["spmv_bv_aggregate_GBs"] = round(tot_b / (tot_ms * 1e-3) / 1e9, 1) if tot_ms else None
        roofline["kernel_time_fraction_of_wall"] = round(tot_ms * 1e-3 / t_total, 3)
        if "multinplace" in kernels:
            mk = kernels["multinplace"]
            roofline["multinplace"] = {"GBs": round(mk["GBs"], 1), "hbm_frac": round(mk["GBs"] / HBM_PEAK_GBS, 4),
                                       "fp64_mfma_TFLOPs": round(mk["TFLOPs"], 3),
                                       "fp64_mfma_frac": round(mk["TFLOPs"] / FP64_MFMA_PEAK_TFLOPS, 5)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(csr, n, args.nev, args.ncv, args.tol, int(np.median(its_list)), args.cpu_seconds)
            log(f"[bench] cpu baseline {cpu['value']:.4f} eigenpairs/s on {cpu['cores']} threads")
        except Exception as ex:   # the oracle is optional on the timing box; report why it is missing
            log(f"[bench] cpu baseline unavailable: {ex}")
            cpu = {"value": None, "unit": "eigenpairs/s", "cores": None, "kind": "port",
                   "sample": f"unavailable: {ex}"}

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D 7-point variable-coefficient Laplacian {N}^3 (lognormal(0,1) faces, seed 1), "
                               f"plain FP64 Krylov-Schur, nev {args.nev} LM, ncv {args.ncv}, tol {args.tol:g}",
                   "n": n, "nnz": nnz, "nev": args.nev, "ncv": args.ncv, "tol": args.tol, "which": "LM",
                   "start_vector": "U(-1,1) seed 0", "restarts": its_list, "gs_passes": passes,
                   "verified_pairs": verified, "eigenvalues": vals[:4]},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    print(json.dumps(out), flush=True)


def _base(args, value, unit, ms_per_step, steps, dtype, workload, extra):
    out = {"metric": METRIC, "value": round(value, 4), "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
           "config": {"workload": workload}}
    out.update(extra)
    return out


def run_mixed(args):
    """Mixed-precision Krylov-Schur vs plain FP64 on the convection-diffusion operator."""
    from slepc_amd import eps as E
    from slepc_amd import matgen as M
    N = args.mixed_grid
    t0 = time.time()
    csr = M.convection_diffusion_2d(N, 10.0, seed=2)
    n = N * N
    log(f"[mixed] matrix {N}^2 n={n} nnz={int(csr[0][-1])} built in {time.time() - t0:.1f}s")
    E.initialize(0)
    A = E.Mat(csr)
    v0 = E.Vec.random(n, 0)
    nev, ncv, tol = 8, 40, 1e-10

    from slepc_amd import ks as K
    L = K.lib()
    ctx = E.context_handle()

    def prof_table():
        K.check(L.ks_prof_collect(ctx), "ks_prof_collect")
        tab = {}
        for i, name in enumerate(K.PROF_CLASSES):
            e = K.ProfEntry()
            K.check(L.ks_prof_get(ctx, i, C_byref(e)), "ks_prof_get")
            if e.launches and e.ms > 0:
                tab[name] = {"launches": e.launches, "ms": round(e.ms, 3),
                             "GBs": round(e.bytes / (e.ms * 1e-3) / 1e9, 1)}
        return tab

    def solve(mixed):
        K.check(L.ks_prof_reset(ctx), "ks_prof_reset")
        K.check(L.ks_prof_enable(ctx, 1), "ks_prof_enable")
        eps = E.EPS()
        eps.setOperators(A)
        eps.setProblemType(E.EPS.NHEP)
        eps.setDimensions(nev, ncv)
        eps.setTolerances(tol)
        eps.setWhichEigenpairs(E.EPS.LARGEST_REAL)
        if mixed:
            eps.setMixedPrecision(1e-5)
        eps.setInitialSpace([v0])
        ts = time.perf_counter()
        eps.solve()
        dt = time.perf_counter() - ts
        prof = prof_table()
        K.check(L.ks_prof_enable(ctx, 0), "ks_prof_enable")
        k = min(eps.getConverged(), nev)
        errs = [eps.computeError(i) for i in range(k)]
        vals = [eps.getEigenvalue(i) for i in range(k)]
        st = eps.stats()
        ms = eps.mixedStats() if mixed else {}
        eps.destroy()
        return dict(t=dt, verified=sum(e <= tol for e in errs), errs=errs, vals=vals, its=st["its"],
                    nmatmult=st["nmatmult"], mixed=ms, prof=prof)

    res = {}
    for form in ("fp64", "mixed"):
        for w in range(args.warmup):
            r = solve(form == "mixed")
            log(f"[mixed] warmup {form}: {r['t']:.3f}s its={r['its']}")
        runs = [solve(form == "mixed") for _ in range(args.steps)]
        for r in runs:
            log(f"[mixed] {form}: {r['t']:.3f}s its={r['its']} verified={r['verified']} {r['mixed']}")
        res[form] = runs
    # per-pair error of the mixed form against the FP64 form (same problem, same start)
    ref = res["fp64"][-1]["vals"]
    got = res["mixed"][-1]["vals"]
    pair_err = [abs(complex(g) - complex(r)) / abs(complex(r)) for g, r in zip(got, ref)]
    rate = {f: sum(r["verified"] for r in res[f]) / sum(r["t"] for r in res[f]) for f in res}
    out = _base(args, rate["mixed"], "eigenpairs/s", 1e3 * np.mean([r["t"] for r in res["mixed"]]), args.steps,
                "f32+f64",
                f"2-D upwind convection-diffusion {N}^2 (Peclet 10, diffusion U(0.5,2) seed 2), mixed-precision "
                f"Krylov-Schur (FP32 basis to 1e-5, then FP64 to 1e-10), nev {nev} LR, ncv {ncv}",
                {"forms": {f: {"eigenpairs_per_s": round(rate[f], 4),
                               "solve_s": [round(r["t"], 4) for r in res[f]],
                               "restarts": [r["its"] for r in res[f]],
                               "spmv": [r["nmatmult"] for r in res[f]],
                               "verified": [r["verified"] for r in res[f]],
                               "kernels": res[f][-1]["prof"]} for f in res},
                 "mixed_phase": res["mixed"][-1]["mixed"],
                 "speedup_vs_fp64": round(rate["mixed"] / rate["fp64"], 3) if rate["fp64"] else None,
                 "pair_rel_error_vs_fp64": [float(f"{e:.3e}") for e in pair_err],
                 "max_rel_residual": float(f"{max(res['mixed'][-1]['errs']):.3e}")})
    out["config"].update({"n": n, "nnz": int(csr[0][-1]), "nev": nev, "ncv": ncv, "tol": tol, "which": "LR"})
    print(json.dumps(out), flush=True)


def run_copycal(args):
    """A streaming kernel of known traffic (k_scale: read and write 64 Mi doubles)
    -- the calibration for the hardware counters FETCH_SIZE / WRITE_SIZE."""
    from slepc_amd import ks as K
    ctx = K.Context(0)
    n = 64 << 20
    x = K.DeviceArray(ctx, (n,))
    K.random(ctx, K.F64, n, x, 1)
    ctx.sync()
    ctx.prof(True)
    ctx.prof_reset()
    for _ in range(max(1, args.steps)):
        K.scale(ctx, K.F64, n, x, 1.0000001)
    ctx.sync()
    e = ctx.prof_get()["scale"]
    ctx.prof(False)
    us = 1e3 * e["ms"] / e["launches"]
    out = {"workload": "copycal", "kernel": "k_scale<double>", "n": n, "bytes_read_per_launch": 8 * n,
           "bytes_written_per_launch": 8 * n, "launches": e["launches"], "avg_us": round(us, 2),
           "GBs": round(16.0 * n / (us * 1e-6) / 1e9, 1)}
    print(json.dumps(out), flush=True)


def _er_job(r):
    from slepc_amd import matgen as M
    return M.er_laplacian(r["n"], 8.0, r["graph_seed"])


def run_serve(args):
    """Poisson stream of ER-Laplacian eigenproblems: served concurrently vs one after another."""
    from concurrent.futures import ProcessPoolExecutor
    from slepc_amd import eps as E
    from slepc_amd import matgen as M
    from slepc_amd import serve as S
    reqs = M.serving_stream(args.requests, args.qps, seed=4)
    t0 = time.time()
    workers = max(1, min(16, os.cpu_count() or 1))
    with ProcessPoolExecutor(workers) as ex:
        mats = list(ex.map(_er_job, reqs, chunksize=8))
    log(f"[serve] {len(reqs)} matrices (total nnz {sum(int(m[0][-1]) for m in mats)}) built in "
        f"{time.time() - t0:.1f}s with {workers} workers")
    E.initialize(0)
    cache = {i: m for i, m in enumerate(mats)}
    for i, r in enumerate(reqs):
        r["_i"] = i
    # served: concurrent, admitted against the budget; first with each request
    # launching its own steps (streams only), then with grouped steps
    srv = S.Server(budget=int(args.budget_gb * (1 << 30)), max_active=args.slots, grouped=False)
    res_ng, wall_ng = S.run_stream(srv, reqs, lambda r: cache[r["_i"]], tol=1e-10)
    srv.close()
    good_ng = sum(r["ngood"] for r in res_ng)
    log(f"[serve] streams only: {wall_ng:.2f}s, {good_ng} verified pairs")
    srv = S.Server(budget=int(args.budget_gb * (1 << 30)), max_active=args.slots, grouped=True)
    results, wall = S.run_stream(srv, reqs, lambda r: cache[r["_i"]], tol=1e-10)
    stats = srv.stats()
    srv.close()
    good = sum(r["ngood"] for r in results)
    lat = np.array([r["latency"] for r in results]) * 1e3
    log(f"[serve] served {len(reqs)} requests in {wall:.2f}s: {good} verified pairs, "
        f"p50 {np.percentile(lat, 50):.1f} ms p99 {np.percentile(lat, 99):.1f} ms")
    # the same requests, one EPSSolve after another on one stream (no arrivals: back to back)
    ts = time.perf_counter()
    seq_good = 0
    seq_lat = []
    for i, r in enumerate(reqs):
        t1 = time.perf_counter()
        out = E.solve(cache[i], nev=r["nev"], ncv=r["ncv"], tol=1e-10, v0=M.uniform_vector(r["v0_seed"], r["n"]),
                      want_vectors=False)
        seq_lat.append(time.perf_counter() - t1)
        seq_good += int(np.sum(np.asarray(out["err"][:r["nev"]]) <= 1e-10))
    seq_wall = time.perf_counter() - ts
    log(f"[serve] sequential: {seq_wall:.2f}s, {seq_good} verified pairs")
    rate = good / wall
    seq_rate = seq_good / seq_wall
    out = _base(args, rate, "eigenpairs/s", 1e3 * wall / len(reqs), len(reqs), "f64",
                f"serving: {len(reqs)} weighted Erdos-Renyi Laplacians (degree 8, n log-uniform [1e3,2e5], "
                f"nev U{{4..32}}, ncv 2nev+16, LM, tol 1e-10), Poisson arrivals at {args.qps:g} QPS, "
                f"{args.slots} slots, {args.budget_gb:g} GiB HBM budget",
                {"latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                                "p99": round(float(np.percentile(lat, 99)), 2)},
                 "served": {"eigenpairs_per_s": round(rate, 4), "wall_s": round(wall, 3), "verified_pairs": good,
                            "requests": len(reqs), "peak_reserved_bytes": stats["peak_reserved"],
                            "grouped_steps": stats["group_steps"],
                            "host_ds_s": round(sum(r["t_host_ds"] for r in results), 3),
                            "host_loop_s": {k: round(v, 3) for k, v in stats["host_s"].items()},
                            "restarts": int(sum(r["its"] for r in results)),
                            "members_per_grouped_step": round(stats["group_member_steps"] / max(1, stats["group_steps"]), 2)},
                 "served_streams_only": {"eigenpairs_per_s": round(good_ng / wall_ng, 4), "wall_s": round(wall_ng, 3),
                                         "verified_pairs": good_ng,
                                         "latency_ms_p50": round(float(np.percentile([r["latency"] * 1e3 for r in res_ng], 50)), 2),
                                         "latency_ms_p99": round(float(np.percentile([r["latency"] * 1e3 for r in res_ng], 99)), 2)},
                 "sequential": {"eigenpairs_per_s": round(seq_rate, 4), "wall_s": round(seq_wall, 3),
                                "verified_pairs": seq_good,
                                "latency_ms_p50": round(float(np.percentile(np.array(seq_lat) * 1e3, 50)), 2)},
                 "speedup_vs_sequential": round(rate / seq_rate, 3) if seq_rate else None})
    out["config"].update({"requests": len(reqs), "qps": args.qps, "slots": args.slots})
    print(json.dumps(out), flush=True)


def C_byref(x):
    import ctypes
    return ctypes.byref(x)


if __name__ == "__main__":
    main()
