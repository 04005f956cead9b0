#!/usr/bin/env python3
"""Benchmark: derived tuples/s of the semi-naive fixpoint engine on one MI355X.

Workload (default ``sg_mesh``): same generation on the synthetic 3D finite-
element mesh -- 35x35x35 hex cells, each split into 5 tetrahedra with
parity-alternating orientation, tetrahedron edges directed low -> high id
(268,380 int32 edges).  It is the largest of the benchmark configurations that
fits one GPU: the 400x400 grid's transitive closure holds >= 2.0e10 tuples
(giant component of ~141k vertices), beyond one device and beyond the
reference's int32 counters.  Other workloads: ``tc_powerlaw`` (3,000 vertices,
7,000 edges), ``tc_usroad`` (the reference's usroad graph, 606 iterations),
``wcc_grid`` (WCC on the 400x400 grid), ``single_join`` (10M rows, range 1M).

One step = one whole fixpoint (input rows already resident in HBM, output
relation resident in HBM at the end).  value = |output relation| * steps / time.
Per-kernel durations are measured with HIP events on the engine's stream
during the timed steps; `roofline` reports the kernel with the largest total
time.  `cpu_baseline` times the CPU oracle on a bounded prefix of the same
workload (the first iterations of the same fixpoint) and checks that the GPU
produces the identical relation on that prefix.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 0x4D4E4D47

# Reported HBM traffic per launch of the dominant kernel, from rocprofv3 --pmc
# FETCH_SIZE + WRITE_SIZE runs of this same command (bytes per launch), keyed
# by (workload, kernel).  Filled from the measured profile summaries.
MEASURED_TRAFFIC = {}


def workload_rows(name):
    import mnmgdatalog_amd as dlg
    if name == "sg_mesh":
        return "sg", dlg.gen_tetmesh(35), {"workload": "sg_tetmesh_35", "cells": 35, "edges": None,
                                           "program": "sg"}
    if name == "tc_powerlaw":
        return "tc", dlg.gen_powerlaw(3000, 7000, 2.1, SEED), {"workload": "tc_powerlaw_3000v_7000e",
                                                             "program": "tc"}
    if name == "tc_usroad":
        return "tc", dlg.read_rows(os.path.join(ROOT, "tests/golden/data/data_165435.bin")), {
            "workload": "tc_usroad_165435e", "program": "tc"}
    if name == "wcc_grid":
        return "wcc", dlg.gen_grid(400, 0.55, 0.01, 20, SEED), {"workload": "wcc_grid_400", "program": "wcc"}
    if name == "single_join":
        return "single_join", dlg.generate_rand_rows(10_000_000, 1_000_000), {
            "workload": "single_join_10M_range1M", "program": "single_join"}
    raise SystemExit(f"unknown workload {name}")


# bounded CPU sample per workload: number of fixpoint iterations the oracle runs
CPU_SAMPLE_ITERS = {"sg_mesh": 3, "tc_powerlaw": 0, "tc_usroad": 40, "wcc_grid": 0, "single_join": 0}


def cpu_baseline(name, prog, rows, ctx, threads):
    import numpy as np

    import oracle
    k = CPU_SAMPLE_ITERS.get(name, 0)
    fn = {"tc": oracle.tc, "sg": oracle.sg, "wcc": oracle.wcc, "single_join": oracle.single_join}[prog]
    t0 = time.perf_counter()
    if k and prog in ("tc", "sg"):
        want = fn(rows, threads=threads, max_iterations=k)
    else:
        want = fn(rows, threads=threads)
    dt = time.perf_counter() - t0
    extra = {"max_iterations": k} if k and prog in ("tc", "sg") else {}
    got = ctx.run(prog, rows, keep_output=True, **extra)
    same = (got["iterations"] == want["iterations"] and got["output_size"] == want["output_size"]
            and np.array_equal(got["rows"], want["rows"]))
    sample = (f"{prog} on the same input, first {want['iterations']} fixpoint iterations"
              if k else f"{prog} on the same input, full fixpoint ({want['iterations']} iterations)")
    return {"value": want["output_size"] / dt, "unit": "derived tuples/s", "cores": threads,
            "kind": "port", "sample": sample, "seconds": round(dt, 3),
            "sample_output_tuples": int(want["output_size"]), "gpu_matches_sample": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sg_mesh")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design "
                                   "(hash partition of the join key + all-to-all), not built"}))
        raise SystemExit(2)

    import mnmgdatalog_amd as dlg

    prog, rows, cfg = workload_rows(args.workload)
    cfg["edges"] = int(rows.shape[0])
    n = rows.shape[0]
    with dlg.Ctx(0) as ctx:
        info = ctx.device_info()
        d_rows = ctx.upload(rows)
        res = None
        for _ in range(args.warmup):
            res = ctx.run(prog, None, keep_output=False, device_rows=d_rows, n=n)
        ctx.kernel_stats_reset()
        ctx.profiling(True)
        ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
        t0 = time.perf_counter()
        dev_s = 0.0
        for _ in range(args.steps):
            res = ctx.run(prog, None, keep_output=False, device_rows=d_rows, n=n)
            dev_s += res["device_seconds"]
        ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
        t1 = time.perf_counter()
        ctx.profiling(False)
        ks = ctx.kernel_stats()
        elapsed = t1 - t0
        out_tuples = res["output_size"]
        value = out_tuples * args.steps / elapsed
        dom = max(ks.items(), key=lambda kv: kv[1]["ms"])
        name, st = dom
        avg_ms = st["ms"] / max(st["launches"], 1)
        alg_per_launch = st["bytes"] / max(st["launches"], 1)
        achieved = alg_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        traffic = MEASURED_TRAFFIC.get((args.workload, name))
        kernels = {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                       "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else 0.0}
                   for k, v in ks.items() if v["launches"]}
        cpu = None
        if not args.no_cpu:
            cpu = cpu_baseline(args.workload, prog, rows, ctx, args.cpu_threads)
        ctx.free(d_rows)
    line = {
        "metric": "derived tuples/s, 1 MI355X, bit-exact; batched/served vs plain speedup + p99",
        "value": value, "unit": "derived tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": dict(cfg, iterations=res["iterations"], output_tuples=int(out_tuples),
                       raw_join_rows=int(res["raw_join_rows"])),
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": alg_per_launch, "avg_launch_ms": avg_ms,
                     "launches": st["launches"]},
        "kernels": kernels,
        "cpu_baseline": cpu,
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
