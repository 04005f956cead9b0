#!/usr/bin/env python3
"""Benchmark: derived tuples/s of the semi-naive fixpoint engine on one MI355X.

Workload (default ``tc_grid``): transitive closure of the road-network-shaped
grid -- 400x400 vertices, each undirected lattice edge kept with p = 0.55 and
stored in both directions, plus shortcut edges (1% of the kept lattice edges)
between vertices within Manhattan distance 20 (354,870 int32 rows).  Its
closure holds 20,691,180,264 tuples (501 iterations): the full relation is
kept as a 160,000 x 160,000-bit map (3.2 GB) and the sorted result (166 GB)
stays resident in HBM.  Other workloads: ``sg_mesh`` (same generation on the
35x35x35 tetrahedral mesh, 268,380 edges), ``tc_powerlaw`` (3,000 vertices,
7,000 edges), ``tc_usroad`` (the reference's usroad graph, 606 iterations),
``wcc_grid`` (WCC on the 400x400 grid), ``single_join`` (10M rows, range 1M),
``batched`` and ``served``.

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

# HBM traffic per launch of the dominant kernel, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of `bench.py --workload W --steps 1
# --warmup 0 --no-cpu`, scaled by the copy-kernel calibration (FETCH x 2.0,
# WRITE x 1.0): bytes per launch, keyed by (workload, kernel id).
MEASURED_TRAFFIC = {("tc_grid", "bitmap_insert"): 4.045e9}


def workload_rows(name):
    import mnmgdatalog_amd as dlg
    if name == "tc_grid":
        return "tc", dlg.gen_grid(400, 0.55, 0.01, 20, SEED), {
            "workload": "tc_grid_400x400_p0.55_shortcuts1pct_d20", "program": "tc"}
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
CPU_SAMPLE_ITERS = {"tc_grid": 30, "sg_mesh": 3, "tc_powerlaw": 0, "tc_usroad": 40, "wcc_grid": 0, "single_join": 0}


def cpu_baseline(name, prog, rows, ctx, threads, strategy=0):
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
    got = ctx.run(prog, rows, keep_output=True, strategy=strategy, **extra)
    same = (got["iterations"] == want["iterations"] and got["output_size"] == want["output_size"]
            and np.array_equal(got["rows"], want["rows"]))
    sample = (f"{prog} on the same input, first {want['iterations']} fixpoint iterations"
              if k else f"{prog} on the same input, full fixpoint ({want['iterations']} iterations)")
    return {"value": want["output_size"] / dt, "unit": "derived tuples/s", "cores": threads,
            "kind": "port", "sample": sample, "seconds": round(dt, 3),
            "sample_output_tuples": int(want["output_size"]), "gpu_matches_sample": bool(same)}


METRIC = "derived tuples/s, 1 MI355X, bit-exact; batched/served vs plain speedup + p99"


def energy_meter(ctx):
    """Reader of the device's energy accumulator (joules) through amdsmi, matched
    to the context's device by PCI address; None when it is not available.
    Reading it changes nothing on the device."""
    try:
        import amdsmi
        amdsmi.amdsmi_init()
        bdf = ctx.pci_bus_id().lower()
        for h in amdsmi.amdsmi_get_processor_handles():
            if amdsmi.amdsmi_get_gpu_device_bdf(h).lower() == bdf:
                def read(h=h):
                    e = amdsmi.amdsmi_get_energy_count(h)
                    return e["energy_accumulator"] * e["counter_resolution"] * 1e-6
                read()
                return read
    except Exception as e:   # no monitor on this host: the line reports null
        log(f"energy: not available ({type(e).__name__}: {e})")
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rmat_scale_for(m):
    """2^ceil(log2(m / 4)) vertices, raised until m distinct edges fit."""
    import math
    s = max(3, math.ceil(math.log2(max(m / 4, 2))))
    while (2 ** s) * (2 ** s - 1) // 4 < m:
        s += 1
    return s


def c4_instances(count, max_edges, seed_base=SEED):
    """Batched-TC instances with edge counts log-uniform in [1k, max_edges]:
    even i Erdos-Renyi (n = m / d, d ~ U[2, 4]), odd i R-MAT(.57, .19, .19, .05) on
    2^ceil(log2(m / 4)) vertices; seed base + i per instance."""
    import numpy as np

    import mnmgdatalog_amd as dlg
    rng = np.random.default_rng(seed_base)
    out = []
    for i in range(count):
        m = int(round(np.exp(rng.uniform(np.log(1e3), np.log(max_edges)))))
        d = rng.uniform(2.0, 4.0)
        if i % 2 == 0:
            n = max(8, int(round(m / d)))
            out.append(dlg.gen_erdos_renyi(n, min(m, n * (n - 1) // 2), seed_base + i))
        else:
            scale = rmat_scale_for(m)
            out.append(dlg.gen_rmat(scale, m, 0.57, 0.19, 0.19, seed_base + i))
    return out


def bench_batched(args):
    """Batched form vs serial plain runs of the same instances (host rows in, results on host)."""
    import numpy as np

    import mnmgdatalog_amd as dlg
    inst = c4_instances(args.instances, args.max_edges)
    log(f"batched: {len(inst)} instances generated")
    with dlg.Ctx(0) as ctx:
        mid = inst[int(np.argsort([len(r) for r in inst])[len(inst) // 2])]
        tune = ctx.autotune("tc", mid, load_factors=[0.3, 0.5, 0.8], radix_bits=[6, 8],
                            block_sizes=[128, 256, 512])
        kw = {"load_factor": tune["load_factor"], "radix_bits": tune["radix_bits"],
              "block_size": tune["block_size"]}
        ctx.batch_run("tc", inst[:8], keep_output=False, **kw)          # warm the pool
        t0 = time.perf_counter()
        res, st = ctx.batch_run("tc", inst, keep_output=False, **kw)
        t_batch = time.perf_counter() - t0
        log(f"batched: batch run {t_batch:.3f} s")
        t0 = time.perf_counter()
        plain = [ctx.run("tc", r, keep_output=False, **kw) for r in inst]
        t_plain = time.perf_counter() - t0
    same = all(a["output_size"] == b["output_size"] and a["iterations"] == b["iterations"]
               for a, b in zip(res, plain))
    total = sum(r["output_size"] for r in res)
    return {
        "metric": METRIC, "value": total / t_batch, "unit": "derived tuples/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"batched_tc_{args.instances}x_edges_1k_to_{args.max_edges}",
                   "instances": args.instances, "edges_total": int(sum(len(r) for r in inst)),
                   "output_tuples": int(total), "rounds": st["rounds"],
                   "cap": f"edge counts capped at {args.max_edges} instead of 200k: with mean out-degree "
                          "2-4, |TC| ~ 0.6 n^2 would sum to ~1e11 tuples over 512 instances"},
        "plain_value": total / t_plain, "plain_ms": t_plain * 1e3, "speedup_vs_plain": t_plain / t_batch,
        "per_instance_equal_to_plain": same, "autotune": tune,
    }


def bench_served(args):
    """Served form: saturated throughput and Poisson-arrival latency vs serial plain runs."""
    import numpy as np

    import mnmgdatalog_amd as dlg
    rng = np.random.default_rng(SEED + 5)
    reqs = []
    for i in range(args.requests):
        m = int(min(50e6, round(np.exp(rng.normal(np.log(50e3), 1.5)))))
        m = max(m, 16)
        prog = "tc" if i % 2 == 0 else "wcc"
        if (i // 2) % 2 == 0:
            n = max(8, int(round(m / rng.uniform(2.0, 4.0))))
            rows = dlg.gen_erdos_renyi(n, min(m, n * (n - 1) // 2), SEED + 1000 + i)
        else:
            scale = rmat_scale_for(m)
            rows = dlg.gen_rmat(scale, m, 0.57, 0.19, 0.19, SEED + 1000 + i)
        reqs.append((prog, rows))
    log(f"served: {len(reqs)} requests generated")
    resident = dlg.gen_rmat(args.rmat_scale, 16 << args.rmat_scale, 0.57, 0.19, 0.19, SEED + 7,
                            undirected=True)
    log(f"served: resident R-MAT graph generated ({len(resident)} rows)")
    n_resident_q = max(1, args.requests // 16)
    cap = args.max_request_gb << 30
    with dlg.Server(0, workers=args.workers, max_request_bytes=cap) as srv:
        gid = srv.add_graph(resident)
        # refusals are decided at submission from the prediction alone
        admitted, refused = [], 0
        for p, r in reqs:
            try:
                t = srv.submit(p, r)
                admitted.append((p, r, t))
            except dlg.DlgError as e:
                if e.status != dlg.DLG_REFUSED:
                    raise
                refused += 1
        # saturated: everything submitted at once (the loop above), resident queries too
        gts = [srv.submit("wcc", graph=gid) for _ in range(n_resident_q)]
        got = [srv.wait(t) for (_, _, t) in admitted] + [srv.wait(t) for t in gts]
        t_sat = max(g["end_s"] for g in got) - min(g["submit_s"] for g in got)
        log(f"served: saturated run done, {len(got)} requests, {refused} refused, {t_sat:.2f} s")
        total = sum(g["output_size"] for g in got)
        work = [(p, r) for (p, r, _) in admitted] + [("wcc_resident", None)] * n_resident_q
        # serial plain execution of the same requests (one context, one request at a time)
        with dlg.Ctx(0) as ctx:
            plain_s, plain_out = [], []
            res_rows = None
            for p, r in work:
                t1 = time.perf_counter()
                if p == "wcc_resident":
                    out = ctx.run("wcc", resident)
                else:
                    out = ctx.run(p, r)
                plain_s.append(time.perf_counter() - t1)
                plain_out.append(out)
                if len(plain_s) % 20 == 0:
                    log(f"served: plain run {len(plain_s)}/{len(work)}")
        same = all(g["status"] == 0 and g["output_size"] == o["output_size"] and
                   g["iterations"] == o["iterations"] and np.array_equal(g["rows"], o["rows"])
                   for g, o in zip(got, plain_out))
        t_plain = sum(plain_s)
        # Poisson arrivals at 70 % of the serial plain capacity
        lam = 0.7 * len(work) / t_plain
        gaps = rng.exponential(1.0 / lam, size=len(work))
        arrive = np.cumsum(gaps) - gaps[0]
        order = rng.permutation(len(work))
        tickets = [None] * len(work)
        start = time.perf_counter()
        for k, j in enumerate(order):
            dt = arrive[k] - (time.perf_counter() - start)
            if dt > 0:
                time.sleep(dt)
            p, r = work[j]
            tickets[j] = srv.submit("wcc", graph=gid) if p == "wcc_resident" else srv.submit(p, r)
        lat = [srv.wait(t)["latency_s"] for t in tickets]
        log("served: poisson run done")
    # the serial plain system under the same arrivals: FIFO, service times measured above
    fin, plat = 0.0, []
    for k, j in enumerate(order):
        begin = max(arrive[k], fin)
        fin = begin + plain_s[j]
        plat.append(fin - arrive[k])
    pct = lambda a, q: float(np.percentile(np.asarray(a) * 1e3, q))
    return {
        "metric": METRIC, "value": total / t_sat, "unit": "derived tuples/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": t_sat * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "served_tc_wcc_mix", "requests": len(work), "refused": refused,
                   "resident_graph": f"rmat_scale{args.rmat_scale}_{len(resident)}_rows",
                   "workers": args.workers, "max_request_gb": args.max_request_gb,
                   "sizes": "log-normal edges, median 50k, sigma 1.5, max 50M; TC:WCC 1:1; "
                            "ER/R-MAT alternating"},
        "plain_value": total / t_plain, "speedup_vs_plain": t_plain / t_sat,
        "results_equal_alone": same,
        "poisson": {"rate_per_s": lam, "served_p50_ms": pct(lat, 50), "served_p99_ms": pct(lat, 99),
                    "plain_fifo_p50_ms": pct(plat, 50), "plain_fifo_p99_ms": pct(plat, 99)},
    }


def bench_copy(args):
    """HBM copy reference: the unpack kernel over 2^27 tuples (1 GiB read + 1 GiB
    written per launch), timed with HIP events; also the calibration run for
    rocprofv3 FETCH_SIZE / WRITE_SIZE (known bytes per dispatch)."""
    import ctypes

    import mnmgdatalog_amd as dlg
    n = 1 << 27
    with dlg.Ctx(0) as c:
        src, dst = c.alloc(8 * n), c.alloc(8 * n)
        for _ in range(args.warmup):
            c.check(c.L.dlg_unpack_rows(c.h, ctypes.c_void_p(src), n, 0, ctypes.c_void_p(dst)), "unpack")
        c.kernel_stats_reset()
        c.profiling(True)
        for _ in range(args.steps):
            c.check(c.L.dlg_unpack_rows(c.h, ctypes.c_void_p(src), n, 0, ctypes.c_void_p(dst)), "unpack")
        c.check(c.L.dlg_ctx_sync(c.h), "sync")
        c.profiling(False)
        st = c.kernel_stats()["pack"]
        c.free(src)
        c.free(dst)
    ms = st["ms"] / max(st["launches"], 1)
    return {"metric": "HBM copy GB/s (unpack kernel, 1 GiB in + 1 GiB out)", "value": 16.0 * n / (ms * 1e-3) / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic", "config": {"workload": "copy_2^27_tuples", "bytes_read": 8 * n,
                                            "bytes_written": 8 * n}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tc_grid")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--instances", type=int, default=512, help="batched: instance count")
    ap.add_argument("--max-edges", type=int, default=10000, help="batched: largest instance")
    ap.add_argument("--requests", type=int, default=160, help="served: request count")
    ap.add_argument("--workers", type=int, default=4, help="served: slots")
    ap.add_argument("--rmat-scale", type=int, default=23, help="served: resident R-MAT scale")
    ap.add_argument("--max-request-gb", type=int, default=8, help="served: per-request limit")
    ap.add_argument("--strategy", type=int, default=0, help="full-relation layout (0 auto, 1..4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design "
                                   "(hash partition of the join key + all-to-all), not built"}))
        raise SystemExit(2)
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "served":
        print(json.dumps(bench_served(args)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return

    import mnmgdatalog_amd as dlg

    prog, rows, cfg = workload_rows(args.workload)
    cfg["edges"] = int(rows.shape[0])
    n = rows.shape[0]
    with dlg.Ctx(0) as ctx:
        info = ctx.device_info()
        d_rows = ctx.upload(rows)
        res = None
        for _ in range(args.warmup):
            res = ctx.run(prog, None, keep_output=False, device_rows=d_rows, n=n, strategy=args.strategy)
        meter = energy_meter(ctx)
        ctx.kernel_stats_reset()
        ctx.profiling(True)
        ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
        e0 = meter() if meter else None
        t0 = time.perf_counter()
        dev_s = 0.0
        for _ in range(args.steps):
            res = ctx.run(prog, None, keep_output=False, device_rows=d_rows, n=n, strategy=args.strategy)
            dev_s += res["device_seconds"]
        ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
        t1 = time.perf_counter()
        e1 = meter() if meter else None
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
            cpu = cpu_baseline(args.workload, prog, rows, ctx, args.cpu_threads, args.strategy)
        ctx.free(d_rows)
    line = {
        "metric": "derived tuples/s, 1 MI355X, bit-exact; batched/served vs plain speedup + p99",
        "value": value, "unit": "derived tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": dict(cfg, iterations=res["iterations"], output_tuples=int(out_tuples),
                       raw_join_rows=int(res["raw_join_rows"]), strategy=args.strategy),
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": alg_per_launch, "avg_launch_ms": avg_ms,
                     "launches": st["launches"]},
        "kernels": kernels,
        "cpu_baseline": cpu,
        "energy": None if e0 is None else {
            "joules_per_step": (e1 - e0) / args.steps, "avg_watts": (e1 - e0) / elapsed,
            "tuples_per_joule": out_tuples * args.steps / (e1 - e0) if e1 > e0 else None,
            "source": "amdsmi energy accumulator of the device, over the timed steps"},
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
