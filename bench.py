#!/usr/bin/env python3
"""Benchmark: derived tuples/s of the semi-naive fixpoint engine on one MI355X.

Workload (default ``tc_grid``): transitive closure of the road-network-shaped
grid -- 400x400 vertices, each undirected lattice edge kept with p = 0.55 and
stored in both directions, plus shortcut edges (1% of the kept lattice edges)
between vertices within Manhattan distance 20 (354,870 int32 rows).  Its
closure holds 20,691,180,264 tuples (501 iterations): the full relation is
kept as a 160,000 x 160,000-bit map (3.2 GB) and the sorted result (166 GB)
stays resident in HBM.

One step = one whole fixpoint (input rows already resident in HBM, output
relation resident in HBM at the end).  value = |output relation| * steps / time.
Per-kernel durations are measured with HIP events on the engine's stream
during the timed steps; `roofline` reports the kernel with the largest total
time.  `cpu_baseline` times the CPU oracle on a bounded prefix of the same
workload (the first iterations of the same fixpoint) on one thread and on the
host's CPU share, and checks that the GPU produces the identical relation on
that prefix.

The default run also reports, in the same line, the metric's second half:
  batched  512 TC instances (edges log-uniform from 1,000 up to the bound whose
           predicted closures sum to 3e9 tuples, ER / R-MAT alternating) in one
           fixpoint against the same instances run one by one; every instance's
           digest equal to its plain run's;
  served   a request loop over a mix of small TC / WCC requests, and the C5 mix
           (sizes log-normal, median 50k edges, plus WCC queries of a resident
           R-MAT scale-23 graph): throughput against serial plain runs, p50/p99
           latency under Poisson arrivals at 70 % of the plain capacity,
           refusals;
  launch   per-round launches against graph-captured rounds (WCC on the grid)
           and the host share of TC's per-round loop on usroad.
Other workloads (``--workload``): sg_mesh, tc_powerlaw, tc_usroad, wcc_grid,
single_join (10M rows, range 1M), single_join_100k (range 100k),
wcc_rmat23, batched, served, served_c5, launch, copy.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 0x4D4E4D47
METRIC = "derived tuples/s, 1 MI355X, bit-exact; batched/served vs plain speedup + p99"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload_rows(name):
    import mnmgdatalog_amd as dlg
    if name == "tc_grid":
        return "tc", dlg.gen_grid(400, 0.55, 0.01, 20, SEED), {
            "workload": "tc_grid_400x400_p0.55_shortcuts1pct_d20", "program": "tc"}
    if name == "sg_mesh":
        return "sg", dlg.gen_tetmesh(35), {"workload": "sg_tetmesh_35", "cells": 35, "program": "sg"}
    if name == "tc_powerlaw":
        return "tc", dlg.gen_powerlaw(3000, 7000, 2.1, SEED), {"workload": "tc_powerlaw_3000v_7000e",
                                                             "program": "tc"}
    if name == "tc_usroad":
        return "tc", dlg.read_rows(os.path.join(ROOT, "tests/golden/data/data_165435.bin")), {
            "workload": "tc_usroad_165435e", "program": "tc"}
    if name == "wcc_grid":
        return "wcc", dlg.gen_grid(400, 0.55, 0.01, 20, SEED), {"workload": "wcc_grid_400", "program": "wcc"}
    if name == "wcc_rmat23":
        return "wcc", dlg.gen_rmat(23, 16 << 23, 0.57, 0.19, 0.19, SEED + 7, undirected=True), {
            "workload": "wcc_rmat23_134M_undirected", "program": "wcc"}
    if name == "single_join":
        return "single_join", dlg.generate_rand_rows(10_000_000, 1_000_000), {
            "workload": "single_join_10M_range1M", "program": "single_join"}
    if name == "single_join_100k":
        return "single_join", dlg.generate_rand_rows(10_000_000, 100_000), {
            "workload": "single_join_10M_range100k", "program": "single_join"}
    raise SystemExit(f"unknown workload {name}")


# Bounded CPU samples per workload: fixpoint iterations the oracle runs on one
# thread and on the host's CPU share (0: the whole fixpoint).
CPU_SAMPLE_ITERS = {"tc_grid": (16, 30), "sg_mesh": (2, 3), "tc_powerlaw": (0, 0), "tc_usroad": (20, 40),
                    "wcc_grid": (0, 0), "single_join": (0, 0)}


def cpu_share():
    """Threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS when set
    (a node's per-GPU share of the host)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    flags = None
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("CXXFLAGS"):
                    flags = line.split("=", 1)[1].strip()
    except OSError:
        pass
    lscpu = {}
    try:   # lscpu's fields of this host (a child process; the GPU is not touched)
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Vendor ID", "CPU family", "Model", "Stepping", "Socket(s)", "Core(s) per socket",
                     "Thread(s) per core", "CPU max MHz", "L3 cache"):
                lscpu[k] = v
            elif k == "Flags":
                fl = set(v.split())
                lscpu["vector_flags"] = sorted(f for f in fl if f in ("avx", "avx2", "avx512f", "avx512bw",
                                                                       "avx512vl", "bmi2", "fma"))
    except (OSError, ValueError, Exception):
        pass
    return {"model": model, "lscpu": lscpu, "nproc": os.cpu_count(), "cpu_share": cpu_share(),
            "oracle_flags": flags}


def cpu_leg(prog, rows, threads, k):
    import oracle
    fn = {"tc": oracle.tc, "sg": oracle.sg, "wcc": oracle.wcc, "single_join": oracle.single_join}[prog]
    extra = {"max_iterations": k} if k and prog in ("tc", "sg") else {}
    t0 = time.perf_counter()
    want = fn(rows, threads=threads, **extra)
    return want, time.perf_counter() - t0, extra


def cpu_baseline(name, prog, rows, ctx, strategy=0):
    import numpy as np
    k1, kn = CPU_SAMPLE_ITERS.get(name, (0, 0))
    share = cpu_share()
    legs = []
    for threads, k in ((1, k1), (share, kn)):
        want, dt, extra = cpu_leg(prog, rows, threads, k)
        got = ctx.run(prog, rows, keep_output=True, strategy=strategy, **extra)
        same = (got["iterations"] == want["iterations"] and got["output_size"] == want["output_size"]
                and np.array_equal(got["rows"], want["rows"]))
        sample = (f"{prog} on the same input, first {want['iterations']} fixpoint iterations"
                  if extra else f"{prog} on the same input, full fixpoint ({want['iterations']} iterations)")
        legs.append({"value": want["output_size"] / dt, "unit": "derived tuples/s", "cores": threads,
                     "kind": "port", "sample": sample, "seconds": round(dt, 3),
                     "sample_output_tuples": int(want["output_size"]), "gpu_matches_sample": bool(same)})
        log(f"cpu leg: {threads} thread(s), {dt:.1f} s")
    out = dict(legs[-1])
    out["legs"] = legs
    out["host"] = host_info()
    return out


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


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from profiles/traffic.json (rocprofv3
    FETCH_SIZE / WRITE_SIZE runs scaled by the copy-kernel calibration), or None when none was
    measured or the kernel's source changed since (its SHA-256 is stored with the figure)."""
    try:
        with open(TRAFFIC_FILE) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    ent = tab.get(f"{workload}/{kernel}")
    if not ent:
        return None
    try:
        with open(os.path.join(ROOT, ent["source"]), "rb") as f:
            if hashlib.sha256(f.read()).hexdigest() != ent["source_sha256"]:
                return None
    except OSError:
        return None
    return ent["bytes_per_launch"]


def pct(a, q):
    import numpy as np
    return float(np.percentile(np.asarray(a) * 1e3, q)) if len(a) else None


# ---------------------------------------------------------------------- batched
def bench_batched(args):
    """512 TC instances (C4 generators, capped by predicted output) as one batched
    fixpoint against serial plain runs; equality by per-instance digests."""
    import mnmgdatalog_amd as dlg
    from mnmgdatalog_amd import workloads as W
    t0 = time.perf_counter()
    inst, info = W.c4_instances(args.instances, output_cap=args.output_cap, workers=cpu_share())
    log(f"batched: {len(inst)} instances generated ({time.perf_counter() - t0:.1f} s), edges up to "
        f"{info['max_edges']}, predicted {info['predicted_output']:.3e} tuples")
    with dlg.Ctx(0) as ctx:
        mid = inst[sorted(range(len(inst)), key=lambda i: len(inst[i]))[len(inst) // 2]]
        tune = ctx.autotune("tc", mid, load_factors=[0.3, 0.5, 0.8], radix_bits=[6, 8],
                            block_sizes=[128, 256, 512])
        kw = {"load_factor": tune["load_factor"], "radix_bits": tune["radix_bits"],
              "block_size": tune["block_size"]}
        ctx.batch_run("tc", inst, keep_output=False, **kw)                 # warm the pool
        t0 = time.perf_counter()
        res, st = ctx.batch_run("tc", inst, keep_output=False, **kw)
        t_batch = time.perf_counter() - t0
        for r in inst[:4]:
            ctx.run("tc", r, keep_output=False, **kw)
        t0 = time.perf_counter()
        plain = [ctx.run("tc", r, keep_output=False, **kw) for r in inst]
        t_plain = time.perf_counter() - t0
        log(f"batched: batch {t_batch:.3f} s, serial plain {t_plain:.3f} s")
        # equality, outside the timing: per-instance digests of batch and plain runs
        chk, _ = ctx.batch_run("tc", inst, keep_output=False, check_output=1, **kw)
        pchk = [ctx.run("tc", r, keep_output=False, check_output=1, **kw) for r in inst]
    same = all(a["output_size"] == b["output_size"] and a["iterations"] == b["iterations"]
               and a["output_sum"] == b["output_sum"] and a["output_order_breaks"] == 0
               and b["output_order_breaks"] == 0 for a, b in zip(chk, pchk))
    same = same and all(a["output_size"] == b["output_size"] for a, b in zip(res, plain))
    total = sum(r["output_size"] for r in res)
    return {
        "workload": f"batched_tc_{len(inst)}x_edges_1k_to_{info['max_edges']}",
        "instances": len(inst), "edges_total": int(sum(len(r) for r in inst)),
        "output_tuples": int(total), "predicted_output_tuples": info["predicted_output"],
        "output_cap": args.output_cap, "rounds": st["rounds"],
        "cap": f"edges log-uniform in [1000, {info['max_edges']}] instead of [1000, 200000]: the largest "
               f"bound whose closures, predicted by SCC condensation, sum to <= {args.output_cap:.0e} tuples",
        "value": total / t_batch, "ms": t_batch * 1e3, "plain_value": total / t_plain, "plain_ms": t_plain * 1e3,
        "speedup_vs_plain": t_plain / t_batch, "per_instance_digest_equal_to_plain": bool(same),
        "autotune": tune,
    }


# ---------------------------------------------------------------------- served
def served_run(reqs, resident=None, n_resident_q=0, workers=16, max_batch=16, max_request_bytes=0,
               rng_seed=SEED + 5, small_rows=20000):
    """Saturated and Poisson-arrival runs of a request list through the server, and the same
    requests run serially in the plain form.  Results equal to the plain runs (rows)."""
    import numpy as np

    import mnmgdatalog_amd as dlg
    rng = np.random.default_rng(rng_seed)
    with dlg.Server(0, workers=workers, max_batch=max_batch, max_request_bytes=max_request_bytes,
                    small_rows=small_rows) as srv:
        gid = srv.add_graph(resident) if resident is not None else -1
        admitted, refused = [], 0
        for p, r in reqs:
            try:
                admitted.append((p, r, srv.submit(p, r)))
            except dlg.DlgError as e:
                if e.status != dlg.DLG_REFUSED:
                    raise
                refused += 1
        gts = [srv.submit("wcc", graph=gid) for _ in range(n_resident_q)]
        got = [srv.wait(t) for (_, _, t) in admitted] + [srv.wait(t) for t in gts]
        t_sat = max(g["end_s"] for g in got) - min(g["submit_s"] for g in got)
        work = [(p, r) for (p, r, _) in admitted] + [("wcc_resident", None)] * n_resident_q
        log(f"served: saturated run, {len(got)} requests, {refused} refused, {t_sat:.2f} s")
        with dlg.Ctx(0) as ctx:
            for p, r in work[:4]:                                   # warm the pool
                ctx.run("wcc" if p == "wcc_resident" else p, resident if r is None else r, keep_output=False)
            plain_s, plain_out = [], []
            for p, r in work:
                t1 = time.perf_counter()
                out = ctx.run("wcc", resident) if p == "wcc_resident" else ctx.run(p, r)
                plain_s.append(time.perf_counter() - t1)
                plain_out.append(out)
        same = all(g["status"] == 0 and g["output_size"] == o["output_size"] and g["iterations"] == o["iterations"]
                   and np.array_equal(g["rows"], o["rows"]) for g, o in zip(got, plain_out))
        t_plain = sum(plain_s)
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
    fin, plat = 0.0, []
    for k, j in enumerate(order):                      # FIFO plain service, same arrivals
        begin = max(arrive[k], fin)
        fin = begin + plain_s[j]
        plat.append(fin - arrive[k])
    total = sum(g["output_size"] for g in got)
    return {"requests": len(work), "workers": workers, "max_batch": max_batch, "refused": refused,
            "value": total / t_sat, "ms": t_sat * 1e3,
            "plain_value": total / t_plain, "plain_ms": t_plain * 1e3, "speedup_vs_plain": t_plain / t_sat,
            "results_equal_plain": bool(same), "batched_requests": int(sum(g["batch_size"] > 1 for g in got)),
            "poisson": {"rate_per_s": lam, "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                        "plain_fifo_p50_ms": pct(plat, 50), "plain_fifo_p99_ms": pct(plat, 99)}}


def bench_served_small(args):
    from mnmgdatalog_amd import workloads as W
    reqs = W.served_requests(args.small_requests, 2_000, 1.0, 50_000, SEED + 11)
    log(f"served (small): {len(reqs)} requests generated")
    out = served_run(reqs, workers=args.workers, max_batch=args.max_batch)
    out["workload"] = "served_small_tc_wcc"
    out["sizes"] = (f"{len(reqs)} requests, edges log-normal median 2k, sigma 1.0, max 50k; TC:WCC 1:1; "
                    "ER / R-MAT alternating by pairs; all arrive together (saturated), then Poisson")
    return out


def bench_served_c5(args):
    import mnmgdatalog_amd as dlg
    from mnmgdatalog_amd import workloads as W
    t0 = time.perf_counter()
    reqs = W.served_requests(args.requests, 50_000, 1.5, 50_000_000, SEED + 5)
    resident = dlg.gen_rmat(args.rmat_scale, 16 << args.rmat_scale, 0.57, 0.19, 0.19, SEED + 7, undirected=True)
    log(f"served (C5): {len(reqs)} requests and the resident graph generated ({time.perf_counter() - t0:.1f} s)")
    out = served_run(reqs, resident=resident, n_resident_q=max(1, args.requests // 16), workers=args.workers,
                     max_batch=args.max_batch, max_request_bytes=args.max_request_gb << 30)
    out["workload"] = "served_c5_tc_wcc_mix"
    out["resident_graph"] = f"rmat_scale{args.rmat_scale}_{len(resident)}_rows"
    out["sizes"] = ("edges log-normal median 50k, sigma 1.5, max 50M; TC:WCC 1:1; ER / R-MAT alternating by "
                    f"pairs; plus WCC queries of the resident graph; requests predicted above "
                    f"{args.max_request_gb} GB refused")
    return out


# ---------------------------------------------------------------------- launch overhead
def bench_launch(args):
    """Per-round launches against graph-captured rounds: WCC on the grid (strategy 1 launches
    every round's kernels from the host; 2 replays captured rounds), and TC on usroad, whose
    rounds are launched from the host (wall time against the summed kernel time)."""
    import mnmgdatalog_amd as dlg
    out = {}
    with dlg.Ctx(0) as ctx:
        grid = ctx.upload(dlg.gen_grid(400, 0.55, 0.01, 20, SEED))
        n = 354_870
        w = {}
        for s in (1, 2):
            ctx.run("wcc", None, keep_output=False, device_rows=grid, n=n, strategy=s)
            t0 = time.perf_counter()
            for _ in range(args.launch_steps):
                r = ctx.run("wcc", None, keep_output=False, device_rows=grid, n=n, strategy=s)
            w[s] = ((time.perf_counter() - t0) / args.launch_steps * 1e3, r["iterations"])
        ctx.free(grid)
        out["wcc_grid"] = {"iterations": w[1][1], "per_round_launch_ms": w[1][0], "graph_captured_ms": w[2][0],
                           "speedup": w[1][0] / w[2][0], "default": "graph_captured"}
        rows = dlg.read_rows(os.path.join(ROOT, "tests/golden/data/data_165435.bin"))
        d = ctx.upload(rows)
        tc = {}
        for q in (1, 0):            # one round at a time; rounds queued ahead of the host (default)
            ctx.run("tc", None, keep_output=False, device_rows=d, n=len(rows), queue=q)
            ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
            t0 = time.perf_counter()
            for _ in range(args.launch_steps):
                r = ctx.run("tc", None, keep_output=False, device_rows=d, n=len(rows), queue=q)
            ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
            wall = (time.perf_counter() - t0) / args.launch_steps * 1e3
            ctx.kernel_stats_reset()                # kernel time from one profiled run
            ctx.profiling(True)
            ctx.run("tc", None, keep_output=False, device_rows=d, n=len(rows), queue=q)
            ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
            ctx.profiling(False)
            ks = ctx.kernel_stats()
            kern = sum(v["ms"] for v in ks.values())
            launches = sum(v["launches"] for v in ks.values())
            tc[q] = {"ms": wall, "kernel_ms": kern, "host_gap_ms": wall - kern, "launches_per_step": launches,
                     "launches_per_round": launches / r["iterations"]}
        ctx.free(d)
        out["tc_usroad"] = {"iterations": r["iterations"], "one_round_at_a_time": tc[1], "queued_4": tc[0],
                            "speedup": tc[1]["ms"] / tc[0]["ms"], "default": "queued_4"}
    return out


# ---------------------------------------------------------------------- copy
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


# ---------------------------------------------------------------------- plain
def kernel_table(ks):
    return {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                "avg_us": round(v["ms"] / v["launches"] * 1e3, 2),
                "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else 0.0,
                "frac": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if v["ms"] > 0 else 0.0,
                "alg_bytes_per_launch": v["bytes"] / v["launches"]}
            for k, v in ks.items() if v["launches"]}


def bench_plain(args):
    import mnmgdatalog_amd as dlg

    prog, rows, cfg = workload_rows(args.workload)
    cfg["edges"] = int(rows.shape[0])
    n = rows.shape[0]
    with dlg.Ctx(0) as ctx:
        info = ctx.device_info()
        d_rows = ctx.upload(rows)
        res = None
        for _ in range(args.warmup):
            res = ctx.run(prog, None, keep_output=False, device_rows=d_rows, n=n, strategy=args.strategy,
                          relax=args.relax)
        meter = energy_meter(ctx)
        ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
        e0 = meter() if meter else None
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = ctx.run(prog, None, keep_output=False, device_rows=d_rows, n=n, strategy=args.strategy,
                          relax=args.relax)
        ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
        t1 = time.perf_counter()
        e1 = meter() if meter else None
        # per-kernel times from one more step, outside the timed region: the HIP
        # events around every launch cost host time that the timed steps do not pay
        ctx.kernel_stats_reset()
        ctx.profiling(True)
        ctx.run(prog, None, keep_output=False, device_rows=d_rows, n=n, strategy=args.strategy, relax=args.relax)
        ctx.check(ctx.L.dlg_ctx_sync(ctx.h), "sync")
        ctx.profiling(False)
        ks = ctx.kernel_stats()
        elapsed = t1 - t0
        out_tuples = res["output_size"]
        value = out_tuples * args.steps / elapsed
        name, st = max(ks.items(), key=lambda kv: kv[1]["ms"])
        avg_ms = st["ms"] / max(st["launches"], 1)
        alg_per_launch = st["bytes"] / max(st["launches"], 1)
        achieved = alg_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        cpu = None
        if not args.no_cpu:
            cpu = cpu_baseline(args.workload, prog, rows, ctx, args.strategy)
        ctx.free(d_rows)
    joules = None if e0 is None or e1 is None or e1 <= e0 else e1 - e0
    return {
        "metric": METRIC, "value": value, "unit": "derived tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": dict(cfg, iterations=res["iterations"], output_tuples=int(out_tuples),
                       raw_join_rows=int(res["raw_join_rows"]), strategy=args.strategy, relax=args.relax),
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.workload, name),
                     "alg_bytes_per_launch": alg_per_launch, "avg_launch_ms": avg_ms, "launches": st["launches"]},
        "kernels": kernel_table(ks),
        "phases_ms": {k: round(v * 1e3, 3) for k, v in res["phases"].items() if v},   # the last step's
        "cpu_baseline": cpu,
        "energy": None if joules is None else {
            "joules_per_step": joules / args.steps, "avg_watts": joules / elapsed,
            "tuples_per_joule": out_tuples * args.steps / joules,
            "source": "amdsmi energy accumulator of the device, over the timed steps"},
        "device": info,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tc_grid")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload only: skip the batched / served / launch summaries")
    ap.add_argument("--instances", type=int, default=512, help="batched: instance count")
    ap.add_argument("--output-cap", type=float, default=3e9, help="batched: predicted output cap (tuples)")
    ap.add_argument("--requests", type=int, default=48, help="served C5: request count")
    ap.add_argument("--small-requests", type=int, default=512, help="served small mix: request count")
    ap.add_argument("--workers", type=int, default=16, help="served: slots")
    ap.add_argument("--max-batch", type=int, default=16, help="served: instances per batched launch")
    ap.add_argument("--rmat-scale", type=int, default=23, help="served C5: resident R-MAT scale")
    ap.add_argument("--max-request-gb", type=int, default=8, help="served C5: per-request limit")
    ap.add_argument("--launch-steps", type=int, default=3, help="launch: timed runs per variant")
    ap.add_argument("--strategy", type=int, default=0, help="full-relation layout (0 auto, 1..4)")
    ap.add_argument("--relax", type=int, default=0, help="WCC round form (0 auto, 1 push, 2 pull)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process and is measured on one GPU, so "
                                   "--gpus must be 1; sharding across GPUs (hash partition of the join key + "
                                   "all-to-all) is a design, not built"}))
        raise SystemExit(2)
    side = {"batched": bench_batched, "served": bench_served_small, "served_c5": bench_served_c5,
            "launch": bench_launch}
    if args.workload in side:
        print(json.dumps(side[args.workload](args)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    line = bench_plain(args)
    if args.workload == "tc_grid" and not args.no_extras:
        t0 = time.perf_counter()
        line["batched"] = bench_batched(args)
        log(f"batched summary: {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        line["served"] = {"small": bench_served_small(args), "c5": bench_served_c5(args)}
        log(f"served summaries: {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        line["launch"] = bench_launch(args)
        log(f"launch summary: {time.perf_counter() - t0:.1f} s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
