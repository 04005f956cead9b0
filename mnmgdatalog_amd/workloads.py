"""Synthetic workloads of the engine's benchmark configurations, shared by
bench.py and the tests (host-side generation through libdlg.so's generators;
deterministic for a seed).

* ``c4_instances``: batched TC, instance i with m_i edges drawn log-uniform in
  [1,000, M]; even i Erdos-Renyi directed (n = m / d, d ~ U[2, 4]), odd i R-MAT
  (a = .57, b = c = .19, d = .05) on 2^ceil(log2(m / 4)) vertices; seed base + i.
  With an output cap, M is the largest bound up to 200,000 whose instances'
  closures, predicted by SCC condensation (dlg_predict_tc_size), sum to at most
  the cap.
* ``served_requests``: a served mix of TC and WCC requests, sizes log-normal
  (median, sigma, max), graphs from the same two generators.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import gen_erdos_renyi, gen_rmat, predict_tc_size

SEED = 0x4D4E4D47
C4_MIN_EDGES, C4_MAX_EDGES = 1_000, 200_000


def rmat_scale_for(m: int) -> int:
    """2^ceil(log2(m / 4)) vertices, raised until m distinct edges fit."""
    s = max(3, math.ceil(math.log2(max(m / 4, 2))))
    while (2 ** s) * (2 ** s - 1) // 4 < m:
        s += 1
    return s


def er_or_rmat(i: int, m: int, d: float, seed: int) -> np.ndarray:
    """Instance i of the batched mix with m edges (ER for even i, R-MAT for odd i)."""
    if i % 2 == 0:
        n = max(8, int(round(m / d)))
        return gen_erdos_renyi(n, min(m, n * (n - 1) // 2), seed)
    return gen_rmat(rmat_scale_for(m), m, 0.57, 0.19, 0.19, seed)


def _c4_draws(count: int, seed_base: int):
    rng = np.random.default_rng(seed_base)
    u = rng.uniform(0.0, 1.0, size=count)
    d = rng.uniform(2.0, 4.0, size=count)
    return u, d


def c4_edges(u: np.ndarray, max_edges: int) -> np.ndarray:
    lo, hi = math.log(C4_MIN_EDGES), math.log(max(max_edges, C4_MIN_EDGES))
    return np.rint(np.exp(lo + u * (hi - lo))).astype(np.int64)


def _generate(count, max_edges, seed_base, workers):
    u, d = _c4_draws(count, seed_base)
    m = c4_edges(u, max_edges)
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda i: er_or_rmat(i, int(m[i]), float(d[i]), seed_base + i), range(count)))


def predicted_outputs(instances, workers: int = 16) -> list:
    with ThreadPoolExecutor(workers) as ex:
        return [k for k, _ in ex.map(predict_tc_size, instances)]


def c4_instances(count: int = 512, output_cap: float | None = None, max_edges: int = C4_MAX_EDGES,
                 seed_base: int = SEED, workers: int = 16):
    """(instances, info): the batched-TC instances, capped by predicted output when asked.
    info = {"max_edges": M, "predicted_output": sum of predictions}."""
    inst = _generate(count, max_edges, seed_base, workers)
    pred = sum(predicted_outputs(inst, workers))
    if output_cap is None or pred <= output_cap:
        return inst, {"max_edges": max_edges, "predicted_output": int(pred)}
    lo, hi = C4_MIN_EDGES, max_edges           # pred(lo) is small; pred(hi) > cap
    best = None
    for _ in range(12):
        mid = int(round(math.sqrt(lo * hi)))
        if mid in (lo, hi):
            break
        cand = _generate(count, mid, seed_base, workers)
        p = sum(predicted_outputs(cand, workers))
        if p <= output_cap:
            lo, best = mid, (cand, p)
        else:
            hi = mid
    if best is None:
        cand = _generate(count, lo, seed_base, workers)
        best = (cand, sum(predicted_outputs(cand, workers)))
    return best[0], {"max_edges": lo, "predicted_output": int(best[1])}


def served_requests(count: int, median_edges: float, sigma: float, max_edges: int, seed: int,
                    min_edges: int = 16):
    """[(prog, rows)]: TC and WCC alternating, sizes log-normal, ER and R-MAT alternating by pairs."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        m = int(min(max_edges, round(np.exp(rng.normal(np.log(median_edges), sigma)))))
        m = max(m, min_edges)
        d = rng.uniform(2.0, 4.0)
        prog = "tc" if i % 2 == 0 else "wcc"
        kind = 0 if (i // 2) % 2 == 0 else 1   # ER for even pairs of requests, R-MAT for odd ones
        reqs.append((prog, er_or_rmat(kind, m, d, seed + 1000 + i)))
    return reqs
