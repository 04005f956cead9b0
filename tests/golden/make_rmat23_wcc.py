"""Fixture for WCC on the served mix's resident graph: R-MAT scale 23
(8,388,608 vertices; 134,217,728 undirected edges drawn with a = .57,
b = c = .19, stored in both directions: gen_rmat(23, 16 << 23, .57, .19, .19,
0x4D4E4D47 + 7, undirected)), as the CPU oracle computes it with the
reference's stop rule (wcc.cu:366-383).

Writes rmat23_wcc.json beside this script: iterations, components, largest
component, distinct nodes and the digest of the output rows (label, 0) in file
order -- dlg_result.output_sum's formula, sum of r0*P + r1*Q + r0*r1 + K
(mod 2^64, r0 and r1 as uint32).  As a second check it counts the true
components with scipy's union of the edges; when the stop rule did not cut
the run short the two counts agree.

Takes several minutes and ~20 GB of host memory.
Run from the repository root: python tests/golden/make_rmat23_wcc.py
"""
import json
import os
import sys
import time

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import mnmgdatalog_amd as dlg  # noqa: E402  (host-side generator only)
import oracle  # noqa: E402

SEED = 0x4D4E4D47
P, Q, K, M = 0x9E3779B97F4A7C15, 0xC2B2AE3D27D4EB4F, 0x632BE59BD9B4E019, 2 ** 64


def rows_digest(rows: np.ndarray) -> int:
    r0 = rows[:, 0].astype(np.int64).view(np.uint64) & np.uint64(0xffffffff)
    r1 = rows[:, 1].astype(np.int64).view(np.uint64) & np.uint64(0xffffffff)
    with np.errstate(over="ignore"):
        s = r0 * np.uint64(P) + r1 * np.uint64(Q) + r0 * r1 + np.uint64(K)
        return int(np.sum(s, dtype=np.uint64))


def main():
    t = time.time()
    rows = dlg.gen_rmat(23, 16 << 23, 0.57, 0.19, 0.19, SEED + 7, undirected=True)
    print(f"generated {len(rows)} rows in {time.time() - t:.1f} s", flush=True)
    t = time.time()
    w = oracle.wcc(rows, threads=os.cpu_count() or 1)
    print(f"oracle wcc in {time.time() - t:.1f} s", flush=True)
    n = int(rows.max()) + 1
    a = coo_matrix((np.ones(len(rows), np.int8), (rows[:, 0], rows[:, 1])), shape=(n, n)).tocsr()
    ncomp, lab = connected_components(a, directed=False)
    present = np.zeros(n, bool)
    present[rows[:, 0]] = True
    true_cc = len(np.unique(lab[present]))
    out = {"graph": "gen_rmat(23, 134217728, 0.57, 0.19, 0.19, 0x4D4E4D47 + 7, undirected)",
           "rows": int(len(rows)), "iterations": int(w["iterations"]), "components": int(w["output_size"]),
           "largest": int(w["secondary"]), "nodes": int(w["nodes"]), "output_sum": rows_digest(w["rows"]),
           "true_components": int(true_cc)}
    print(out, flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rmat23_wcc.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
