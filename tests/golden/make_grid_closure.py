"""Closed-form facts of the transitive closure of the 400 x 400 road-shaped grid.

The grid (gen_grid(400, 0.55, 0.01, 20, 0x4D4E4D47): lattice edges kept with
p = 0.55 in both directions, 1% shortcut edges within Manhattan distance 20)
is symmetric, so its closure is C x C for every connected component C with at
least two vertices:
  * |TC| = sum |C|^2;
  * the output digest (dlg_result.output_sum) = sum over C of
      P*|C|*S + Q*|C|*S + S*S + K*|C|^2  (mod 2^64), S = sum of the ids in C;
  * the semi-naive iteration count = the largest shortest-path distance
    between two vertices of one component, counting a vertex's distance to
    itself as 2 (the diameter, 501 here) -- found exactly with iFUB
    (Crescenzi et al., "On computing the diameter of real-world undirected
    graphs", TCS 2013): BFS from a central vertex u, then eccentricities of
    the vertices in decreasing distance from u until the lower bound reaches
    twice the current level.
The diameter takes a few minutes of BFS; its value is stored in
grid_closure.json beside this script, with the other two figures.
Run from the repository root: python tests/golden/make_grid_closure.py
"""
import json
import os
import sys

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components, shortest_path

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import mnmgdatalog_amd as dlg  # noqa: E402  (host-side generator only)

SEED = 0x4D4E4D47
P, Q, K, M = 0x9E3779B97F4A7C15, 0xC2B2AE3D27D4EB4F, 0x632BE59BD9B4E019, 2 ** 64


def closure_facts(rows):
    n = int(rows.max()) + 1
    a = coo_matrix((np.ones(len(rows)), (rows[:, 0], rows[:, 1])), shape=(n, n)).tocsr()
    _, lab = connected_components(a, directed=False)
    present = np.bincount(rows.ravel(), minlength=n) > 0
    ids = np.nonzero(present)[0]
    comps = {}
    for u in ids:
        comps.setdefault(int(lab[u]), []).append(int(u))
    size, digest = 0, 0
    for nodes in comps.values():
        k, s = len(nodes), sum(nodes)
        size += k * k
        digest = (digest + P * k * s + Q * k * s + s * s + K * k * k) % M
    return a, comps, size, digest


def diameter(a, comps):
    def dist(u):
        return shortest_path(a, unweighted=True, directed=True, indices=[u])[0]

    best = 0
    for nodes in sorted(comps.values(), key=len, reverse=True):
        if len(nodes) < 2 or len(nodes) - 1 <= best and best >= 2:
            break
        nodes = np.asarray(nodes)
        d1 = dist(nodes[0])
        x = nodes[np.argmax(d1[nodes])]
        dx = dist(x)
        y = nodes[np.argmax(dx[nodes])]
        dy = dist(y)
        lb = max(2, int(dx[nodes].max()), int(dy[nodes].max()))
        u = nodes[np.argmin(np.maximum(dx[nodes], dy[nodes]))]
        du = dist(u)
        i = int(du[nodes].max())
        lb = max(lb, i)
        while i > 0 and lb < 2 * i:   # vertices within i of u are within 2i of each other
            for v in nodes[du[nodes] == i]:
                lb = max(lb, int(dist(v)[nodes].max()))
            i -= 1
        best = max(best, lb)
    return best


def main():
    L = 400
    rows = dlg.gen_grid(L, 0.55, 0.01, 20, SEED)
    a, comps, size, digest = closure_facts(rows)
    d = diameter(a, comps)
    out = {"grid": L, "p": 0.55, "shortcut_frac": 0.01, "shortcut_dist": 20, "seed": SEED,
           "edges": int(len(rows)), "tc_size": size, "tc_output_sum": str(digest), "iterations": d}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "grid_closure.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
