"""Closed forms of the transitive closure of a symmetric graph, shared by the
CPU pin (test_grid_closure.py) and the GPU full-size test: for every
connected component C with at least two vertices the closure holds C x C."""
import numpy as np

P, Q, K, M = 0x9E3779B97F4A7C15, 0xC2B2AE3D27D4EB4F, 0x632BE59BD9B4E019, 2 ** 64


def components(rows):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    n = int(rows.max()) + 1
    a = coo_matrix((np.ones(len(rows)), (rows[:, 0], rows[:, 1])), shape=(n, n)).tocsr()
    _, lab = connected_components(a, directed=False)
    present = np.nonzero(np.bincount(rows.ravel(), minlength=n) > 0)[0]
    sizes = np.bincount(lab[present])
    sums = np.bincount(lab[present], weights=present.astype(np.float64))   # exact below 2^53
    return [(int(k), int(s)) for k, s in zip(sizes, sums) if k > 0]


def closure_size_and_digest(rows):
    size, digest = 0, 0
    for k, s in components(rows):
        size += k * k
        digest = (digest + P * k * s + Q * k * s + s * s + K * k * k) % M
    return size, digest


def rows_digest(rows):
    """dlg_result.output_sum of host rows (r0, r1): sum of r0*P + r1*Q + r0*r1 + K mod 2^64."""
    r0 = rows[:, 0].astype(np.uint32).astype(np.uint64)
    r1 = rows[:, 1].astype(np.uint32).astype(np.uint64)
    with np.errstate(over="ignore"):
        h = r0 * np.uint64(P) + r1 * np.uint64(Q) + r0 * r1 + np.uint64(K)
        return int(h.sum(dtype=np.uint64))
