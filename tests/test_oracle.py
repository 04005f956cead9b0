"""Pins of the CPU oracle (oracle/oracle.cpp) against what the reference
published and what the mathematics fixes.  CPU only.

Sources of the expected values:
  README.md:37-45, 56-58, 67 (reference tables), data/data_5_TC.txt and
  data/dummy_TC.txt (golden TC tuple sets), data/cc_process.txt:339 (WCC trace),
  logs/small-merged-job-output.txt and logs/single-join-small-job-output.txt
  (one-rank runs), rough/README.md:446-447 (WCC SF.cedge), rough/compares.md:19.
The brute-force checks below evaluate NAIVELY (all tuples every round, Python
sets) -- a different algorithm from the oracle's semi-naive loop -- and closed
forms of chains, cycles, trees and paths.
"""
from collections import defaultdict

import numpy as np
import pytest

from conftest import read_bin, read_pairs_txt


def as_set(rows):
    return set(map(tuple, np.asarray(rows).tolist()))


# ------------------------------------------------------------------ brute force
def naive_tc(rows):
    E = set(map(tuple, rows))
    succ = defaultdict(set)
    for a, b in E:
        succ[a].add(b)
    S, it = set(E), 0
    while True:
        new = {(x, z) for (x, y) in S for z in succ.get(y, ())}
        it += 1
        if new <= S:
            return S, it
        S |= new


def naive_sg(rows):
    ch = defaultdict(set)
    for p, x in set(map(tuple, rows)):
        ch[p].add(x)
    S = {(x, y) for p in ch for x in ch[p] for y in ch[p] if x != y}
    it = 0
    while True:
        new = {(x, y) for (a, b) in S for x in ch.get(a, ()) for y in ch.get(b, ())}
        it += 1
        if new <= S:
            return S, it
        S |= new


def naive_sj(rows):
    R = set(map(tuple, rows))
    succ = defaultdict(set)
    for a, b in R:
        succ[a].add(b)
    return {(a, c) for (a, b) in R for c in succ.get(b, ())}


def union_find_labels(rows):
    parent = {}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in map(tuple, rows):
        parent.setdefault(a, a)
        parent.setdefault(b, b)
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    comp = defaultdict(list)
    for x in parent:
        comp[find(x)].append(x)
    return {min(v): len(v) for v in comp.values()}


def random_graph(rng, n, m, allow_self=False, neg=False):
    r = rng.integers(-n if neg else 0, n, size=(m, 2)).astype(np.int32)
    if not allow_self:
        r = r[r[:, 0] != r[:, 1]]
    return r


# ------------------------------------------------------------------ TC
@pytest.mark.parametrize("name,its,size", [
    ("hipc_2019.bin", 3, 9), ("data_10.bin", 3, 18), ("data_7035.bin", 64, 146120),
    ("data_23874.bin", 58, 481121), ("data_5.bin", 3, 9), ("dummy.bin", 3, 13),
])
def test_tc_published_counts(orc, name, its, size):
    r = orc.tc(read_bin(name))
    assert (r["iterations"], r["output_size"]) == (its, size)


def test_tc_golden_sets(orc):
    assert as_set(orc.tc(read_bin("data_5.bin"))["rows"]) == read_pairs_txt("data_5_TC.txt")
    assert as_set(orc.tc(read_bin("dummy.bin"))["rows"]) == read_pairs_txt("dummy_TC.txt")


def test_tc_chain_closed_form(orc):
    # data_22: chain 1 -> 2 -> ... -> 23: C(23, 2) pairs, one round per path length
    r = orc.tc(read_bin("data_22.bin"))
    assert (r["iterations"], r["output_size"]) == (22, 253)
    assert as_set(r["rows"]) == {(i, j) for i in range(1, 24) for j in range(i + 1, 24)}


@pytest.mark.parametrize("n", [1, 2, 5, 17])
def test_tc_cycle_closed_form(orc, n):
    rows = np.array([(i, (i + 1) % n) for i in range(n)], np.int32)
    r = orc.tc(rows)
    assert r["output_size"] == n * n
    assert r["iterations"] == n


def test_tc_row_order_is_dst_then_src(orc):
    rows = orc.tc(read_bin("data_7035.bin"))["rows"]
    key = rows[:, 1].astype(np.int64) * (1 << 32) + (rows[:, 0].astype(np.int64) + (1 << 31))
    assert np.all(np.diff(key) > 0)


@pytest.mark.parametrize("seed", range(6))
def test_tc_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    rows = random_graph(rng, 25, 40, allow_self=seed % 2 == 0, neg=seed % 3 == 0)
    rows = np.concatenate([rows, rows[:3]])  # duplicate rows must not matter
    S, its = naive_tc(rows)
    r = orc.tc(rows)
    assert as_set(r["rows"]) == S
    assert r["iterations"] == its


def test_tc_empty(orc):
    r = orc.tc(np.zeros((0, 2), np.int32))
    assert (r["iterations"], r["output_size"]) == (1, 0)


# ------------------------------------------------------------------ SG
@pytest.mark.parametrize("name,its,size", [
    ("hipc_2019.bin", 3, 4), ("data_10.bin", 2, 11), ("data_7035.bin", 56, 285431),
])
def test_sg_published_counts(orc, name, its, size):
    r = orc.sg(read_bin(name))
    assert (r["iterations"], r["output_size"]) == (its, size)


def test_sg_ego_facebook_published(orc):
    # logs/small-merged-job-output.txt:269 -> 13 iterations, 15,018,986 tuples
    r = orc.sg(read_bin("data_88234.bin"), threads=8)
    assert (r["iterations"], r["output_size"]) == (13, 15018986)


@pytest.mark.parametrize("depth", [1, 2, 3, 5])
def test_sg_binary_tree_closed_form(orc, depth):
    rows = np.array([(p, c) for p in range(1, 2 ** depth) for c in (2 * p, 2 * p + 1)], np.int32)
    r = orc.sg(rows)
    assert r["output_size"] == sum((2 ** k) * (2 ** k - 1) for k in range(1, depth + 1))
    assert r["iterations"] == depth


@pytest.mark.parametrize("seed", range(6))
def test_sg_brute_force(orc, seed):
    rng = np.random.default_rng(100 + seed)
    rows = random_graph(rng, 20, 35, allow_self=seed % 2 == 1, neg=seed % 3 == 1)
    S, its = naive_sg(rows)
    r = orc.sg(rows)
    assert as_set(r["rows"]) == S
    assert r["iterations"] == its
    assert np.all(np.diff(r["rows"][:, 0].astype(np.int64) * (1 << 33) + r["rows"][:, 1]) > 0)


# ------------------------------------------------------------------ WCC
@pytest.mark.parametrize("name,its,cc,largest", [
    ("paper.bin", 4, 2, 3),         # data/cc_process.txt:339
    ("dummy.bin", 4, 3, 3),         # README.md:163
    ("data_214078.bin", 12, 547, 56739),     # README.md:67 (loc-Brightkite)
    ("data_223001.bin", 306, 17, 115921),    # rough/README.md:446-447 (SF.cedge; stop rule fires early)
])
def test_wcc_published_counts(orc, name, its, cc, largest):
    r = orc.wcc(read_bin(name))
    assert (r["iterations"], r["output_size"], r["secondary"]) == (its, cc, largest)


def test_wcc_stop_rule_changes_sf_cedge(orc):
    # The true fixpoint of SF.cedge is one component; the reference stops at 306.
    r = orc.wcc_fixpoint(read_bin("data_223001.bin"))
    assert (r["iterations"], r["output_size"], r["secondary"]) == (524, 1, 174956)


@pytest.mark.parametrize("n", [2, 3, 10, 40])
def test_wcc_path_closed_form(orc, n):
    rows = np.array([(i, i + 1) for i in range(n - 1)], np.int32)
    r = orc.wcc(rows)
    assert (r["iterations"], r["output_size"], r["secondary"]) == (n + 1, 1, n)
    assert r["rows"].tolist() == [[0, 0]]


@pytest.mark.parametrize("seed", range(8))
def test_wcc_union_find(orc, seed):
    rng = np.random.default_rng(200 + seed)
    rows = random_graph(rng, 60, 45, allow_self=True, neg=seed % 2 == 0)
    comps = union_find_labels(rows)
    fx = orc.wcc_fixpoint(rows)
    assert fx["output_size"] == len(comps)
    assert fx["secondary"] == max(comps.values())
    r = orc.wcc(rows)
    if r["iterations"] == fx["iterations"] + 1:   # the stop rule did not cut the run short
        assert r["rows"][:, 0].tolist() == sorted(comps)
        assert r["secondary"] == max(comps.values())
    else:
        assert r["iterations"] <= fx["iterations"]


def test_wcc_empty(orc):
    r = orc.wcc(np.zeros((0, 2), np.int32))
    assert (r["iterations"], r["output_size"], r["secondary"]) == (1, 0, -1)


# ------------------------------------------------------------------ single join
def test_single_join_published_10m(orc):
    # logs/single-join-small-job-output.txt:34 (10M rows, range 1M)
    rows = orc.generate_rand_rows(10_000_000, 1_000_000)
    r = orc.single_join(rows, threads=8)
    assert r["output_size"] == 99_998_429


@pytest.mark.parametrize("seed", range(5))
def test_single_join_brute_force(orc, seed):
    rng = np.random.default_rng(300 + seed)
    rows = rng.integers(-20 if seed == 4 else 0, 40, size=(120, 2)).astype(np.int32)
    r = orc.single_join(rows)
    assert as_set(r["rows"]) == naive_sj(rows)
    rr = r["rows"]
    assert np.all(np.diff(rr[:, 1].astype(np.int64) * (1 << 33) + rr[:, 0]) > 0)


def test_rand_rows_are_glibc_rand(orc):
    rows = orc.generate_rand_rows(5, 100)
    assert rows.shape == (5, 2) and rows.min() >= 1 and rows.max() <= 100


# ------------------------------------------------------------------ slow pins (DLG_SLOW=1)
@pytest.mark.slow
@pytest.mark.parametrize("name,its,size", [
    ("data_223001.bin", 287, 80498014),   # logs/small-merged-job-output.txt:108
    ("data_163734.bin", 188, 156120489),  # rough/compares.md:19
])
def test_tc_published_large(orc, name, its, size):
    r = orc.tc(read_bin(name), threads=8)
    assert (r["iterations"], r["output_size"]) == (its, size)


@pytest.mark.slow
def test_sg_ca_hepth_published(orc):
    r = orc.sg(read_bin("data_51971.bin"), threads=8)    # logs/small-merged-job-output.txt:250
    assert (r["iterations"], r["output_size"]) == (9, 74618689)
