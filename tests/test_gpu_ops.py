"""Operator-level parity of the HIP kernels (through the C ABI) against numpy.

Covers the roles of the reference's Thrust calls and join:
sort / sort+unique / set_difference / merge / exclusive_scan, and
get_hash_table + get_join (multiset of join rows, both projections)."""
import ctypes

import numpy as np
import pytest

import mnmgdatalog_amd as dlg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dlg.Ctx(0)
    yield c
    c.close()


def rand_tuples(rng, n, kbits=32, vbits=32, neg=False):
    k = rng.integers(0, 2 ** kbits, size=n, dtype=np.uint64)
    v = rng.integers(0, 2 ** vbits, size=n, dtype=np.uint64)
    if neg:
        k ^= np.uint64(0x80000000)
    return (k << np.uint64(32)) | v


SORT_CASES = [(0, 32, 32), (1, 32, 32), (2, 3, 3), (1000, 17, 17), (4097, 32, 32), (65536, 1, 0),
              (100_003, 20, 12), (1_000_000, 32, 32), (2_500_000, 17, 17), (300_000, 0, 5)]


@pytest.mark.parametrize("n,kb,vb", SORT_CASES)
def test_sort(ctx, n, kb, vb):
    rng = np.random.default_rng(n + kb)
    t = rand_tuples(rng, n, kb, vb)
    assert np.array_equal(ctx.sort(t), np.sort(t))


def test_sort_all_equal_and_presorted(ctx):
    t = np.full(50_000, 0x1234567800000001, np.uint64)
    assert np.array_equal(ctx.sort(t), t)
    s = np.sort(rand_tuples(np.random.default_rng(5), 70_000))
    assert np.array_equal(ctx.sort(s), s)
    assert np.array_equal(ctx.sort(s[::-1].copy()), s)


@pytest.mark.parametrize("n,kb,vb", [(2_097_152, 32, 32), (3_000_001, 9, 3), (20_000_000, 17, 17),
                                     (8_500_000, 32, 0), (4_200_000, 1, 1)])
def test_sort_onesweep_sizes(ctx, n, kb, vb):
    """Large sorts (>= 2^21 keys) take the one-sweep passes with decoupled look-back: equal
    to numpy's sort at several key widths (1 to 8 passes), and stable under presorted and
    reversed input."""
    rng = np.random.default_rng(n + 7 * kb)
    t = rand_tuples(rng, n, kb, vb)
    assert np.array_equal(ctx.sort(t), np.sort(t))
    if n == 3_000_001:
        s = np.sort(t)
        assert np.array_equal(ctx.sort(s), s)
        assert np.array_equal(ctx.sort(s[::-1].copy()), s)
        assert np.array_equal(ctx.sort_unique(t), np.unique(t))


def test_sort_more_than_2_32_keys(ctx):
    """One buffer of 2^32 + 2^20 distinct keys (the one-sweep passes count digits in 64
    bits): 256 shuffled copies of a block b_i = (i * A mod 2^56) << 8 (A odd, so the b_i
    are distinct), copy j with its low byte set to j.  The sorted buffer is sorted(b)
    with each entry repeated as b | 0 .. b | 255, checked chunk by chunk on the host.
    Uses about 70 GB of device memory."""
    if ctx.device_info()["hbm_bytes"] < (120 << 30):
        pytest.skip("needs about 70 GB of device memory")
    m, copies = (1 << 24) + (1 << 12), 256
    n = m * copies
    assert n > 0xffffffff
    A = np.uint64(0x9E3779B97F4A7C15)
    b = ((np.arange(m, dtype=np.uint64) * A) & np.uint64((1 << 56) - 1)) << np.uint64(8)
    order = np.random.default_rng(11).permutation(copies)
    p = ctx.alloc(n * 8)
    try:
        for slot, j in enumerate(order):   # the copies in shuffled order
            blk = b | np.uint64(j)
            ctx.check(ctx.L.dlg_copy_h2d(ctx.h, ctypes.c_void_p(p + slot * m * 8), blk.ctypes.data, blk.nbytes),
                      "h2d")
        del blk
        ctx.check(ctx.L.dlg_sort(ctx.h, ctypes.c_void_p(p), n), "sort")
        sb = np.sort(b)
        low = np.arange(copies, dtype=np.uint64)[None, :]
        rows = 1 << 20   # sorted entries per checked chunk (2^28 keys, 2 GiB)
        for k0 in range(0, m, rows):
            k1 = min(m, k0 + rows)
            got = ctx.download(p + k0 * copies * 8, (k1 - k0) * copies, np.uint64).reshape(-1, copies)
            assert np.array_equal(got, sb[k0:k1, None] | low), f"sorted entries {k0}..{k1}"
    finally:
        ctx.free(p)


def test_sort_unique_more_than_2_32_keys(ctx):
    """Sort + unique over one buffer of 2^32 + 2^20 keys: 256 copies of a block of distinct
    scattered keys (i * A mod 2^64, A odd), so the result is the sorted block.  Uses about
    105 GB of device memory (the keys, the sort's second buffer, the unique output)."""
    if ctx.device_info()["hbm_bytes"] < (160 << 30):
        pytest.skip("needs about 105 GB of device memory")
    m, copies = (1 << 24) + (1 << 12), 256
    n = m * copies
    b = np.arange(m, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    p = ctx.alloc(n * 8)
    try:
        for slot in range(copies):
            ctx.check(ctx.L.dlg_copy_h2d(ctx.h, ctypes.c_void_p(p + slot * m * 8), b.ctypes.data, b.nbytes), "h2d")
        k = ctypes.c_int64()
        ctx.check(ctx.L.dlg_sort_unique(ctx.h, ctypes.c_void_p(p), n, ctypes.byref(k)), "sort_unique")
        assert k.value == m
        assert np.array_equal(ctx.download(p, m, np.uint64), np.sort(b))
    finally:
        ctx.free(p)


@pytest.mark.parametrize("n,kb,vb", SORT_CASES)
def test_sort_unique(ctx, n, kb, vb):
    rng = np.random.default_rng(7 * n + vb)
    t = rand_tuples(rng, n, kb, vb)
    assert np.array_equal(ctx.sort_unique(t), np.unique(t))


@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8])
def test_sort_digit_widths_via_program(ctx, bits):
    rng = np.random.default_rng(bits)
    rows = rng.integers(0, 3000, size=(40_000, 2)).astype(np.int32)
    r = ctx.run("single_join", rows, radix_bits=bits)
    t = dlg.pack(rows)
    R = np.unique(t)
    # brute force of the single join with numpy: (c, a) for R(a, b), R(b, c)
    a, b = R >> np.uint64(32), R & np.uint64(0xffffffff)
    order = np.argsort(a, kind="stable")
    ka = a[order]
    lo, hi = np.searchsorted(ka, b, "left"), np.searchsorted(ka, b, "right")
    cnt = hi - lo
    rep = np.repeat(np.arange(len(R)), cnt)
    idx = np.concatenate([np.arange(l, h) for l, h in zip(lo, hi)]) if cnt.sum() else np.zeros(0, int)
    c = b[order][idx]
    out = np.unique((c << np.uint64(32)) | a[rep])
    assert r["output_size"] == len(out)
    assert np.array_equal(dlg.pack(r["rows"], reverse=True), out)


@pytest.mark.parametrize("na,nb,overlap", [(0, 10, 0.0), (10, 0, 0.0), (1000, 1000, 0.5),
                                          (200_000, 3_000, 0.3), (3_000, 200_000, 0.9),
                                          (1_000_000, 1_000_000, 0.1)])
def test_set_difference_and_merge(ctx, na, nb, overlap):
    rng = np.random.default_rng(na + 3 * nb)
    a = np.unique(rand_tuples(rng, na, 20, 20))
    b = np.unique(np.concatenate([rand_tuples(rng, nb, 20, 20),
                                  rng.choice(a, size=int(overlap * min(len(a), nb)), replace=False)
                                  if len(a) else np.zeros(0, np.uint64)]))
    d = ctx.set_difference(a, b)
    assert np.array_equal(d, np.setdiff1d(a, b))
    m = ctx.merge(a, b)
    assert np.array_equal(m, np.sort(np.concatenate([a, b])))


@pytest.mark.parametrize("na,nb,vals", [(2048, 0, 5), (2047, 1, 5), (1, 4095, 2), (3000, 5000, 3),
                                       (100_000, 70_000, 40), (1_500_000, 600_000, 1 << 40)])
def test_merge_repeated_values(ctx, na, nb, vals):
    """Merges of sorted inputs with long runs of equal values crossing tile edges,
    including the largest u64 value."""
    rng = np.random.default_rng(na * 7 + nb)
    top = np.uint64(0xFFFFFFFFFFFFFFFF)
    a = np.sort(top - rng.integers(0, vals, size=na, dtype=np.uint64))
    b = np.sort(top - rng.integers(0, vals, size=nb, dtype=np.uint64))
    m = ctx.merge(a, b)
    assert np.array_equal(m, np.sort(np.concatenate([a, b])))


@pytest.mark.parametrize("na,nb,vals", [(0, 0, 5), (0, 7, 3), (2048, 0, 5), (2047, 1, 5), (1, 4095, 2),
                                       (3000, 5000, 3), (4096, 4096, 1), (100_000, 70_000, 40),
                                       (1_500_000, 600_000, 1 << 40), (3_000_000, 3_000_000, 1 << 22)])
def test_merge_unique(ctx, na, nb, vals):
    """Merge keeping one copy of each value against numpy's unique of the
    concatenation: runs of equal values inside and across the inputs and
    across tile edges, the largest u64 value included."""
    rng = np.random.default_rng(na * 3 + nb * 5 + vals)
    top = np.uint64(0xFFFFFFFFFFFFFFFF)
    a = np.sort(top - rng.integers(0, vals, size=na, dtype=np.uint64))
    b = np.sort(top - rng.integers(0, vals, size=nb, dtype=np.uint64))
    assert np.array_equal(ctx.merge_unique(a, b), np.unique(np.concatenate([a, b])))


@pytest.mark.parametrize("n", [1, 100, 2048, 2049, 1_000_000, 5_000_001, 40_000_003])
def test_exclusive_scan(ctx, n):
    rng = np.random.default_rng(n)
    c = rng.integers(0, 5000, size=n, dtype=np.uint32)
    offs, tot = ctx.exclusive_scan(c)
    ref = np.concatenate([[0], np.cumsum(c.astype(np.uint64))[:-1]])
    assert tot == int(c.astype(np.uint64).sum())
    assert np.array_equal(offs, ref)


def numpy_join(rel, probe, proj):
    rk, rv = rel >> np.uint64(32), rel & np.uint64(0xffffffff)
    pk, pv = probe >> np.uint64(32), probe & np.uint64(0xffffffff)
    order = np.argsort(rk, kind="stable")
    sk, sv = rk[order], rv[order]
    lo, hi = np.searchsorted(sk, pk, "left"), np.searchsorted(sk, pk, "right")
    cnt = hi - lo
    rep = np.repeat(np.arange(len(probe)), cnt)
    idx = np.concatenate([np.arange(l, h) for l, h in zip(lo, hi)]) if cnt.sum() else np.zeros(0, int)
    hv, pvv = sv[idx], pv[rep]
    out = (hv << np.uint64(32)) | pvv if proj == dlg.PROJ_HV_PV else (pvv << np.uint64(32)) | hv
    return np.sort(out)


@pytest.mark.parametrize("nrel,nprobe,keys,lf", [(0, 100, 50, 0.0), (100, 0, 50, 0.0), (1, 1, 1, 0.0),
                                                (5000, 7000, 300, 0.6), (50_000, 20_000, 40_000, 0.3),
                                                (200_000, 100_000, 2_000, 0.9), (30_000, 30_000, 30, 0.5)])
@pytest.mark.parametrize("proj", [dlg.PROJ_HV_PV, dlg.PROJ_PV_HV])
def test_hash_join_multiset(ctx, nrel, nprobe, keys, lf, proj):
    rng = np.random.default_rng(nrel + nprobe + keys)
    k = rng.integers(0, keys, size=nrel).astype(np.uint64) ^ np.uint64(0x80000000)
    rel = (k << np.uint64(32)) | rng.integers(0, 2 ** 32, size=nrel, dtype=np.uint64)
    if nrel > 10:
        rel[:5] = rel[5:10]                                   # duplicate relation rows
    pk = rng.integers(0, 2 * keys, size=nprobe).astype(np.uint64) ^ np.uint64(0x80000000)
    probe = (pk << np.uint64(32)) | rng.integers(0, 2 ** 32, size=nprobe, dtype=np.uint64)
    got, info = ctx.hash_join(rel, probe, proj, lf)
    assert np.array_equal(np.sort(got), numpy_join(rel, probe, proj))
    assert info["rows"] == nrel
    if nrel:
        assert info["keys"] == len(np.unique(rel >> np.uint64(32)))
        assert info["slots"] & (info["slots"] - 1) == 0 and info["slots"] > info["keys"]


def test_hash_join_extreme_ids(ctx):
    ids = np.array([-2**31, -1, 0, 2**31 - 1, -2], np.int32)
    rows = np.array([(a, b) for a in ids for b in ids], np.int32)
    rel, probe = dlg.pack(rows), dlg.pack(rows, reverse=True)
    got, _ = ctx.hash_join(rel, probe)
    assert np.array_equal(np.sort(got), numpy_join(rel, probe, dlg.PROJ_HV_PV))


def test_invalid_arguments_report_status(ctx):
    k = ctypes.c_int64()
    s = ctx.L.dlg_sort_unique(ctx.h, None, 5, ctypes.byref(k))
    assert s == dlg.DLG_INVALID_INPUT
    with pytest.raises(dlg.DlgError) as e:
        ctx.run("tc", np.zeros((3, 2), np.int32), load_factor=2.0)
    assert e.value.status == dlg.DLG_INVALID_INPUT


@pytest.mark.parametrize("proj", [dlg.PROJ_HV_PV_DROP_EQ, dlg.PROJ_PV_HV_DROP_EQ])
def test_hash_join_drop_eq(ctx, proj):
    """get_join + remove_if(is_key_equal_value) (sg.cu:166-182): the plain join's rows, in the
    same order, minus the rows with key == value."""
    rng = np.random.default_rng(proj)
    rows = rng.integers(0, 300, size=(20_000, 2)).astype(np.int32)
    rows[:50, 1] = rows[:50, 0]                                   # self-loops give diagonal rows
    rel, probe = dlg.pack(rows), dlg.pack(rows)
    with ctx.hash_index(rel) as hx:
        plain = hx.join(probe, proj - 2)
        got = hx.join(probe, proj)
    diag = (plain >> np.uint64(32)) == (plain & np.uint64(0xffffffff))
    assert diag.any()
    assert np.array_equal(got, plain[~diag])
    want = numpy_join(rel, probe, proj - 2)
    want = want[(want >> np.uint64(32)) != (want & np.uint64(0xffffffff))]
    assert np.array_equal(np.sort(got), want)


def symmetric_edges(rows: np.ndarray) -> np.ndarray:
    """E u E^-1, sorted and deduplicated (wcc.cu:140-176)."""
    return np.unique(np.concatenate([dlg.pack(rows), dlg.pack(rows, reverse=True)]))


def numpy_wcc_round(E: np.ndarray, delta: np.ndarray, cc: np.ndarray):
    """The reference's round on sorted tuples: J = delta join E on the node, cc' = per-node least
    label of cc u J, delta' = cc' minus cc."""
    J = numpy_join(E, delta, dlg.PROJ_HV_PV)                       # (m, l)
    allt = np.sort(np.concatenate([cc, J]))
    keys = allt >> np.uint64(32)
    first = np.ones(len(allt), bool)
    first[1:] = keys[1:] != keys[:-1]
    new_cc = allt[first]
    return new_cc, np.setdiff1d(new_cc, cc)


def test_wcc_relax_round_matches_numpy(ctx):
    rng = np.random.default_rng(11)
    rows = rng.integers(-500, 500, size=(3000, 2)).astype(np.int32)
    E = symmetric_edges(rows)
    nodes = np.unique(E >> np.uint64(32))
    # labels: any node id not above the node's own (as the loop keeps them), Δ a random subset
    lab = np.minimum(nodes, rng.choice(nodes, size=len(nodes)))
    cc = (nodes << np.uint64(32)) | (lab & np.uint64(0xffffffff))
    delta = cc[rng.random(len(cc)) < 0.3]
    with ctx.hash_index(E) as hx:
        got_cc, got_d = hx.wcc_relax(delta, cc)
    want_cc, want_d = numpy_wcc_round(E, delta, cc)
    assert np.array_equal(got_cc, want_cc)
    assert np.array_equal(got_d, want_d)


@pytest.mark.parametrize("name", ["paper.bin", "dummy.bin", "data_7035.bin", "data_214078.bin"])
def test_wcc_relax_drives_the_reference_loop(ctx, orc, name):
    """wcc.cu:160-383 written against the C ABI roles (dlg_hash_build of E u E^-1, then
    dlg_wcc_relax per round, stop when two consecutive |Δ| are equal) gives the oracle's
    iterations, components and largest component."""
    from conftest import read_bin
    rows = read_bin(name)
    E = symmetric_edges(rows)
    nodes = np.unique(E >> np.uint64(32))
    cc = (nodes << np.uint64(32)) | nodes
    delta, prev, it = cc.copy(), len(cc), 0
    with ctx.hash_index(E) as hx:
        while True:
            cc, delta = hx.wcc_relax(delta, cc)
            it += 1
            if len(delta) == prev:
                break
            prev = len(delta)
    labels, counts = np.unique(cc & np.uint64(0xffffffff), return_counts=True)
    want = orc.wcc(rows)
    assert it == want["iterations"]
    assert len(labels) == want["output_size"]
    assert int(counts.max()) == want["secondary"]
    assert np.array_equal(dlg.unpack(labels << np.uint64(32) | np.uint64(0x80000000)), want["rows"])


def test_wcc_relax_rejects_bad_input(ctx):
    rows = np.array([[1, 2], [2, 3]], np.int32)
    E = np.unique(dlg.pack(rows))                                   # not symmetric: 3 is no key
    nodes = np.unique(E >> np.uint64(32))
    cc = (nodes << np.uint64(32)) | nodes
    with ctx.hash_index(E) as hx:
        with pytest.raises(dlg.DlgError) as e:
            hx.wcc_relax(cc, cc)
        assert e.value.status == dlg.DLG_INVALID_INPUT
        with pytest.raises(dlg.DlgError) as e:
            hx.wcc_relax(cc, cc[:1])                                # n_cc != key count
        assert e.value.status == dlg.DLG_INVALID_INPUT
