"""Program-level parity on the GPU: TC, SG, WCC and the single join through
the C ABI (and the CLI executables) against the CPU oracle -- bit-exact rows
and equal iteration counts -- and against the reference's published counts at
full size."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import mnmgdatalog_amd as dlg
from conftest import DATA, read_bin

pytestmark = pytest.mark.gpu

SMALL = ["hipc_2019.bin", "data_5.bin", "data_10.bin", "data_22.bin", "dummy.bin", "tc_paper.bin",
         "paper.bin", "quiz3.bin", "data_7035.bin", "data_23874.bin"]


@pytest.fixture(scope="module")
def ctx():
    c = dlg.Ctx(0)
    yield c
    c.close()


def assert_same(got, want, secondary=False):
    assert got["iterations"] == want["iterations"]
    assert got["output_size"] == want["output_size"]
    if secondary:
        assert got["secondary"] == want["secondary"]
    assert got["rows"].shape == want["rows"].shape
    assert np.array_equal(got["rows"], want["rows"])


def rand_rows(seed, n, m, neg=False, self_loops=True, dup=True):
    rng = np.random.default_rng(seed)
    r = rng.integers(-n if neg else 0, n, size=(m, 2)).astype(np.int32)
    if not self_loops:
        r = r[r[:, 0] != r[:, 1]]
    if dup and len(r) > 4:
        r = np.concatenate([r, r[:4]])
    return r


# ------------------------------------------------------------------ fixtures vs oracle
@pytest.mark.parametrize("name", SMALL)
def test_tc_fixture(ctx, orc, name):
    rows = read_bin(name)
    assert_same(ctx.run("tc", rows), orc.tc(rows))


@pytest.mark.parametrize("name", SMALL)
def test_sg_fixture(ctx, orc, name):
    rows = read_bin(name)
    assert_same(ctx.run("sg", rows), orc.sg(rows))


@pytest.mark.parametrize("name", SMALL + ["data_214078.bin", "data_223001.bin", "data_88234.bin"])
def test_wcc_fixture(ctx, orc, name):
    rows = read_bin(name)
    assert_same(ctx.run("wcc", rows), orc.wcc(rows), secondary=True)


@pytest.mark.parametrize("name,its,cc,largest", [
    ("paper.bin", 4, 2, 3), ("dummy.bin", 4, 3, 3), ("data_214078.bin", 12, 547, 56739),
    ("data_223001.bin", 306, 17, 115921)])
def test_wcc_published(ctx, name, its, cc, largest):
    r = ctx.run("wcc", read_bin(name), keep_output=False)
    assert (r["iterations"], r["output_size"], r["secondary"]) == (its, cc, largest)


# ------------------------------------------------------------------ random graphs vs oracle
@pytest.mark.parametrize("seed", range(8))
def test_tc_random(ctx, orc, seed):
    rows = rand_rows(seed, 60 + 40 * seed, 90 + 60 * seed, neg=seed % 2 == 1)
    assert_same(ctx.run("tc", rows), orc.tc(rows))


@pytest.mark.parametrize("seed", range(8))
def test_sg_random(ctx, orc, seed):
    rows = rand_rows(100 + seed, 50 + 30 * seed, 70 + 40 * seed, neg=seed % 2 == 0)
    assert_same(ctx.run("sg", rows), orc.sg(rows))


@pytest.mark.parametrize("seed", range(8))
def test_wcc_random(ctx, orc, seed):
    rows = rand_rows(200 + seed, 400 * (seed + 1), 300 * (seed + 1), neg=seed % 3 == 0)
    assert_same(ctx.run("wcc", rows), orc.wcc(rows), secondary=True)


@pytest.mark.parametrize("seed", range(6))
def test_single_join_random(ctx, orc, seed):
    rng = np.random.default_rng(300 + seed)
    rows = rng.integers(-50 if seed % 2 else 1, 400, size=(3000 * (seed + 1), 2)).astype(np.int32)
    assert_same(ctx.run("single_join", rows), orc.single_join(rows))


@pytest.mark.parametrize("prog", ["tc", "sg", "wcc", "single_join"])
def test_extreme_ids(ctx, orc, prog):
    ids = np.array([-2**31, -2**31 + 1, -1, 0, 1, 2**31 - 2, 2**31 - 1], np.int32)
    rng = np.random.default_rng(9)
    rows = ids[rng.integers(0, len(ids), size=(40, 2))]
    want = getattr(orc, prog)(rows)
    for strategy in (0, 1, 2, 3):
        assert_same(ctx.run(prog, rows, strategy=strategy), want, secondary=prog == "wcc")


@pytest.mark.parametrize("prog,its,size,second", [("tc", 1, 0, 0), ("sg", 1, 0, 0), ("wcc", 1, 0, -1),
                                                  ("single_join", 1, 0, 0)])
def test_empty_input(ctx, prog, its, size, second):
    r = ctx.run(prog, np.zeros((0, 2), np.int32))
    assert (r["iterations"], r["output_size"], r["secondary"]) == (its, size, second)


@pytest.mark.parametrize("lf,bits", [(0.3, 4), (0.6, 6), (0.9, 8), (0.05, 7)])
def test_tuning_does_not_change_results(ctx, orc, lf, bits):
    rows = read_bin("data_23874.bin")
    want = orc.tc(rows)
    assert_same(ctx.run("tc", rows, load_factor=lf, radix_bits=bits), want)


def test_repeated_runs_reuse_pool(ctx, orc):
    rows = read_bin("data_7035.bin")
    want = orc.sg(rows)
    for _ in range(3):
        assert_same(ctx.run("sg", rows), want)


# ------------------------------------------------------------------ configs
def test_config_powerlaw_tc(ctx, orc):
    rows = dlg.gen_powerlaw(3000, 7000, 2.1, 0x4D4E4D47)
    got = ctx.run("tc", rows)
    want = orc.tc(rows)
    assert_same(got, want)
    assert want["output_size"] == 7_778_066 and want["iterations"] == 43


def test_config_tetmesh_sg_small(ctx, orc):
    rows = dlg.gen_tetmesh(6)
    assert_same(ctx.run("sg", rows), orc.sg(rows))


def test_config_grid_wcc(ctx, orc):
    rows = dlg.gen_grid(400, 0.55, 0.01, 20, 0x4D4E4D47)
    assert_same(ctx.run("wcc", rows), orc.wcc(rows), secondary=True)


def capped_degree_rows(seed, n, dmax, dup):
    """A chain 0 -> 1 -> ... -> n-1 (n - 1 rounds) with a few chords, and one
    hub of out-degree exactly dmax (repeated targets when dup: a duplicate
    edge counts towards the degree and is joined twice)."""
    rng = np.random.default_rng(seed)
    e = [(u, u + 1) for u in range(n - 1)]
    hub = n // 2
    tg = rng.integers(0, n, size=dmax - 1)
    if dup:
        tg[1] = tg[0]
    e += [(hub, int(t)) for t in tg]
    for u in rng.choice(np.arange(n - 1), size=n // 10, replace=False):
        if u != hub:
            e.append((int(u), int(rng.integers(0, n))))
    return np.array(e, dtype=np.int32)


@pytest.mark.parametrize("dmax,dup", [(4, False), (8, True), (9, False), (16, True), (17, False), (40, True)])
def test_tc_rows_join_paths(ctx, orc, dmax, dup):
    # max degree <= 8: ELL rows; 9..16: CSR with new-end bounds; above: CSR with exact bounds
    rows = capped_degree_rows(dmax, 1500, dmax, dup)
    assert np.unique(rows[:, 0], return_counts=True)[1].max() == dmax
    want = orc.tc(rows)
    for strategy in (0, 4):
        assert_same(ctx.run("tc", rows, strategy=strategy), want)


@pytest.mark.parametrize("dmax,dup", [(9, False), (16, True), (17, False), (40, True)])
def test_sg_rows_join_paths(ctx, orc, dmax, dup):
    # SG's first join reads 16-wide ELL rows up to out-degree 16, CSR lists above
    rows = capped_degree_rows(dmax, 400, dmax, dup)
    want = orc.sg(rows)
    for strategy in (0, 4):
        got = ctx.run("sg", rows, strategy=strategy)
        assert_same(got, want)
        assert got["raw_join_rows"] == want["raw_join_rows"]


# ------------------------------------------------------------------ published counts, full size
@pytest.mark.parametrize("name,its,size", [
    ("data_223001.bin", 287, 80_498_014),   # logs/small-merged-job-output.txt:108 (SF.cedge)
    ("data_163734.bin", 188, 156_120_489),  # rough/compares.md:19 (fe_body)
])
def test_tc_published_full(ctx, name, its, size):
    r = ctx.run("tc", read_bin(name))
    assert (r["iterations"], r["output_size"]) == (its, size)
    rows = r["rows"]
    key = rows[:, 1].astype(np.int64) * (1 << 32) + rows[:, 0].astype(np.int64)
    assert np.all(np.diff(key) > 0)                                # sorted by (dst, src), no duplicates

    def member(x, y):
        k = y.astype(np.int64) * (1 << 32) + x.astype(np.int64)
        p = np.searchsorted(key, k)
        return (p < len(key)) & (key[np.minimum(p, len(key) - 1)] == k)

    e = read_bin(name)
    assert member(e[:, 0], e[:, 1]).all()                          # E is in TC
    # closure on a sample: tc(x, y), e(y, z) => tc(x, z)
    rng = np.random.default_rng(0)
    smp = rows[rng.integers(0, len(rows), size=20000)]
    es = e[np.argsort(e[:, 0], kind="stable")]
    lo, hi = np.searchsorted(es[:, 0], smp[:, 1], "left"), np.searchsorted(es[:, 0], smp[:, 1], "right")
    xs = np.repeat(smp[:, 0], hi - lo)
    zs = es[np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)]), 1]
    assert member(xs, zs).all()


@pytest.mark.parametrize("name,its,size", [
    ("data_88234.bin", 13, 15_018_986),    # logs/small-merged-job-output.txt:269 (ego-Facebook)
    ("data_51971.bin", 9, 74_618_689),     # logs/small-merged-job-output.txt:250 (CA-HepTh)
])
def test_sg_published_full(ctx, name, its, size):
    r = ctx.run("sg", read_bin(name), keep_output=False)
    assert (r["iterations"], r["output_size"]) == (its, size)


@pytest.mark.parametrize("rows,rng_,out", [
    (10_000_000, 1_000_000, 99_998_429),   # logs/single-join-small-job-output.txt:34
    (10_000_000, 500_000, 199_915_594),    # :58
    (10_000_000, 100_000, 950_725_901),    # :10
])
def test_single_join_published(ctx, rows, rng_, out):
    r = ctx.run("single_join", dlg.generate_rand_rows(rows, rng_), keep_output=False)
    assert r["output_size"] == out
    if rng_ == 1_000_000:
        assert r["raw_join_rows"] == 100_003_467


@pytest.mark.parametrize("lo,hi,n,seed", [(1, 400, 6000, 0), (-50, 400, 3000, 1), (0, 33, 500, 2),
                                         (7, 8, 10, 3), (-2000, 3000, 4000, 4), (0, 64, 20000, 5)])
def test_single_join_map_and_sorted_agree(ctx, orc, lo, hi, n, seed):
    """The map form (each output row built in LDS, read out of the map in
    order) and the row form (each row's candidates sorted in LDS) against the
    sorted raw join (strategy 1) and the oracle: rows, sizes and raw join
    rows, automatic and forced (strategies 4 and 2)."""
    rng = np.random.default_rng(700 + seed)
    rows = rng.integers(lo, hi, size=(n, 2)).astype(np.int32)
    want = orc.single_join(rows)
    runs = [ctx.run("single_join", rows, strategy=s) for s in (0, 1, 2, 4)]
    for r in runs:
        assert_same(r, want)
    assert len({r["raw_join_rows"] for r in runs}) == 1


@pytest.mark.parametrize("lo,hi,n,seed", [(0, 1_000_000, 60_000, 0), (-500_000, 500_000, 200_000, 1),
                                         (0, 300_000, 30_000, 2), (5, 70_000, 400_000, 3)])
def test_single_join_row_form_wide_ids(ctx, orc, lo, hi, n, seed):
    """Id ranges too wide for the map: the row form (automatic) against the
    oracle and the sorted raw join."""
    rng = np.random.default_rng(900 + seed)
    rows = rng.integers(lo, hi, size=(n, 2)).astype(np.int32)
    want = orc.single_join(rows)
    auto, srt = ctx.run("single_join", rows), ctx.run("single_join", rows, strategy=1)
    assert_same(auto, want)
    assert_same(srt, want)
    assert auto["raw_join_rows"] == srt["raw_join_rows"]


def test_single_join_row_form_row_shapes(ctx, orc):
    """Rows of every shape the row form sorts: more than 64 b's with more
    than 256 candidates (bitonic sort in LDS), more than 64 b's with few
    candidates, and small rows (sorted in registers), over an id range too
    wide for the map."""
    rng = np.random.default_rng(21)
    parts = []
    for c in range(60):                      # 100 b's of 3 a's each: 300 candidates
        bs = 10_000 + c * 100 + np.arange(100)
        parts.append(np.stack([bs, np.full(100, c)], 1))
        parts.append(np.stack([rng.integers(200_000, 300_000, 300), np.repeat(bs, 3)], 1))
    for c in range(60, 90):                  # 70 b's of 2 a's each: 140 candidates
        bs = 20_000 + c * 70 + np.arange(70)
        parts.append(np.stack([bs, np.full(70, c)], 1))
        parts.append(np.stack([rng.integers(200_000, 300_000, 140), np.repeat(bs, 2)], 1))
    parts.append(rng.integers(0, 300_000, size=(30_000, 2)))   # small rows
    rows = np.concatenate(parts).astype(np.int32)
    want = orc.single_join(rows)
    for st in (0, 2, 1):
        assert_same(ctx.run("single_join", rows, strategy=st), want)


def test_single_join_row_form_heavy_rows(ctx, orc):
    """A hub whose rows have 3,000 candidates each (above the 2,048 a wave
    sorts in LDS): those rows are materialised, sorted and deduplicated
    together, the others sorted by their waves; same result as the oracle."""
    rng = np.random.default_rng(11)
    hub = np.stack([np.arange(1, 3001), np.zeros(3000, np.int64)], 1)          # (a, 0)
    out = np.stack([np.zeros(5, np.int64), np.arange(3001, 3006)], 1)          # (0, c)
    noise = rng.integers(0, 400_000, size=(20_000, 2))
    rows = np.concatenate([hub, out, noise]).astype(np.int32)
    want = orc.single_join(rows)
    for st in (0, 2):
        assert_same(ctx.run("single_join", rows, strategy=st), want)


@pytest.mark.parametrize("seed", range(3))
def test_single_join_row_form_power_law(ctx, orc, seed):
    """Power-law degrees over a wide id range: many rows above the LDS cap
    beside light ones, against the oracle and the sorted raw join."""
    rng = np.random.default_rng(1300 + seed)
    n = 20_000
    hubs = rng.zipf(1.6, size=(n, 2)).clip(max=400_000) - 1      # heavy-tailed ids, many repeats
    rows = np.where(rng.random((n, 2)) < 0.5, hubs, rng.integers(0, 400_000, size=(n, 2))).astype(np.int32)
    want = orc.single_join(rows)
    auto, srt = ctx.run("single_join", rows), ctx.run("single_join", rows, strategy=1)
    assert_same(auto, want)
    assert_same(srt, want)
    assert auto["raw_join_rows"] == srt["raw_join_rows"]


def test_single_join_map_range_100k_equals_sorted(ctx):
    """10M rows over [1, 100,000] (950,725,901 tuples, logs/single-join-small-job-output.txt:10):
    the map form and the sorted raw join give the same count, raw join rows and
    row digest, both in order."""
    rows = dlg.generate_rand_rows(10_000_000, 100_000)
    a = ctx.run("single_join", rows, keep_output=False, check_output=1)
    b = ctx.run("single_join", rows, keep_output=False, check_output=1, strategy=1)
    assert a["output_size"] == b["output_size"] == 950_725_901
    assert a["raw_join_rows"] == b["raw_join_rows"]
    assert a["output_sum"] == b["output_sum"]
    assert a["output_order_breaks"] == b["output_order_breaks"] == 0


def test_single_join_map_too_wide(ctx):
    rows = np.array([[0, 1], [1, 400_000]], np.int32)
    with pytest.raises(dlg.DlgError):
        ctx.run("single_join", rows, strategy=4)


def test_single_join_10m_rows_equal_oracle(ctx, orc):
    rows = dlg.generate_rand_rows(10_000_000, 1_000_000)
    got = ctx.run("single_join", rows)
    want = orc.single_join(rows, threads=16)
    assert_same(got, want)


# ------------------------------------------------------------------ CLI executables
def run_cli(prog, *args):
    p = subprocess.run([os.path.join(dlg.BIN_DIR, prog), *map(str, args)], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    return p.stdout.strip().splitlines()


@pytest.mark.parametrize("prog,suffix", [("tc", "_tc.bin"), ("sg", "_sg.bin"), ("wcc", "_cc.bin"),
                                         ("tc_nl", "_tc.bin"), ("sg_nl", "_sg.bin")])
@pytest.mark.parametrize("name", ["hipc_2019.bin", "dummy.bin", "data_7035.bin"])
def test_cli_output_file_and_csv(orc, tmp_path, prog, suffix, name):
    src = tmp_path / name
    shutil.copy(os.path.join(DATA, name), src)
    out = run_cli(prog, src, 0, 0, 0)
    assert out[0] == f"Rank 0 wrote local tuples on file: {src}{suffix}"
    want = getattr(orc, prog.replace("_nl", ""))(read_bin(name))
    header, row = out[1].split(","), out[2].split(",")
    assert header[:4] == ["# Input", "# Process", "# Iterations", "# CC" if prog == "wcc" else "# TC"]
    assert len(row) == len(header) == (16 if prog == "wcc" else 15)
    assert int(row[0]) == len(read_bin(name)) and int(row[1]) == 1
    assert int(row[2]) == want["iterations"] and int(row[3]) == want["output_size"]
    if prog == "wcc":
        assert int(row[4]) == want["secondary"]
    data = np.fromfile(f"{src}{suffix}", dtype=np.int32).reshape(-1, 2)
    assert np.array_equal(data, want["rows"])


def test_cli_job_run_1_writes_nothing(tmp_path):
    src = tmp_path / "hipc_2019.bin"
    shutil.copy(os.path.join(DATA, "hipc_2019.bin"), src)
    out = run_cli("tc", src, 0, 0, 1)
    assert len(out) == 2 and not os.path.exists(f"{src}_tc.bin")


def test_cli_single_join():
    out = run_cli("single_join", 100000, 0, 0, 50000)
    header, row = out[0].split(","), out[1].split(",")
    assert header[3] == "# Output" and len(header) == len(row) == 17
    assert int(row[0]) == 100000 and int(row[2]) == 1


def test_cli_missing_file_fails(tmp_path):
    p = subprocess.run([os.path.join(dlg.BIN_DIR, "tc"), str(tmp_path / "nope.bin")], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode != 0 and "Error opening file" in p.stderr


# ------------------------------------------------------------------ sorted-index join (tc_nl / sg_nl)
@pytest.mark.parametrize("name", ["data_7035.bin", "data_23874.bin", "paper.bin"])
def test_sorted_index_join_matches_oracle(ctx, orc, name):
    rows = read_bin(name)
    uniq = np.unique(rows, axis=0)
    want_tc, want_sg = orc.tc(rows), orc.sg(rows)
    for strategy in (0, 1, 3, 4):
        got = ctx.run("tc", rows, join_method=1, strategy=strategy)
        assert_same(got, want_tc)
        # tc_nl.cu probes the deduplicated edges: its |J| is TC's on unique(E)
        assert got["raw_join_rows"] == orc.tc(uniq)["raw_join_rows"]
        got = ctx.run("sg", rows, join_method=1, strategy=strategy)
        assert_same(got, want_sg)
        assert got["raw_join_rows"] == want_sg["raw_join_rows"]
    assert_same(ctx.run("wcc", rows, join_method=1), orc.wcc(rows), secondary=True)
    assert_same(ctx.run("single_join", rows, join_method=1), orc.single_join(rows))


@pytest.mark.parametrize("seed", range(3))
def test_sorted_index_join_random(ctx, orc, seed):
    rows = rand_rows(700 + seed, 200, 400, neg=seed == 1)
    assert_same(ctx.run("tc", rows, join_method=1), orc.tc(rows))
    assert_same(ctx.run("sg", rows, join_method=1), orc.sg(rows))
    assert_same(ctx.run("wcc", rows, join_method=1), orc.wcc(rows), secondary=True)


def test_join_method_rejects_unknown(ctx):
    with pytest.raises(dlg.DlgError) as e:
        ctx.run("tc", read_bin("dummy.bin"), join_method=2)
    assert e.value.status == dlg.DLG_INVALID_INPUT


# ------------------------------------------------------------------ strategies
@pytest.mark.parametrize("name", ["data_223001.bin", "data_214078.bin", "paper.bin", "dummy.bin"])
def test_wcc_graph_captured_and_eager_agree(ctx, orc, name):
    rows = read_bin(name)
    want = orc.wcc(rows)
    for strategy in (1, 2, 0):                       # eager rounds, captured rounds, auto
        assert_same(ctx.run("wcc", rows, strategy=strategy), want, secondary=True)


@pytest.mark.parametrize("name", ["data_223001.bin", "data_214078.bin", "paper.bin", "dummy.bin",
                                  "data_88234.bin"])
def test_wcc_round_forms_agree(ctx, orc, name):
    """Push rounds, pull rounds, and auto (round one off the sorted lists, then pull or push by
    frontier size) give the oracle's labels and iteration count."""
    rows = read_bin(name)
    want = orc.wcc(rows)
    for relax in (1, 2, 0):
        assert_same(ctx.run("wcc", rows, strategy=1, relax=relax), want, secondary=True)


@pytest.mark.parametrize("seed", range(6))
def test_wcc_round_forms_random(ctx, orc, seed):
    rows = rand_rows(700 + seed, 500 * (seed + 1), 400 * (seed + 1), neg=seed % 2 == 0)
    want = orc.wcc(rows)
    for relax in (2, 0):
        assert_same(ctx.run("wcc", rows, strategy=1, relax=relax), want, secondary=True)


def test_wcc_round_forms_truncated(ctx):
    """Stopped after k rounds, every form leaves the same labels (the pull round is the push
    round, not just the same fixpoint)."""
    rows = read_bin("data_214078.bin")
    for k in (1, 2, 3, 5):
        runs = [ctx.run("wcc", rows, strategy=1, relax=relax, max_iterations=k) for relax in (1, 2, 0)]
        for r in runs[1:]:
            assert r["iterations"] == runs[0]["iterations"] == k
            assert r["output_size"] == runs[0]["output_size"]
            assert np.array_equal(r["rows"], runs[0]["rows"])


@pytest.mark.parametrize("name", ["data_7035.bin", "data_23874.bin"])
def test_tc_queued_rounds_agree(ctx, orc, name):
    """Rounds queued ahead of the host (depths 2, 4 and 8; growing outputs stall rounds, which
    are queued again) give the oracle's relation, iteration count and join rows, as one round
    at a time does."""
    rows = read_bin(name)
    want = orc.tc(rows)
    for q in (1, 2, 0, 8):
        assert_same(ctx.run("tc", rows, queue=q), want)


@pytest.mark.parametrize("name,its,size", [
    ("data_223001.bin", 287, 80_498_014),   # logs/small-merged-job-output.txt:108 (SF.cedge)
    ("data_163734.bin", 188, 156_120_489)])  # rough/compares.md:19 (fe_body)
def test_tc_queued_rounds_published(ctx, name, its, size):
    """Long closures, where outputs grow over many rounds (and queued rounds stall): every
    queue depth gives the published size and round count, and the same relation (digest)."""
    rows = read_bin(name)
    runs = [ctx.run("tc", rows, keep_output=False, check_output=1, queue=q) for q in (1, 0, 8)]
    for r in runs:
        assert (r["iterations"], r["output_size"]) == (its, size)
        assert (r["output_sum"], r["output_order_breaks"]) == (runs[0]["output_sum"], 0)
        assert r["raw_join_rows"] == runs[0]["raw_join_rows"]


def test_tc_queued_rounds_truncated(ctx):
    """Stopped after k rounds, no queued round past the k-th changes the relation."""
    rows = read_bin("data_165435.bin")
    for k in (1, 2, 5, 9):
        one = ctx.run("tc", rows, queue=1, max_iterations=k)
        for q in (0, 8):
            r = ctx.run("tc", rows, queue=q, max_iterations=k)
            assert r["iterations"] == one["iterations"] == k
            assert r["raw_join_rows"] == one["raw_join_rows"]
            assert np.array_equal(r["rows"], one["rows"])


def test_tc_queue_out_of_range(ctx):
    with pytest.raises(dlg.DlgError) as e:
        ctx.run("tc", read_bin("dummy.bin"), queue=9)
    assert e.value.status == dlg.DLG_INVALID_INPUT


def test_wcc_presorted_edges(ctx, orc):
    """Above 2^20 rows an edge list that arrives sorted by (src, dst) builds its symmetric
    graph by sorting only the reversed half (by key) and merging; shuffled, by the full sort.
    Both give the oracle's labels, iterations and largest component."""
    rows = dlg.gen_rmat(17, 1 << 20, 0.57, 0.19, 0.19, 31)            # sorted by (src, dst)
    key = rows[:, 0].astype(np.int64) * (1 << 32) + rows[:, 1].astype(np.int64)
    assert rows.shape[0] >= 1 << 20 and np.all(np.diff(key) >= 0)
    want = orc.wcc(rows)
    shuffled = rows[np.random.default_rng(3).permutation(rows.shape[0])]
    for r in (rows, shuffled):
        for strategy in (1, 0):
            assert_same(ctx.run("wcc", r, strategy=strategy), want, secondary=True)
    neg = rows.copy()
    neg[:, 0] -= 70_000                                                # sorted, with negative ids
    neg[:, 1] -= 70_000
    assert_same(ctx.run("wcc", neg, strategy=1), orc.wcc(neg), secondary=True)


def test_wcc_relax_out_of_range(ctx):
    with pytest.raises(dlg.DlgError) as e:
        ctx.run("wcc", read_bin("dummy.bin"), relax=3)
    assert e.value.status == dlg.DLG_INVALID_INPUT


def test_wcc_graph_captured_many_blocks(ctx, orc):
    """Captured rounds on ~1.5M nodes: over 700 scan workgroups, each summing the block
    sums before its own in several strides, against eager rounds and the oracle."""
    rows = dlg.gen_erdos_renyi(1_500_000, 3_000_000, 41)
    want = orc.wcc(rows)
    for strategy in (2, 1):
        assert_same(ctx.run("wcc", rows, strategy=strategy), want, secondary=True)


def test_wcc_graph_captured_max_iterations(ctx):
    rows = read_bin("data_223001.bin")
    a = ctx.run("wcc", rows, strategy=1, max_iterations=37)
    b = ctx.run("wcc", rows, strategy=2, max_iterations=37)
    assert a["iterations"] == b["iterations"] == 37
    assert np.array_equal(a["rows"], b["rows"])


@pytest.mark.parametrize("prog,name", [("tc", "data_23874.bin"), ("sg", "data_7035.bin"),
                                       ("tc", "data_22.bin")])
def test_full_relation_strategies_agree(ctx, orc, prog, name):
    rows = read_bin(name)
    want = getattr(orc, prog)(rows)
    for strategy in (1, 2, 3, 4):                    # merged array, tiered runs, hash set, bitmap
        assert_same(ctx.run(prog, rows, strategy=strategy), want)


@pytest.mark.parametrize("prog", ["tc", "sg"])
def test_max_iterations_prefix_matches_oracle(ctx, orc, prog):
    rows = read_bin("data_7035.bin")
    for k in (1, 2, 5):
        want = getattr(orc, prog)(rows, max_iterations=k)
        for strategy in (1, 2, 3, 4):
            assert_same(ctx.run(prog, rows, strategy=strategy, max_iterations=k), want)


@pytest.mark.parametrize("seed", range(4))
def test_hash_strategy_random(ctx, orc, seed):
    rows = rand_rows(500 + seed, 300, 500, neg=seed % 2 == 0)
    assert_same(ctx.run("tc", rows, strategy=3, load_factor=0.9), orc.tc(rows))
    assert_same(ctx.run("sg", rows, strategy=3, load_factor=0.3), orc.sg(rows))
    assert_same(ctx.run("tc", rows, strategy=4), orc.tc(rows))
    assert_same(ctx.run("sg", rows, strategy=4), orc.sg(rows))


def test_bitmap_strategy_rejects_wide_id_range(ctx):
    rows = np.array([[0, 1], [1, 2**30]], np.int32)
    with pytest.raises(dlg.DlgError) as e:
        ctx.run("tc", rows, strategy=4)
    assert e.value.status == dlg.DLG_INVALID_INPUT
    rows = np.array([[0, 1], [1, 370_720]], np.int32)   # W = 370,721: word indexes past 32 bits
    with pytest.raises(dlg.DlgError) as e:
        ctx.run("tc", rows, strategy=4)
    assert e.value.status == dlg.DLG_INVALID_INPUT


@pytest.mark.parametrize("prog,lo", [("tc", -2**31), ("sg", 2**31 - 370_720)])
def test_bitmap_widest_id_range(ctx, orc, prog, lo):
    """The widest id range the bitmap layout takes, W = 370,720 (370,720 x
    11,585 words, just under 2^32), at either end of the int32 ids."""
    W = 370_720
    rng = np.random.default_rng(5)
    off = np.sort(rng.choice(W, size=2000, replace=False)).astype(np.int64)
    off[0], off[-1] = 0, W - 1                                   # the whole id range
    src = rng.integers(0, len(off) - 1, size=3000)
    dst = np.minimum(src + 1 + rng.integers(0, 30, size=3000), len(off) - 1)   # forward edges: a DAG
    rows = (np.stack([off[src], off[dst]], 1) + lo).astype(np.int32)
    assert_same(ctx.run(prog, rows, strategy=4), getattr(orc, prog)(rows))


# ------------------------------------------------------------------ the road-shaped grid
GRID_SEED = 0x4D4E4D47


@pytest.mark.parametrize("L", [20, 45])
def test_tc_grid_all_layouts(ctx, orc, L):
    rows = dlg.gen_grid(L, 0.55, 0.01, 20, GRID_SEED)
    want = orc.tc(rows, threads=4)
    for s in (0, 1, 2, 3, 4):
        assert_same(ctx.run("tc", rows, strategy=s), want)


@pytest.mark.parametrize("W", [8_200, 20_017])
def test_tc_bitmap_tiled_transpose(ctx, orc, W):
    """Id ranges of more than 256 words per row: the read-out transposes the
    map by tiles of 16 x 16 blocks, with partial tiles at the edges."""
    rng = np.random.default_rng(W)
    ids = np.sort(rng.choice(W, size=1500, replace=False)).astype(np.int32)
    ids[0], ids[-1] = 0, W - 1                                   # the whole id range
    src = rng.integers(0, len(ids) - 1, size=2500)
    dst = np.minimum(src + 1 + rng.integers(0, 40, size=2500), len(ids) - 1)   # forward edges: a DAG
    rows = np.stack([ids[src], ids[dst]], 1).astype(np.int32)
    assert_same(ctx.run("tc", rows, strategy=4), orc.tc(rows))


def test_output_digest_matches_rows(ctx, orc):
    from grid_closure import rows_digest
    rows = read_bin("data_7035.bin")
    for prog in ("tc", "sg", "wcc", "single_join"):
        got = ctx.run(prog, rows, check_output=1)
        assert got["output_sum"] == rows_digest(got["rows"])
        assert got["output_order_breaks"] == 0
    got = ctx.run("tc", rows, strategy=4, check_output=1)
    assert got["output_sum"] == rows_digest(got["rows"]) and got["output_order_breaks"] == 0


def test_tc_grid_400_full_size():
    """TC of the 400 x 400 grid: 20,691,180,264 tuples (166 GB resident),
    501 iterations, against the closed forms (size, digest) and the diameter
    computed independently (grid_closure.json)."""
    import json
    with open(os.path.join(os.path.dirname(DATA), "grid_closure.json")) as f:
        fx = json.load(f)
    rows = dlg.gen_grid(fx["grid"], fx["p"], fx["shortcut_frac"], fx["shortcut_dist"], fx["seed"])
    assert len(rows) == fx["edges"]
    with dlg.Ctx(0) as c:   # its own pool: released when the test ends
        got = c.run("tc", rows, keep_output=False, check_output=1)
    assert got["output_size"] == fx["tc_size"]
    assert got["output_sum"] == int(fx["tc_output_sum"])
    assert got["output_order_breaks"] == 0
    assert got["iterations"] == fx["iterations"]


# ------------------------------------------------------------------ locality order (TC bitmap layout)
@pytest.mark.parametrize("name", ["data_7035.bin", "data_23874.bin", "data_22.bin", "paper.bin",
                                  "hipc_2019.bin"])
def test_tc_locality_order_forced_matches_oracle(ctx, orc, name):
    """reorder = 2 renumbers the ids by recursive BFS bisection before the fixpoint and maps the
    relation back at read-out: rows, size and iterations stay the oracle's."""
    rows = read_bin(name)
    got = ctx.run("tc", rows, reorder=2, strategy=4)
    assert_same(got, orc.tc(rows))
    assert got["raw_join_rows"] == ctx.run("tc", rows, reorder=1, strategy=4)["raw_join_rows"]


@pytest.mark.parametrize("seed", range(4))
def test_tc_locality_order_random(ctx, orc, seed):
    """Negative ids, self-loops, duplicates, isolated ids inside the range, several components."""
    rng = np.random.default_rng(100 + seed)
    parts = [rand_rows(seed, 300, 700, neg=True)]
    parts.append(np.array([(j, j + 1) for j in range(2000, 2400)] + [(2400, 2000)], np.int32))
    parts.append(np.array([[5000, 5000], [5000, 5001], [-7000, 5000]], np.int32))
    rows = np.concatenate(parts)[rng.permutation(sum(len(p) for p in parts))]
    assert_same(ctx.run("tc", rows, reorder=2, strategy=4), orc.tc(rows))
