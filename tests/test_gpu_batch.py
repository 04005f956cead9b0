"""Batched form: many instances in one fixpoint must give, per instance, the
same rows, sizes and iteration counts as that instance's plain run and the
oracle."""
import numpy as np
import pytest

import mnmgdatalog_amd as dlg
from conftest import read_bin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dlg.Ctx(0)
    yield c
    c.close()


def instances(seed, count, big=False):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 5
        if kind == 0:
            n = int(rng.integers(20, 400 if big else 120))
            out.append(dlg.gen_erdos_renyi(n, int(n * rng.uniform(1.0, 1.8)), seed * 1000 + i))
        elif kind == 1:
            out.append(dlg.gen_rmat(7, int(rng.integers(50, 300)), 0.57, 0.19, 0.19, seed * 1000 + i))
        elif kind == 2:   # arbitrary ids, duplicates, self loops, negatives
            r = rng.integers(-5000, 5000, size=(int(rng.integers(1, 60)), 2)).astype(np.int32)
            out.append(np.concatenate([r, r[:2]]))
        elif kind == 3:
            out.append(np.zeros((0, 2), np.int32))
        else:
            L = int(rng.integers(2, 30))
            out.append(np.array([(j, j + 1) for j in range(L)], np.int32) + int(rng.integers(0, 10**6)))
    return out


def check(got, want, secondary=False):
    assert got["iterations"] == want["iterations"]
    assert got["output_size"] == want["output_size"]
    if secondary:
        assert got["secondary"] == want["secondary"]
    assert np.array_equal(got["rows"], want["rows"])


@pytest.mark.parametrize("prog", ["tc", "sg", "wcc", "single_join"])
def test_batch_matches_oracle(ctx, orc, prog):
    inst = instances(7, 40)
    res, st = ctx.batch_run(prog, inst)
    assert len(res) == len(inst)
    ref = getattr(orc, prog)
    for r, rows in zip(res, inst):
        check(r, ref(rows), secondary=prog == "wcc")
    assert st["total_output"] == sum(r["output_size"] for r in res)


@pytest.mark.parametrize("prog", ["tc", "wcc"])
def test_batch_matches_plain_gpu(ctx, prog):
    inst = instances(11, 25, big=True) + [read_bin("data_7035.bin"), read_bin("paper.bin")]
    res, st = ctx.batch_run(prog, inst)
    for r, rows in zip(res, inst):
        check(r, ctx.run(prog, rows), secondary=prog == "wcc")
    assert st["rounds"] == max(r["iterations"] for r in res)


def test_batch_wcc_stop_rule_per_instance(ctx, orc):
    # SF.cedge stops early (306 rounds) by the reference rule while a long path
    # beside it needs more rounds: each must keep its own count.
    path = np.array([(j, j + 1) for j in range(400)], np.int32)
    inst = [read_bin("data_223001.bin"), path, read_bin("data_214078.bin")]
    res, _ = ctx.batch_run("wcc", inst)
    assert (res[0]["iterations"], res[0]["output_size"], res[0]["secondary"]) == (306, 17, 115921)
    assert (res[1]["iterations"], res[1]["output_size"], res[1]["secondary"]) == (402, 1, 401)
    assert (res[2]["iterations"], res[2]["output_size"], res[2]["secondary"]) == (12, 547, 56739)
    check(res[2], orc.wcc(inst[2]), secondary=True)


@pytest.mark.parametrize("prog", ["tc", "sg", "wcc", "single_join"])
def test_batch_digests_equal_plain(ctx, prog):
    """check_output in the batched form: each instance's digest over its own ids equals its
    plain run's (the per-instance row check of runs too large to download)."""
    inst = instances(13, 30)
    res, _ = ctx.batch_run(prog, inst, keep_output=False, check_output=1)
    for r, rows in zip(res, inst):
        p = ctx.run(prog, rows, keep_output=False, check_output=1)
        assert (r["output_sum"], r["output_order_breaks"]) == (p["output_sum"], p["output_order_breaks"])
        assert r["output_size"] == p["output_size"] and r["output_order_breaks"] == 0


def test_batch_keep_output_off(ctx):
    inst = instances(3, 10)
    res, _ = ctx.batch_run("tc", inst, keep_output=False)
    full, _ = ctx.batch_run("tc", inst)
    assert [r["output_size"] for r in res] == [r["output_size"] for r in full]
    assert all(r["rows"] is None for r in res)


def test_batch_id_space_overflow_is_reported(ctx):
    a = np.array([[-2**31, 2**31 - 1]], np.int32)
    with pytest.raises(dlg.DlgError) as e:
        ctx.batch_run("tc", [a, a])
    assert e.value.status == dlg.DLG_OVERFLOW


@pytest.mark.parametrize("prog", ["tc", "wcc"])
def test_batch_empty_instances_after_the_whole_id_space(ctx, orc, prog):
    """An instance spanning every int32 id, then empty ones: the empties start past the
    last id and must not take the tuples keyed by INT32_MAX."""
    whole = np.array([[-2**31, 2**31 - 1], [2**31 - 1, 5], [5, -2**31]], np.int32)
    empty = np.zeros((0, 2), np.int32)
    inst = [whole, empty, empty]
    res, st = ctx.batch_run(prog, inst)
    for r, rows in zip(res, inst):
        check(r, ctx.run(prog, rows), secondary=prog == "wcc")
    check(res[0], getattr(orc, prog)(whole), secondary=prog == "wcc")
    assert st["total_output"] == res[0]["output_size"]


def test_autotune_picks_a_candidate_and_keeps_results(ctx, orc):
    rows = read_bin("data_23874.bin")
    best = ctx.autotune("tc", rows, load_factors=[0.4, 0.8], radix_bits=[6, 8], block_sizes=[128, 256])
    assert best["load_factor"] in (0.4, 0.8) and best["radix_bits"] in (6, 8)
    assert best["block_size"] in (128, 256) and best["seconds"] > 0
    got = ctx.run("tc", rows, load_factor=best["load_factor"], radix_bits=best["radix_bits"],
                  block_size=best["block_size"])
    check(got, orc.tc(rows))
