"""Served form: requests of several programs and sizes in flight together
(continuous batching, several streams, a resident graph) must each give the
oracle's result, and the same result as the request run alone."""
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


def same(got, want, secondary=False):
    assert got["status"] == dlg.DLG_OK
    assert got["iterations"] == want["iterations"]
    assert got["output_size"] == want["output_size"]
    if secondary:
        assert got["secondary"] == want["secondary"]
    assert np.array_equal(got["rows"], want["rows"])


def mixed_requests(seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(60):
        m = int(np.exp(rng.uniform(np.log(50), np.log(3000))))
        if i % 2 == 0:
            rows = dlg.gen_erdos_renyi(max(8, int(m / rng.uniform(2, 4))), m, 1000 + i)
        else:
            scale = max(4, int(np.ceil(np.log2(max(m / 4, 2)))))
            while (2 ** scale) * (2 ** scale - 1) // 4 < m:
                scale += 1
            rows = dlg.gen_rmat(scale, m, 0.57, 0.19, 0.19, 1000 + i)
        prog = ["tc", "wcc", "sg", "tc", "wcc", "single_join"][i % 6]
        reqs.append((prog, rows))
    reqs.append(("tc", read_bin("data_7035.bin")))
    reqs.append(("wcc", read_bin("data_214078.bin")))
    reqs.append(("sg", read_bin("data_7035.bin")))
    return reqs


def test_served_results_equal_alone(ctx, orc):
    reqs = mixed_requests(5)
    resident = dlg.gen_rmat(14, 60000, 0.57, 0.19, 0.19, 77, undirected=True)
    with dlg.Server(0, workers=4, small_rows=5000) as srv:
        gid = srv.add_graph(resident)
        tickets = [srv.submit(p, r) for p, r in reqs]
        gt = [srv.submit("wcc", graph=gid) for _ in range(3)]
        got = [srv.wait(t) for t in tickets]
        ggot = [srv.wait(t) for t in gt]
    for (prog, rows), g in zip(reqs, got):
        same(g, getattr(orc, prog)(rows), secondary=prog == "wcc")    # the oracle
        same(g, ctx.run(prog, rows), secondary=prog == "wcc")         # and the plain run
    alone = ctx.run("wcc", resident)
    for g in ggot:
        same(g, alone, secondary=True)
    assert max(g["batch_size"] for g in got) > 1            # small requests were batched
    assert len({g["slot"] for g in got}) > 1                 # several slots ran concurrently
    for g in got + ggot:
        assert g["submit_s"] <= g["start_s"] <= g["end_s"]


def test_served_refuses_what_cannot_fit(orc):
    """Admission by predicted output: an Erdos-Renyi graph of mean out-degree 4 on 300,000
    vertices closes over ~8.6e10 tuples (SCC condensation) and is refused for TC, not for WCC;
    400,000 random edges over the whole id range close over few tuples and run (the bound
    nodes^2 alone would have refused them)."""
    huge = dlg.gen_erdos_renyi(300_000, 1_200_000, 5)
    assert dlg.predict_tc_size(huge)[0] > 5e10
    rng = np.random.default_rng(1)
    sparse = rng.integers(0, 2**31 - 1, size=(400_000, 2)).astype(np.int32)
    with dlg.Server(0, workers=2) as srv:
        with pytest.raises(dlg.DlgError) as e:
            srv.submit("tc", huge)
        assert e.value.status == dlg.DLG_REFUSED
        t = srv.submit("wcc", huge)                                  # the same rows fit for WCC
        ts = srv.submit("tc", sparse)
        r, rs = srv.wait(t), srv.wait(ts)
    assert r["status"] == dlg.DLG_OK and r["nodes"] > 0
    want = orc.tc(sparse)
    assert rs["status"] == dlg.DLG_OK and rs["output_size"] == want["output_size"]
    assert np.array_equal(rs["rows"], want["rows"])


def test_served_wait_timeout_and_unknown_ticket():
    with dlg.Server(0, workers=1) as srv:
        t = srv.submit("tc", read_bin("data_23874.bin"))
        first = srv.wait(t, timeout=0.0)
        if first is None:                                           # still running
            first = srv.wait(t)
        assert first["status"] == dlg.DLG_OK and first["output_size"] == 481121
        with pytest.raises(dlg.DlgError):
            srv.wait(t)                                             # handed over once


def test_requests_re_run_alone_one_after_another(ctx):
    """Requests that outgrow their reservation are re-run alone, each with the
    whole budget: with one slot they do so one after the other, and every one
    completes with the result of its plain run (the second used to wait for
    good: a request run alone was not counted in the budget in use, so the
    count never returned to zero when it ended)."""
    rng = np.random.default_rng(77)
    reqs = [rng.integers(1, 1000, size=(100_000, 2)).astype(np.int32) for _ in range(3)]
    with dlg.Server(0, workers=1, hbm_budget=2 << 30, retain_bytes=1, strategy=1) as srv:
        tickets = [srv.submit("single_join", r) for r in reqs]
        got = [srv.wait(t, timeout=300.0) for t in tickets]
    for r, g in zip(reqs, got):
        assert g is not None, "request still pending"
        assert g["attempts"] == 2 and g["batch_size"] == 1
        same(g, ctx.run("single_join", r, strategy=1))
