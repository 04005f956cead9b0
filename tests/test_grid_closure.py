"""The closed forms used to check the 400 x 400 grid's closure at full size
(166 GB of output, beyond the oracle), pinned against the oracle on smaller
grids of the same generator, and the committed full-size figures re-derived
from the generator."""
import json
import os

import numpy as np
import pytest

import mnmgdatalog_amd as dlg
from conftest import ROOT
from grid_closure import closure_size_and_digest, rows_digest

SEED = 0x4D4E4D47


@pytest.mark.parametrize("L", [12, 25, 40])
def test_closed_form_matches_oracle(orc, L):
    rows = dlg.gen_grid(L, 0.55, 0.01, 20, SEED)
    want = orc.tc(rows, threads=4)
    size, digest = closure_size_and_digest(rows)
    assert want["output_size"] == size
    assert rows_digest(want["rows"]) == digest
    # sorted in the file order of TC rows: by (dst, src)
    key = want["rows"][:, 1].astype(np.int64) * 2 ** 32 + want["rows"][:, 0]
    assert np.all(np.diff(key) > 0)


def test_closed_form_catches_a_missing_tuple(orc):
    rows = dlg.gen_grid(12, 0.55, 0.01, 20, SEED)
    want = orc.tc(rows, threads=2)["rows"]
    _, digest = closure_size_and_digest(rows)
    assert rows_digest(want[1:]) != digest
    bad = want.copy()
    bad[3, 0] += 1
    assert rows_digest(bad) != digest


def test_full_size_fixture_rederived():
    with open(os.path.join(ROOT, "tests", "golden", "grid_closure.json")) as f:
        fx = json.load(f)
    rows = dlg.gen_grid(fx["grid"], fx["p"], fx["shortcut_frac"], fx["shortcut_dist"], fx["seed"])
    assert len(rows) == fx["edges"]
    size, digest = closure_size_and_digest(rows)
    assert size == fx["tc_size"] == 20_691_180_264
    assert digest == int(fx["tc_output_sum"])
