"""dlg_join_noalloc (the join with all scratch in a static device array) against dlg_join
through the C ABI: same rows in the same order, on two calls in a row, the second smaller
(the array keeps the first call's contents), plus the limits it reports."""
import ctypes
from ctypes import POINTER, byref, c_int, c_int64, c_void_p

import numpy as np
import pytest

import mnmgdatalog_amd as dlg

pytestmark = pytest.mark.gpu

MAX_PROBES = 4_000_000


def bind():
    L = dlg.lib()
    f = L.dlg_join_noalloc
    f.restype = c_int
    f.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_int64, POINTER(c_int64)]
    return f


def rand_rel(rng, n, keys):
    k = rng.integers(0, keys, size=n).astype(np.uint64) ^ np.uint64(0x80000000)
    return (k << np.uint64(32)) | rng.integers(0, 2 ** 32, size=n, dtype=np.uint64)


def noalloc_join(ctx, hx, probe, proj, cap):
    f = bind()
    pp = ctx.upload(probe)
    po = ctx.alloc(8 * max(cap, 1))
    n = c_int64()
    try:
        s = f(ctx.h, hx.h, c_void_p(pp), len(probe), proj, c_void_p(po), cap, byref(n))
        out = ctx.download(po, n.value, np.uint64) if s == dlg.DLG_OK else None
        return s, n.value, out
    finally:
        ctx.free(pp)
        ctx.free(po)


@pytest.mark.parametrize("proj", [dlg.PROJ_HV_PV, dlg.PROJ_PV_HV])
def test_noalloc_join_equals_join_twice(proj):
    rng = np.random.default_rng(31 + proj)
    with dlg.Ctx(0) as ctx:
        rel = rand_rel(rng, 300_000, 50_000)
        with ctx.hash_index(rel) as hx:
            for n_probe in (MAX_PROBES, 123_457):        # the second call is smaller
                probe = rand_rel(rng, n_probe, 60_000)
                want = hx.join(probe, proj)
                s, n, got = noalloc_join(ctx, hx, probe, proj, len(want) + 17)
                assert s == dlg.DLG_OK, ctx.L.dlg_last_error(ctx.h)
                assert n == len(want)
                assert np.array_equal(got, want)


def test_noalloc_join_limits():
    rng = np.random.default_rng(5)
    with dlg.Ctx(0) as ctx:
        rel = rand_rel(rng, 10_000, 100)
        with ctx.hash_index(rel) as hx:
            probe = rand_rel(rng, 5_000, 100)
            want = hx.join(probe)
            s, n, _ = noalloc_join(ctx, hx, probe, dlg.PROJ_HV_PV, len(want) - 1)
            assert s == dlg.DLG_OVERFLOW and n == len(want)
            s, n, got = noalloc_join(ctx, hx, probe, dlg.PROJ_HV_PV, len(want))
            assert s == dlg.DLG_OK and np.array_equal(got, want)
            s, n, got = noalloc_join(ctx, hx, probe[:0], dlg.PROJ_HV_PV, 0)
            assert s == dlg.DLG_OK and n == 0
            f = bind()
            k = c_int64()
            pp, po = ctx.alloc(8 * (MAX_PROBES + 1)), ctx.alloc(64)
            try:
                # beyond the probe limit: refused before any device work
                s = f(ctx.h, hx.h, c_void_p(pp), MAX_PROBES + 1, 0, c_void_p(po), 1, byref(k))
                assert s == dlg.DLG_INVALID_INPUT
                assert b"exceed the limit" in ctx.L.dlg_last_error(ctx.h)
                s = f(ctx.h, hx.h, c_void_p(pp), 10, dlg.PROJ_HV_PV_DROP_EQ, c_void_p(po), 1, byref(k))
                assert s == dlg.DLG_INVALID_INPUT
            finally:
                ctx.free(pp)
                ctx.free(po)
