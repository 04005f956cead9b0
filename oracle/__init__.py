"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The engine never uses it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_int, c_int32, c_int64, c_void_p

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")


class Stats(ctypes.Structure):
    _fields_ = [("iterations", c_int32), ("output_size", c_int64), ("secondary", c_int64),
                ("raw_join_rows", c_int64), ("nodes", c_int64)]


_L = None


def build() -> None:
    subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name in ("orc_tc", "orc_sg", "orc_wcc", "orc_single_join"):
            f = getattr(L, name)
            f.restype = c_int
            f.argtypes = [c_void_p, c_int64, c_int, POINTER(Stats), POINTER(POINTER(c_int32))]
        for name in ("orc_tc_bounded", "orc_sg_bounded"):
            f = getattr(L, name)
            f.restype = c_int
            f.argtypes = [c_void_p, c_int64, c_int, c_int32, POINTER(Stats), POINTER(POINTER(c_int32))]
        L.orc_wcc_fixpoint.restype = c_int
        L.orc_wcc_fixpoint.argtypes = [c_void_p, c_int64, POINTER(Stats), POINTER(POINTER(c_int32))]
        L.orc_generate_rand_rows.restype = c_int
        L.orc_generate_rand_rows.argtypes = [c_int64, c_int32, c_void_p]
        L.orc_free.restype = None
        L.orc_free.argtypes = [c_void_p]
        _L = L
    return _L


def _take(ptr, n: int) -> np.ndarray:
    if n:
        a = np.ctypeslib.as_array(ptr, shape=(2 * n,)).copy().reshape(n, 2)
    else:
        a = np.zeros((0, 2), np.int32)
    lib().orc_free(ctypes.cast(ptr, c_void_p))
    return a


def _run(name: str, rows: np.ndarray, threads: int = 1, max_iterations: int = 0) -> dict:
    rows = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1, 2))
    st, out = Stats(), POINTER(c_int32)()
    L = lib()
    if max_iterations:
        rc = getattr(L, name + "_bounded")(rows.ctypes.data, rows.shape[0], threads, max_iterations,
                                           byref(st), byref(out))
    else:
        rc = getattr(L, name)(rows.ctypes.data, rows.shape[0], threads, byref(st), byref(out))
    if rc != 0:
        raise ValueError(f"{name} rejected its input")
    n_out = st.output_size
    return {"iterations": st.iterations, "output_size": st.output_size, "secondary": st.secondary,
            "raw_join_rows": st.raw_join_rows, "nodes": st.nodes, "rows": _take(out, n_out)}


def tc(rows, threads: int = 1, max_iterations: int = 0) -> dict:
    """TC: rows (src, dst) ordered by (dst, src) -- the reference's one-rank file order."""
    return _run("orc_tc", rows, threads, max_iterations)


def sg(rows, threads: int = 1, max_iterations: int = 0) -> dict:
    """SG: rows (x, y) ordered by (x, y)."""
    return _run("orc_sg", rows, threads, max_iterations)


def wcc(rows, threads: int = 1) -> dict:
    """WCC with the reference's stop rule: rows (label, 0) by label; secondary = largest."""
    return _run("orc_wcc", rows, threads)


def single_join(rows, threads: int = 1) -> dict:
    """Single join: rows (a, c) ordered by (c, a)."""
    return _run("orc_single_join", rows, threads)


def wcc_fixpoint(rows) -> dict:
    """WCC run to the true fixpoint (stop on an empty delta): (node, label) rows."""
    rows = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1, 2))
    st, out = Stats(), POINTER(c_int32)()
    if lib().orc_wcc_fixpoint(rows.ctypes.data, rows.shape[0], byref(st), byref(out)) != 0:
        raise ValueError("orc_wcc_fixpoint rejected its input")
    return {"iterations": st.iterations, "output_size": st.output_size, "secondary": st.secondary,
            "nodes": st.nodes, "rows": _take(out, st.nodes)}


def generate_rand_rows(rows: int, rand_range: int) -> np.ndarray:
    out = np.empty(2 * rows, dtype=np.int32)
    if lib().orc_generate_rand_rows(rows, rand_range, out.ctypes.data) != 0:
        raise ValueError("bad generator arguments")
    return out.reshape(rows, 2)
