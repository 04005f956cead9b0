"""mnmgdatalog_amd -- MI355X-native semi-naive fixpoint engine (TC, SG, WCC, single join).

The engine is the C ABI library ``libdlg.so`` (HIP kernels for gfx950 + C++
host code, declared in ``include/dlg.h``) and the ``bin/{tc,sg,wcc,single_join}``
executables that replace the reference's ``tc.cu`` / ``sg.cu`` / ``wcc.cu`` /
``single_join.cu`` programs.  This module is a thin ctypes binding of that ABI
for Python harnesses (tests, bench).  It never falls back to a CPU path: if the
library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libdlg.so")
BIN_DIR = os.path.join(PKG_DIR, "bin")

DLG_OK, DLG_INVALID_INPUT, DLG_OUT_OF_MEMORY, DLG_OVERFLOW, DLG_DEVICE_ERROR, DLG_IO_ERROR, DLG_REFUSED = range(7)
STATUS_NAMES = ["OK", "INVALID_INPUT", "OUT_OF_MEMORY", "OVERFLOW", "DEVICE_ERROR", "IO_ERROR", "REFUSED"]

PHASES = ["join", "buffer_preparation", "communication", "deduplication", "merge", "initialization",
          "hashtable", "finalization", "clear", "file_io"]
KERNELS = ["pack", "hash_build", "join_count", "scan", "join_expand", "sort_minmax", "sort_upsweep",
           "sort_downsweep", "select", "merge", "wcc_relax", "wcc_final", "hashset_insert",
           "bitmap_insert", "misc"]

PROJ_HV_PV, PROJ_PV_HV, PROJ_HV_PV_DROP_EQ, PROJ_PV_HV_DROP_EQ = 0, 1, 2, 3
ORIENT_KV, ORIENT_VK = 0, 1


class DlgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        self.status = status
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{name}: {msg}")


class Opts(ctypes.Structure):
    _fields_ = [("load_factor", c_double), ("block_size", c_int32), ("radix_bits", c_int32),
                ("strategy", c_int32), ("max_iterations", c_int32), ("keep_output", c_int32),
                ("check_output", c_int32), ("join_method", c_int32), ("reorder", c_int32),
                ("relax", c_int32), ("queue", c_int32), ("reserved", c_int32 * 1)]


class Result(ctypes.Structure):
    _fields_ = [("iterations", c_int32), ("output_size", c_int64), ("secondary", c_int64),
                ("nodes", c_int64), ("raw_join_rows", c_int64), ("input_rows", c_int64),
                ("rows", POINTER(c_int32)), ("n_rows", c_int64), ("device_seconds", c_double),
                ("phase_seconds", c_double * 10), ("output_sum", ctypes.c_uint64),
                ("output_order_breaks", c_int64)]


class BatchStats(ctypes.Structure):
    _fields_ = [("device_seconds", c_double), ("rounds", c_int32), ("total_output", c_int64),
                ("union_rows", c_int64)]


PROGS = {"tc": 0, "sg": 1, "wcc": 2, "single_join": 3}
DLG_PENDING = 7


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", c_int32), ("workers", c_int32), ("hbm_budget", c_int64), ("max_batch", c_int32),
                ("small_rows", c_int64), ("retain_bytes", c_int64), ("max_request_bytes", c_int64),
                ("tune", Opts)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", c_int32), ("batch_size", c_int32), ("attempts", c_int32), ("slot", c_int32),
                ("submit_s", c_double), ("start_s", c_double), ("end_s", c_double),
                ("reserved_bytes", c_int64)]

_LIB = None


def lib() -> ctypes.CDLL:
    """Load libdlg.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}` "
                           "or __graft_entry__.build(); there is no fallback path")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64 = c_void_p, c_int64
    sig = {
        "dlg_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
        "dlg_ctx_destroy": (None, [vp]),
        "dlg_last_error": (c_char_p, [vp]),
        "dlg_ctx_sync": (c_int, [vp]),
        "dlg_device_info": (c_int, [vp, POINTER(c_int32), POINTER(c_int64), POINTER(c_int32), c_char_p, c_int32]),
        "dlg_device_pci_bus_id": (c_int, [vp, c_char_p, c_int32]),
        "dlg_profiling": (c_int, [vp, c_int]),
        "dlg_kernel_name": (c_char_p, [c_int]),
        "dlg_kernel_stats": (c_int, [vp, c_int, POINTER(c_int64), POINTER(c_double), POINTER(c_double)]),
        "dlg_kernel_stats_reset": (c_int, [vp]),
        "dlg_dev_alloc": (c_int, [vp, c_size_t, POINTER(c_void_p)]),
        "dlg_dev_free": (c_int, [vp, vp]),
        "dlg_copy_h2d": (c_int, [vp, vp, vp, c_size_t]),
        "dlg_copy_d2h": (c_int, [vp, vp, vp, c_size_t]),
        "dlg_pack_rows": (c_int, [vp, vp, i64, c_int, vp]),
        "dlg_unpack_rows": (c_int, [vp, vp, i64, c_int, vp]),
        "dlg_hash_build": (c_int, [vp, vp, i64, c_double, POINTER(c_void_p)]),
        "dlg_hash_info": (c_int, [vp, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64)]),
        "dlg_hash_free": (None, [vp, vp]),
        "dlg_join": (c_int, [vp, vp, vp, i64, c_int, POINTER(c_void_p), POINTER(c_int64)]),
        "dlg_wcc_relax": (c_int, [vp, vp, vp, i64, vp, i64, vp, POINTER(c_int64)]),
        "dlg_sort_unique": (c_int, [vp, vp, i64, POINTER(c_int64)]),
        "dlg_sort": (c_int, [vp, vp, i64]),
        "dlg_set_difference": (c_int, [vp, vp, i64, vp, i64, vp, POINTER(c_int64)]),
        "dlg_merge": (c_int, [vp, vp, i64, vp, i64, vp]),
        "dlg_merge_unique": (c_int, [vp, vp, i64, vp, i64, vp, POINTER(i64)]),
        "dlg_exclusive_scan": (c_int, [vp, vp, i64, vp, POINTER(c_uint64)]),
        "dlg_result_free": (None, [POINTER(Result)]),
        "dlg_read_rows": (c_int, [c_char_p, POINTER(POINTER(c_int32)), POINTER(c_int64)]),
        "dlg_write_rows": (c_int, [c_char_p, vp, i64]),
        "dlg_generate_rand_rows": (c_int, [i64, c_int32, POINTER(POINTER(c_int32))]),
        "dlg_host_free": (None, [vp]),
        "dlg_predict_tc_size": (c_int, [vp, i64, POINTER(c_int64), POINTER(c_int32)]),
        "dlg_gen_powerlaw": (c_int, [c_int32, i64, c_double, c_uint64, POINTER(POINTER(c_int32)), POINTER(c_int64)]),
        "dlg_gen_grid": (c_int, [c_int32, c_double, c_double, c_int32, c_uint64, POINTER(POINTER(c_int32)), POINTER(c_int64)]),
        "dlg_gen_tetmesh": (c_int, [c_int32, POINTER(POINTER(c_int32)), POINTER(c_int64)]),
        "dlg_gen_erdos_renyi": (c_int, [c_int32, i64, c_uint64, POINTER(POINTER(c_int32)), POINTER(c_int64)]),
        "dlg_gen_rmat": (c_int, [c_int32, i64, c_double, c_double, c_double, c_uint64, c_int, POINTER(POINTER(c_int32)), POINTER(c_int64)]),
    }
    sig["dlg_batch_run"] = (c_int, [vp, c_int, c_int32, POINTER(c_void_p), POINTER(c_int64), POINTER(Opts),
                                     POINTER(Result), POINTER(BatchStats)])
    sig["dlg_autotune"] = (c_int, [vp, c_int, vp, i64, POINTER(c_double), POINTER(c_int32), POINTER(c_int32),
                                    c_int32, POINTER(Opts), POINTER(c_double)])
    sig["dlg_server_create"] = (c_int, [POINTER(ServerOpts), POINTER(c_void_p)])
    sig["dlg_server_add_graph"] = (c_int, [vp, vp, i64, POINTER(c_int32)])
    sig["dlg_server_submit"] = (c_int, [vp, c_int, vp, i64, c_int32, POINTER(c_uint64)])
    sig["dlg_server_wait"] = (c_int, [vp, c_uint64, c_double, POINTER(Result), POINTER(RequestInfo)])
    sig["dlg_server_destroy"] = (None, [vp])
    sig["dlg_server_error"] = (c_char_p, [vp])
    for prog in ("tc", "sg", "wcc", "single_join"):
        sig[f"dlg_run_{prog}"] = (c_int, [vp, vp, i64, POINTER(Opts), POINTER(Result)])
        sig[f"dlg_run_{prog}_device"] = (c_int, [vp, vp, i64, POINTER(Opts), POINTER(Result)])
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


class _HostBlock:
    """Owner of a host block malloc'd by libdlg: freed with dlg_host_free when the
    last numpy view of it goes away."""

    def __init__(self, addr: int):
        self.addr = addr

    def __del__(self):
        if _LIB is not None and self.addr:
            _LIB.dlg_host_free(c_void_p(self.addr))
            self.addr = 0


def _host_rows(ptr, n: int) -> np.ndarray:
    """(n, 2) int32 view of a malloc'd row block (a pointer or its address), which it takes
    over (no copy)."""
    addr = ptr if isinstance(ptr, int) else ctypes.cast(ptr, c_void_p).value
    if n == 0 or not addr:
        if addr:
            lib().dlg_host_free(c_void_p(addr))
        return np.zeros((0, 2), dtype=np.int32)
    buf = (c_int32 * (2 * n)).from_address(addr)
    buf._owner = _HostBlock(addr)
    return np.frombuffer(buf, dtype=np.int32).reshape(n, 2)


def _adopt_rows(r) -> np.ndarray | None:
    """The rows of a dlg_result as a numpy view that owns them (the result no longer does)."""
    if not r.rows:
        return None
    addr, n = ctypes.cast(r.rows, c_void_p).value, r.n_rows   # the address, before the field is cleared
    r.rows = POINTER(c_int32)()
    r.n_rows = 0
    return _host_rows(addr, n)


# ----------------------------------------------------------------- formats / generators
def read_rows(path: str) -> np.ndarray:
    p, n = POINTER(c_int32)(), c_int64()
    s = lib().dlg_read_rows(path.encode(), byref(p), byref(n))
    if s != DLG_OK:
        raise DlgError(s, f"read {path}")
    return _host_rows(p, n.value)


def write_rows(path: str, rows: np.ndarray) -> None:
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    s = lib().dlg_write_rows(path.encode(), rows.ctypes.data, rows.shape[0])
    if s != DLG_OK:
        raise DlgError(s, f"write {path}")


def generate_rand_rows(rows: int, rand_range: int) -> np.ndarray:
    p = POINTER(c_int32)()
    s = lib().dlg_generate_rand_rows(rows, rand_range, byref(p))
    if s != DLG_OK:
        raise DlgError(s, "generate_rand_rows")
    return _host_rows(p, rows)


def _gen(fn, *args) -> np.ndarray:
    p, n = POINTER(c_int32)(), c_int64()
    s = fn(*args, byref(p), byref(n))
    if s != DLG_OK:
        raise DlgError(s, fn.__name__)
    return _host_rows(p, n.value)


def predict_tc_size(rows: np.ndarray):
    """|TC| of the rows by SCC condensation (host only): (size, exact)."""
    rows = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1, 2))
    k, ex = c_int64(), c_int32()
    s = lib().dlg_predict_tc_size(rows.ctypes.data, rows.shape[0], byref(k), byref(ex))
    if s != DLG_OK:
        raise DlgError(s, "predict_tc_size")
    return k.value, bool(ex.value)


def gen_powerlaw(n: int, m: int, alpha: float, seed: int) -> np.ndarray:
    return _gen(lib().dlg_gen_powerlaw, n, m, alpha, seed)


def gen_grid(side: int, p: float, shortcut_frac: float, max_dist: int, seed: int) -> np.ndarray:
    return _gen(lib().dlg_gen_grid, side, p, shortcut_frac, max_dist, seed)


def gen_tetmesh(cells: int) -> np.ndarray:
    return _gen(lib().dlg_gen_tetmesh, cells)


def gen_erdos_renyi(n: int, m: int, seed: int) -> np.ndarray:
    return _gen(lib().dlg_gen_erdos_renyi, n, m, seed)


def gen_rmat(scale: int, m: int, a: float, b: float, c: float, seed: int, undirected: bool = False) -> np.ndarray:
    return _gen(lib().dlg_gen_rmat, scale, m, a, b, c, seed, 1 if undirected else 0)


# ----------------------------------------------------------------- tuples
def pack(rows: np.ndarray, reverse: bool = False) -> np.ndarray:
    """int32 (n, 2) rows -> packed uint64 tuples (the library's order-preserving layout)."""
    rows = np.asarray(rows, dtype=np.int32).reshape(-1, 2)
    a = rows[:, 1] if reverse else rows[:, 0]
    b = rows[:, 0] if reverse else rows[:, 1]
    ka = (a.view(np.uint32) ^ np.uint32(0x80000000)).astype(np.uint64)
    kb = (b.view(np.uint32) ^ np.uint32(0x80000000)).astype(np.uint64)
    return (ka << np.uint64(32)) | kb


def unpack(t: np.ndarray) -> np.ndarray:
    t = np.asarray(t, dtype=np.uint64)
    k = ((t >> np.uint64(32)).astype(np.uint32) ^ np.uint32(0x80000000)).view(np.int32)
    v = ((t & np.uint64(0xffffffff)).astype(np.uint32) ^ np.uint32(0x80000000)).view(np.int32)
    return np.stack([k, v], axis=1)


# ----------------------------------------------------------------- context
class Ctx:
    """One device context (HIP stream + memory pool)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        self.h = c_void_p()
        s = self.L.dlg_ctx_create(device, byref(self.h))
        if s != DLG_OK:
            raise DlgError(s, "dlg_ctx_create (is a GPU visible?)")

    def close(self):
        if self.h:
            self.L.dlg_ctx_destroy(self.h)
            self.h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, s: int, what: str):
        if s != DLG_OK:
            raise DlgError(s, f"{what}: {self.L.dlg_last_error(self.h).decode(errors='replace')}")

    def device_info(self) -> dict:
        cu, hbm, lds = c_int32(), c_int64(), c_int32()
        arch = ctypes.create_string_buffer(64)
        self.check(self.L.dlg_device_info(self.h, byref(cu), byref(hbm), byref(lds), arch, 64), "device_info")
        return {"num_cus": cu.value, "hbm_bytes": hbm.value, "lds_per_block": lds.value,
                "arch": arch.value.decode()}

    def pci_bus_id(self) -> str:
        b = ctypes.create_string_buffer(64)
        self.check(self.L.dlg_device_pci_bus_id(self.h, b, 64), "device_pci_bus_id")
        return b.value.decode()

    # ---- device memory
    def alloc(self, nbytes: int) -> int:
        p = c_void_p()
        self.check(self.L.dlg_dev_alloc(self.h, max(nbytes, 1), byref(p)), "dev_alloc")
        return p.value

    def free(self, p: int):
        if p:
            self.check(self.L.dlg_dev_free(self.h, c_void_p(p)), "dev_free")

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.check(self.L.dlg_copy_h2d(self.h, c_void_p(p), arr.ctypes.data, arr.nbytes), "h2d")
        return p

    def download(self, p: int, n: int, dtype) -> np.ndarray:
        out = np.empty(n, dtype=dtype)
        if n:
            self.check(self.L.dlg_copy_d2h(self.h, out.ctypes.data, c_void_p(p), out.nbytes), "d2h")
        return out

    # ---- operators (device tuples in, numpy out)
    def sort_unique(self, t: np.ndarray) -> np.ndarray:
        p = self.upload(np.asarray(t, dtype=np.uint64))
        k = c_int64()
        try:
            self.check(self.L.dlg_sort_unique(self.h, c_void_p(p), len(t), byref(k)), "sort_unique")
            return self.download(p, k.value, np.uint64)
        finally:
            self.free(p)

    def sort(self, t: np.ndarray) -> np.ndarray:
        p = self.upload(np.asarray(t, dtype=np.uint64))
        try:
            self.check(self.L.dlg_sort(self.h, c_void_p(p), len(t)), "sort")
            return self.download(p, len(t), np.uint64)
        finally:
            self.free(p)

    def set_difference(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        pa, pb = self.upload(np.asarray(a, np.uint64)), self.upload(np.asarray(b, np.uint64))
        po = self.alloc(8 * max(len(a), 1))
        k = c_int64()
        try:
            self.check(self.L.dlg_set_difference(self.h, c_void_p(pa), len(a), c_void_p(pb), len(b),
                                                 c_void_p(po), byref(k)), "set_difference")
            return self.download(po, k.value, np.uint64)
        finally:
            for p in (pa, pb, po):
                self.free(p)

    def merge(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        pa, pb = self.upload(np.asarray(a, np.uint64)), self.upload(np.asarray(b, np.uint64))
        po = self.alloc(8 * max(len(a) + len(b), 1))
        try:
            self.check(self.L.dlg_merge(self.h, c_void_p(pa), len(a), c_void_p(pb), len(b), c_void_p(po)), "merge")
            return self.download(po, len(a) + len(b), np.uint64)
        finally:
            for p in (pa, pb, po):
                self.free(p)

    def merge_unique(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        pa, pb = self.upload(np.asarray(a, np.uint64)), self.upload(np.asarray(b, np.uint64))
        po = self.alloc(8 * max(len(a) + len(b), 1))
        try:
            k = c_int64()
            self.check(self.L.dlg_merge_unique(self.h, c_void_p(pa), len(a), c_void_p(pb), len(b), c_void_p(po),
                                               byref(k)), "merge_unique")
            return self.download(po, k.value, np.uint64)
        finally:
            for p in (pa, pb, po):
                self.free(p)

    def exclusive_scan(self, counts: np.ndarray):
        counts = np.asarray(counts, np.uint32)
        pi = self.upload(counts)
        po = self.alloc(8 * max(len(counts), 1))
        tot = c_uint64()
        try:
            self.check(self.L.dlg_exclusive_scan(self.h, c_void_p(pi), len(counts), c_void_p(po), byref(tot)), "scan")
            return self.download(po, len(counts), np.uint64), tot.value
        finally:
            self.free(pi)
            self.free(po)

    def hash_join(self, rel: np.ndarray, probe: np.ndarray, proj: int = PROJ_HV_PV, load_factor: float = 0.0):
        """get_hash_table(rel) then get_join(probe): returns (join tuples, hash info)."""
        pr, pp = self.upload(np.asarray(rel, np.uint64)), self.upload(np.asarray(probe, np.uint64))
        hh = c_void_p()
        out, n = c_void_p(), c_int64()
        try:
            self.check(self.L.dlg_hash_build(self.h, c_void_p(pr), len(rel), load_factor, byref(hh)), "hash_build")
            slots, keys, rows = c_int64(), c_int64(), c_int64()
            self.L.dlg_hash_info(hh, byref(slots), byref(keys), byref(rows))
            self.check(self.L.dlg_join(self.h, hh, c_void_p(pp), len(probe), proj, byref(out), byref(n)), "join")
            res = self.download(out.value, n.value, np.uint64) if n.value else np.zeros(0, np.uint64)
            return res, {"slots": slots.value, "keys": keys.value, "rows": rows.value}
        finally:
            if out.value:
                self.free(out.value)
            if hh.value:
                self.L.dlg_hash_free(self.h, hh)
            self.free(pr)
            self.free(pp)

    def hash_index(self, rel: np.ndarray, load_factor: float = 0.0) -> "HashIndex":
        """get_hash_table(rel): an index kept on the device for several joins / WCC rounds."""
        return HashIndex(self, rel, load_factor)

    # ---- profiling
    def profiling(self, on: bool):
        self.check(self.L.dlg_profiling(self.h, 1 if on else 0), "profiling")

    def kernel_stats(self) -> dict:
        out = {}
        for k, name in enumerate(KERNELS):
            n, ms, by = c_int64(), c_double(), c_double()
            self.check(self.L.dlg_kernel_stats(self.h, k, byref(n), byref(ms), byref(by)), "kernel_stats")
            out[name] = {"launches": n.value, "ms": ms.value, "bytes": by.value}
        return out

    def kernel_stats_reset(self):
        self.check(self.L.dlg_kernel_stats_reset(self.h), "kernel_stats_reset")

    # ---- batched form
    def batch_run(self, prog: str, instances, keep_output: bool = True, **tune):
        """Run many independent instances of one program as one fixpoint; returns (results, stats)."""
        arrs = [np.ascontiguousarray(np.asarray(r, dtype=np.int32).reshape(-1, 2)) for r in instances]
        B = len(arrs)
        ptrs = (c_void_p * max(B, 1))(*[a.ctypes.data for a in arrs])
        ns = (c_int64 * max(B, 1))(*[a.shape[0] for a in arrs])
        o = Opts()
        o.keep_output = 1 if keep_output else 0
        for k, v in tune.items():
            setattr(o, k, v)
        res = (Result * max(B, 1))()
        st = BatchStats()
        self.check(self.L.dlg_batch_run(self.h, PROGS[prog], B, ptrs, ns, byref(o), res, byref(st)), "batch_run")
        out = []
        try:
            for i in range(B):
                r = res[i]
                out.append({"iterations": r.iterations, "output_size": r.output_size, "secondary": r.secondary,
                            "nodes": r.nodes, "rows": _adopt_rows(r),
                            "output_sum": r.output_sum, "output_order_breaks": r.output_order_breaks})
        finally:
            for i in range(B):
                self.L.dlg_result_free(byref(res[i]))
        return out, {"device_seconds": st.device_seconds, "rounds": st.rounds,
                     "total_output": st.total_output, "union_rows": st.union_rows}

    def autotune(self, prog: str, rows: np.ndarray, load_factors=None, radix_bits=None, block_sizes=None,
                 repeats: int = 1):
        rows = np.ascontiguousarray(rows, dtype=np.int32).reshape(-1, 2)

        def arr(vals, ctype):
            return None if vals is None else (ctype * (len(vals) + 1))(*vals, 0)

        best, sec = Opts(), c_double()
        self.check(self.L.dlg_autotune(self.h, PROGS[prog], rows.ctypes.data, rows.shape[0],
                                       arr(load_factors, c_double), arr(radix_bits, c_int32),
                                       arr(block_sizes, c_int32), repeats, byref(best), byref(sec)), "autotune")
        return {"load_factor": best.load_factor, "radix_bits": best.radix_bits,
                "block_size": best.block_size, "seconds": sec.value}

    # ---- whole programs
    def run(self, prog: str, rows: np.ndarray, keep_output: bool = True, device_rows: int | None = None,
            n: int | None = None, **tune) -> dict:
        """Run tc / sg / wcc / single_join on host rows (or on device rows when device_rows is given)."""
        o = Opts()
        o.keep_output = 1 if keep_output else 0
        for k, v in tune.items():
            setattr(o, k, v)
        r = Result()
        if device_rows is not None:
            fn = getattr(self.L, f"dlg_run_{prog}_device")
            s = fn(self.h, c_void_p(device_rows), n, byref(o), byref(r))
        else:
            rows = np.ascontiguousarray(rows, dtype=np.int32).reshape(-1, 2)
            fn = getattr(self.L, f"dlg_run_{prog}")
            s = fn(self.h, rows.ctypes.data, rows.shape[0], byref(o), byref(r))
        self.check(s, f"run_{prog}")
        try:
            out = {"iterations": r.iterations, "output_size": r.output_size, "secondary": r.secondary,
                   "nodes": r.nodes, "raw_join_rows": r.raw_join_rows, "device_seconds": r.device_seconds,
                   "phases": {PHASES[i]: r.phase_seconds[i] for i in range(10)},
                   "output_sum": r.output_sum, "output_order_breaks": r.output_order_breaks}
            out["rows"] = _adopt_rows(r)
            return out
        finally:
            self.L.dlg_result_free(byref(r))


class HashIndex:
    """A dlg_hash (dlg_hash_build over a copy of `rel`) bound to one context."""

    def __init__(self, ctx: Ctx, rel: np.ndarray, load_factor: float = 0.0):
        self.ctx = ctx
        self.h = c_void_p()
        pr = ctx.upload(np.asarray(rel, np.uint64))
        try:
            ctx.check(ctx.L.dlg_hash_build(ctx.h, c_void_p(pr), len(rel), load_factor, byref(self.h)), "hash_build")
        finally:
            ctx.free(pr)

    def info(self) -> dict:
        slots, keys, rows = c_int64(), c_int64(), c_int64()
        self.ctx.L.dlg_hash_info(self.h, byref(slots), byref(keys), byref(rows))
        return {"slots": slots.value, "keys": keys.value, "rows": rows.value}

    def join(self, probe: np.ndarray, proj: int = PROJ_HV_PV) -> np.ndarray:
        """get_join(ht, probe) with the given projection, rows in the library's output order."""
        c = self.ctx
        pp = c.upload(np.asarray(probe, np.uint64))
        out, n = c_void_p(), c_int64()
        try:
            c.check(c.L.dlg_join(c.h, self.h, c_void_p(pp), len(probe), proj, byref(out), byref(n)), "join")
            return c.download(out.value, n.value, np.uint64) if n.value else np.zeros(0, np.uint64)
        finally:
            if out.value:
                c.free(out.value)
            c.free(pp)

    def wcc_relax(self, delta: np.ndarray, cc: np.ndarray):
        """One reference WCC round (dlg_wcc_relax): returns (cc', delta')."""
        c = self.ctx
        cc = np.ascontiguousarray(cc, np.uint64)
        pd = c.upload(np.asarray(delta, np.uint64))
        pc = c.upload(cc)
        pn = c.alloc(8 * max(len(cc), 1))
        k = c_int64()
        try:
            c.check(c.L.dlg_wcc_relax(c.h, self.h, c_void_p(pd), len(delta), c_void_p(pc), len(cc), c_void_p(pn),
                                      byref(k)), "wcc_relax")
            return c.download(pc, len(cc), np.uint64), c.download(pn, k.value, np.uint64)
        finally:
            for p in (pd, pc, pn):
                c.free(p)

    def close(self):
        if self.h:
            self.ctx.L.dlg_hash_free(self.ctx.h, self.h)
            self.h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Served form: request loop with HBM-budget admission and continuous batching."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget: int = 0, max_batch: int = 64,
                 small_rows: int = 20000, retain_bytes: int = 0, max_request_bytes: int = 0, **tune):
        self.L = lib()
        o = ServerOpts()
        o.device, o.workers, o.hbm_budget, o.max_batch = device, workers, hbm_budget, max_batch
        o.small_rows, o.retain_bytes, o.max_request_bytes = small_rows, retain_bytes, max_request_bytes
        for k, v in tune.items():
            setattr(o.tune, k, v)
        self.h = c_void_p()
        s = self.L.dlg_server_create(byref(o), byref(self.h))
        if s != DLG_OK:
            raise DlgError(s, "dlg_server_create")

    def _err(self, s, what):
        if s != DLG_OK:
            raise DlgError(s, f"{what}: {self.L.dlg_server_error(self.h).decode(errors='replace')}")

    def add_graph(self, rows: np.ndarray) -> int:
        rows = np.ascontiguousarray(rows, dtype=np.int32).reshape(-1, 2)
        gid = c_int32()
        self._err(self.L.dlg_server_add_graph(self.h, rows.ctypes.data, rows.shape[0], byref(gid)), "add_graph")
        return gid.value

    def submit(self, prog: str, rows=None, graph: int = -1) -> int:
        """Queue a request; returns its ticket.  Raises DlgError(DLG_REFUSED) when it can never fit."""
        t = c_uint64()
        if rows is None:
            s = self.L.dlg_server_submit(self.h, PROGS[prog], None, 0, graph, byref(t))
        else:
            rows = np.ascontiguousarray(rows, dtype=np.int32).reshape(-1, 2)
            s = self.L.dlg_server_submit(self.h, PROGS[prog], rows.ctypes.data, rows.shape[0], graph, byref(t))
        self._err(s, "submit")
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0):
        r, info = Result(), RequestInfo()
        s = self.L.dlg_server_wait(self.h, ticket, timeout, byref(r), byref(info))
        if s == DLG_PENDING:
            return None
        self._err(s, "wait")
        try:
            out = {"status": info.status, "iterations": r.iterations, "output_size": r.output_size,
                   "secondary": r.secondary, "nodes": r.nodes,
                   "rows": _adopt_rows(r),
                   "submit_s": info.submit_s, "start_s": info.start_s, "end_s": info.end_s,
                   "latency_s": info.end_s - info.submit_s, "batch_size": info.batch_size,
                   "attempts": info.attempts, "slot": info.slot, "reserved_bytes": info.reserved_bytes}
        finally:
            self.L.dlg_result_free(byref(r))
        return out

    def close(self):
        if self.h:
            self.L.dlg_server_destroy(self.h)
            self.h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
