"""CPU-side checks of libdlg.so: it loads, exports every entry point that
include/dlg.h declares, and its host code (file format, glibc generator,
synthetic graph generators, argument checks) behaves.  No GPU needed."""
import os
import re
import subprocess

import numpy as np
import pytest

import mnmgdatalog_amd as dlg
from conftest import DATA, ROOT, read_bin

HEADER = os.path.join(ROOT, "include", "dlg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dlg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", dlg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (dlg_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = dlg.lib()
    for n in names:
        assert getattr(L, n) is not None


def test_library_carries_gfx950_code_objects():
    blob = open(dlg.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_kernel_names():
    L = dlg.lib()
    assert [L.dlg_kernel_name(i).decode() for i in range(len(dlg.KERNELS))] == dlg.KERNELS


def test_read_rows_matches_numpy(tmp_path):
    r = dlg.read_rows(os.path.join(DATA, "data_7035.bin"))
    assert np.array_equal(r, read_bin("data_7035.bin"))
    # trailing bytes are ignored: rows = filesize / 8
    p = tmp_path / "odd.bin"
    p.write_bytes(np.arange(5, dtype=np.int32).tobytes())
    assert dlg.read_rows(str(p)).tolist() == [[0, 1], [2, 3]]


def test_write_rows_truncates(tmp_path):
    p = str(tmp_path / "out.bin")
    dlg.write_rows(p, np.arange(20, dtype=np.int32).reshape(10, 2))
    dlg.write_rows(p, np.array([[7, 8]], np.int32))
    assert dlg.read_rows(p).tolist() == [[7, 8]]


def test_read_missing_file():
    with pytest.raises(dlg.DlgError) as e:
        dlg.read_rows("/nonexistent/file.bin")
    assert e.value.status == dlg.DLG_IO_ERROR


def test_rand_rows_same_as_oracle(orc):
    a = dlg.generate_rand_rows(100_000, 1_000_000)
    b = orc.generate_rand_rows(100_000, 1_000_000)
    assert np.array_equal(a, b)


def test_powerlaw_config_graph():
    r = dlg.gen_powerlaw(3000, 7000, 2.1, 0x4D4E4D47)
    assert r.shape == (7000, 2)
    assert len(set(map(tuple, r.tolist()))) == 7000
    assert (r[:, 0] != r[:, 1]).all() and r.min() >= 0 and r.max() < 3000
    assert np.array_equal(r, dlg.gen_powerlaw(3000, 7000, 2.1, 0x4D4E4D47))
    assert not np.array_equal(r, dlg.gen_powerlaw(3000, 7000, 2.1, 0x4D4E4D48))
    deg = np.bincount(r[:, 0], minlength=3000)
    assert deg.min() >= 1 and deg.max() > 20    # every vertex has out-edges; heavy tail
    assert (deg == 1).sum() > 1500              # most vertices have out-degree 1


def test_grid_config_graph():
    r = dlg.gen_grid(400, 0.55, 0.01, 20, 0x4D4E4D47)
    s = set(map(tuple, r.tolist()))
    assert all((b, a) in s for a, b in s)          # stored in both directions
    assert 330_000 < len(r) < 370_000
    d = np.abs(r[:, 0] // 400 - r[:, 1] // 400) + np.abs(r[:, 0] % 400 - r[:, 1] % 400)
    assert d.min() >= 1 and d.max() <= 20
    assert (d == 1).sum() > 0.97 * len(r)


def test_tetmesh_config_graph():
    r = dlg.gen_tetmesh(35)
    # 3*35*36*36 lattice edges + one diagonal per face (3*35*35*36)
    assert len(r) == 3 * 35 * 36 * 36 + 3 * 35 * 35 * 36
    assert (r[:, 0] < r[:, 1]).all()
    assert len(set(map(tuple, r.tolist()))) == len(r)
    small = dlg.gen_tetmesh(1)
    assert len(small) == 12 + 6   # cube edges + 6 face diagonals of the central tetrahedron


def test_erdos_renyi_and_rmat():
    e = dlg.gen_erdos_renyi(1000, 3000, 7)
    assert e.shape == (3000, 2) and (e[:, 0] != e[:, 1]).all()
    assert len(set(map(tuple, e.tolist()))) == 3000
    m = dlg.gen_rmat(12, 20000, 0.57, 0.19, 0.19, 9)
    assert m.shape == (20000, 2) and (m[:, 0] != m[:, 1]).all() and m.max() < 4096
    assert len(set(map(tuple, m.tolist()))) == 20000
    u = dlg.gen_rmat(10, 3000, 0.57, 0.19, 0.19, 9, undirected=True)
    s = set(map(tuple, u.tolist()))
    assert len(u) == 6000 and all((b, a) in s for a, b in s)


def test_generator_argument_checks():
    with pytest.raises(dlg.DlgError):
        dlg.gen_powerlaw(1, 1, 2.1, 0)
    with pytest.raises(dlg.DlgError):
        dlg.gen_tetmesh(0)


def test_pack_order_matches_signed_lexicographic():
    rng = np.random.default_rng(1)
    rows = rng.integers(-2**31, 2**31, size=(2000, 2), dtype=np.int64).astype(np.int32)
    t = dlg.pack(rows)
    order = np.lexsort((rows[:, 1], rows[:, 0]))
    assert np.array_equal(np.argsort(t, kind="stable"), order) or np.array_equal(t[order], np.sort(t))
    assert np.array_equal(dlg.unpack(t), rows)


def test_no_gpu_context_fails_cleanly():
    import ctypes
    L = dlg.lib()
    if L.dlg_ctx_create(0, ctypes.byref(ctypes.c_void_p())) == dlg.DLG_OK:
        pytest.skip("a GPU is visible")
    with pytest.raises(dlg.DlgError) as e:
        dlg.Ctx()
    assert e.value.status in (dlg.DLG_DEVICE_ERROR, dlg.DLG_INVALID_INPUT)


def test_cli_reports_error_without_gpu(tmp_path):
    import ctypes
    L = dlg.lib()
    h = ctypes.c_void_p()
    if L.dlg_ctx_create(0, ctypes.byref(h)) == dlg.DLG_OK:
        L.dlg_ctx_destroy(h)
        pytest.skip("a GPU is visible")
    p = subprocess.run([os.path.join(dlg.BIN_DIR, "tc"), os.path.join(DATA, "hipc_2019.bin")],
                       capture_output=True, text=True)
    assert p.returncode != 0 and "failed" in p.stderr


def test_ctypes_structs_match_the_c_header(tmp_path):
    """Size and field offsets of every struct the Python binding mirrors, as
    gcc lays out include/dlg.h."""
    import ctypes
    mirrors = {"dlg_opts": dlg.Opts, "dlg_result": dlg.Result, "dlg_batch_stats": dlg.BatchStats,
               "dlg_server_opts": dlg.ServerOpts, "dlg_request_info": dlg.RequestInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dlg.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


# ------------------------------------------------------------------ format tools
TWINS = ["data_5", "data_10", "paper", "quiz3", "tc_paper", "data_22", "data_7035"]


def tool(*args):
    return subprocess.run([os.path.join(dlg.BIN_DIR, args[0]), *map(str, args[1:])], capture_output=True,
                          text=True, timeout=60)


@pytest.mark.parametrize("name", TWINS)
def test_txt_to_bin_matches_reference_twins(tmp_path, name):
    """The reference ships each small graph as .txt and .bin; converting the
    text must give the binary byte for byte (binary_file_utils.py:5-18)."""
    out = tmp_path / f"{name}.bin"
    p = tool("binary_file_utils", "txt_to_bin", os.path.join(DATA, f"{name}.txt"), out)
    assert p.returncode == 0 and "Converted" in p.stdout
    assert out.read_bytes() == open(os.path.join(DATA, f"{name}.bin"), "rb").read()


@pytest.mark.parametrize("name", TWINS + ["hipc_2019", "dummy"])
def test_bin_to_txt_round_trip(tmp_path, name):
    src = os.path.join(DATA, f"{name}.bin")
    txt, back = tmp_path / "x.txt", tmp_path / "x.bin"
    assert tool("binary_file_utils", "bin_to_txt", src, txt).returncode == 0
    rows = read_bin(f"{name}.bin")
    assert txt.read_text() == "".join(f"{a}\t{b}\n" for a, b in rows)   # binary_file_utils.py:21-33
    assert tool("binary_file_utils", "txt_to_bin", txt, back).returncode == 0
    assert back.read_bytes() == open(src, "rb").read()


def test_txt_to_bin_field_rules(tmp_path):
    src, out = tmp_path / "in.txt", tmp_path / "out.bin"
    src.write_text("1 2 3\n-4\t+5\nx 7 y 8\n\n9\n1_0 2_000\nfoo\n")
    assert tool("binary_file_utils", "txt_to_bin", src, out).returncode == 0
    assert np.fromfile(out, np.int32).tolist() == [1, 2, -4, 5, 7, 8, 9, 10, 2000]


def test_format_tool_errors(tmp_path):
    src, out = tmp_path / "in.txt", tmp_path / "out.bin"
    src.write_text("1 2147483648\n")
    assert tool("binary_file_utils", "txt_to_bin", src, out).returncode != 0
    odd = tmp_path / "odd.bin"
    odd.write_bytes(b"\x01\x00\x00\x00\x02\x00\x00\x00\x03")
    assert tool("binary_file_utils", "bin_to_txt", odd, tmp_path / "o.txt").returncode != 0
    p = tool("binary_file_utils", "txt_to_bin", src)
    assert p.returncode == 0 and p.stdout.startswith("Usage:")
    p = tool("binary_file_utils", "nope", src, out)
    assert p.returncode == 0 and "Invalid conversion type" in p.stdout


def test_row_size():
    p = tool("row_size", os.path.join(DATA, "data_7035.bin"))
    assert p.returncode == 0 and p.stdout.strip() == f"{os.path.join(DATA, 'data_7035.bin')}: 7035"


@pytest.mark.parametrize("name,want", [("paper.bin", 18), ("dummy.bin", 13), ("hipc_2019.bin", 9),
                                       ("data_22.bin", 253), ("data_7035.bin", 146120)])
def test_predict_tc_size_exact_on_fixtures(name, want):
    """SCC-condensation count of |TC| equals the README / oracle counts (exact below 16,384
    components)."""
    from conftest import read_bin
    assert dlg.predict_tc_size(read_bin(name)) == (want, True)


def test_predict_tc_size_closed_forms_and_random(orc):
    n = 300
    chain = np.array([(i, i + 1) for i in range(n)], np.int32)
    assert dlg.predict_tc_size(chain) == (n * (n + 1) // 2, True)           # path: no self pairs
    cycle = np.concatenate([chain, [[n, 0]]]).astype(np.int32)
    assert dlg.predict_tc_size(cycle) == ((n + 1) ** 2, True)               # one SCC reaches itself
    loops = np.array([[5, 5], [5, 6], [7, 8]], np.int32)
    assert dlg.predict_tc_size(loops) == (3, True)                          # (5,5), (5,6), (7,8)
    assert dlg.predict_tc_size(np.zeros((0, 2), np.int32)) == (0, True)
    rng = np.random.default_rng(3)
    for seed in range(4):
        rows = rng.integers(-400, 400, size=(int(rng.integers(100, 1500)), 2)).astype(np.int32)
        got, exact = dlg.predict_tc_size(rows)
        assert exact and got == orc.tc(rows)["output_size"]
    for seed in range(2):
        rows = dlg.gen_erdos_renyi(3000, 7000, seed)
        assert dlg.predict_tc_size(rows) == (orc.tc(rows, threads=8)["output_size"], True)
        rows = dlg.gen_rmat(11, 6000, 0.57, 0.19, 0.19, seed)
        assert dlg.predict_tc_size(rows) == (orc.tc(rows, threads=8)["output_size"], True)


def test_predict_tc_size_estimate_on_a_wide_graph():
    """Above 16,384 components the count is estimated from sampled reach sets: SF.cedge's
    published closure is 80,498,014 tuples (logs/small-merged-job-output.txt:108)."""
    from conftest import read_bin
    got, exact = dlg.predict_tc_size(read_bin("data_223001.bin"))
    assert not exact
    assert abs(got - 80_498_014) < 0.05 * 80_498_014
