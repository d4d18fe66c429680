"""C ABI checks that need no GPU: every declared symbol is exported, and the host I/O
entry points (bl_mtx_read, bl_initf_read, bl_output_write) reproduce the reference's
file handling (IO.h:60-155, algorithms.cpp:98-119, :1244-1256)."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden, golden_files, golden_input


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "batchlayout.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(bl_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported(bl):
    syms = declared_symbols()
    assert len(syms) >= 14
    L = bl.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


@pytest.mark.parametrize("fx", golden_files("csr_*.npz"))
def test_mtx_read_matches_reference_csr(bl, fx):
    d = golden(fx)
    rp, ci = bl.mtx_read(golden_input(fx[4:-4] + ".mtx"))
    assert np.array_equal(rp, d["rowptr"]) and np.array_equal(ci, d["colids"])


def test_mtx_read_matches_oracle_on_random_files(bl, oracle, tmp_path):
    rng = np.random.default_rng(3)
    for t in range(20):
        n = int(rng.integers(2, 60))
        m = int(rng.integers(0, 3 * n))
        sym = bool(rng.integers(0, 2))
        e = rng.integers(1, n + 1, size=(m, 2))
        head = "%%MatrixMarket matrix coordinate pattern " + ("symmetric" if sym else "general")
        p = tmp_path / f"r{t}.mtx"
        with open(p, "w") as f:
            f.write(head + "\n% comment\n" + f"{n} {n} {m}\n")
            for a, b in e:
                f.write(f"{a} {b}\n")
        r1 = bl.mtx_read(str(p))
        r2 = oracle.mtx_read(str(p))
        assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])


@pytest.mark.parametrize("threads", ["1", "3", "16"])
def test_mtx_read_threaded_chunks(bl, oracle, tmp_path, monkeypatch, threads):
    # a body large enough to split into chunks, with CRLF ends, tabs and values; lines after
    # the declared count (including a malformed one) are never read
    rng = np.random.default_rng(11)
    n, m = 3000, 60000
    e = rng.integers(1, n + 1, size=(m, 2))

    def write(path, blank, declared):
        with open(path, "w", newline="") as f:
            f.write("%%MatrixMarket matrix coordinate real symmetric\r\n% c\r\n" + f"{n} {n} {declared}\r\n")
            for k, (a, b) in enumerate(e):
                if blank and k % 997 == 0:
                    f.write("\r\n   \n")
                f.write(f"{a}\t {b} {k % 7}.5\r\n" if k % 5 == 0 else f"{a} {b}\n")
            f.write("1 x\nnot an entry\n")

    p, pb, q = tmp_path / "big.mtx", tmp_path / "blank.mtx", tmp_path / "bad.mtx"
    write(p, False, m)
    write(pb, True, m)
    # declaring one entry more than the file holds reaches the malformed line
    write(q, True, m + 1)
    monkeypatch.setenv("BL_IO_THREADS", threads)
    r1 = bl.mtx_read(str(p))
    r2 = oracle.mtx_read(str(p))
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
    # blank lines (where the reference would fail) are skipped
    r3 = bl.mtx_read(str(pb))
    assert np.array_equal(r1[0], r3[0]) and np.array_equal(r1[1], r3[1])
    with pytest.raises(bl.BLError) as ei:
        bl.mtx_read(str(q))
    assert ei.value.code == -3


def test_mtx_read_errors(bl, tmp_path):
    cases = {
        "short.mtx": ("%%MatrixMarket matrix coordinate pattern symmetric\n4 4 3\n2 1\n3 2\n", -3),
        "range.mtx": ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n4 1\n", -3),
        "rect.mtx": ("%%MatrixMarket matrix coordinate pattern general\n3 4 1\n1 2\n", -3),
    }
    for name, (text, code) in cases.items():
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(bl.BLError) as ei:
            bl.mtx_read(str(p))
        assert ei.value.code == code
    with pytest.raises(bl.BLError) as ei:
        bl.mtx_read(str(tmp_path / "does_not_exist.mtx"))
    assert ei.value.code == -2


def test_initf_read(bl, tmp_path):
    p = tmp_path / "init.txt"
    xy = np.random.default_rng(1).random((5, 2))
    with open(p, "w") as f:
        for i, (x, y) in enumerate(xy):
            f.write("%.17g %.17g %d\n" % (x, y, i + 1))
            if i == 1:
                f.write("\n")  # a line that does not parse repeats the previous coordinate
    got, k = bl.initf_read(str(p), 7)
    assert k == 6
    assert np.array_equal(got[:2], xy[:2]) and np.array_equal(got[2], xy[1]) and np.array_equal(got[3:6], xy[2:])
    assert np.array_equal(got[6], [0.0, 0.0])
    got, k = bl.initf_read(str(p), 3)  # at most n lines are read
    assert k == 3


@pytest.mark.parametrize("fx", golden_files("out_*.txt"))
def test_output_write_is_byte_identical_to_reference(bl, fx, tmp_path):
    case, k = fx[4:-4].rsplit("_k", 1)
    d = golden(f"traj_{case}.npz")
    kk = list(d["ks"]).index(int(k))
    out = tmp_path / "o.txt"
    bl.output_write(str(out), d["xy"][kk])
    assert out.read_bytes() == open(os.path.join(ROOT, "tests", "golden", fx), "rb").read()


def test_params_default_is_reference_cli_default(bl):
    p = bl.params()
    assert (p.algo, p.iters, p.batch, p.bht, p.engt, p.weight, p.step0) == (2, 600, 256, 1.2, 0.01, 1.0, 1.0)


E_ARG, E_GRAPH = -1, -7


@pytest.mark.parametrize("rp,ci,code", [
    ([1, 1, 2], [1, 0], E_GRAPH),          # rowptr[0] != 0
    ([0, 2, 1, 2], [1, 0], E_GRAPH),       # rowptr not monotone
    ([0, 1, 2], [1, 2], E_GRAPH),          # colid == n
    ([0, 1, 2], [-1, 0], E_GRAPH),         # negative colid
    ([0, 1, 3], [1, 0], E_GRAPH),          # rowptr[n] != nnz
])
def test_graph_create_rejects_bad_csr_before_device_work(bl, rp, ci, code):
    """bl_graph_create validates the CSR on the host before it touches a device, so each of
    these returns its error with no GPU present."""
    import ctypes as C

    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    h = C.c_void_p()
    assert bl.lib().bl_graph_create(len(rp) - 1, len(ci), rp, ci.ctypes.data, None, C.byref(h)) == code
    assert not h.value


@pytest.mark.parametrize("case", ["eoff0", "voff_empty_graph", "eoff_decreasing", "rowptr_offset",
                                  "colid_global", "rowptr_not_monotone", "algo4"])
def test_run_batched_rejects_bad_input_before_device_work(bl, case):
    """bl_run_batched checks its offsets, then each graph's CSR, on the host before any device
    work: two triangles, each spoiled in one way, return BL_E_ARG or BL_E_GRAPH with no GPU."""
    import ctypes as C

    voff = np.array([0, 3, 6], np.int32)
    eoff = np.array([0, 6, 12], np.int64)
    rowptr = np.array([0, 2, 4, 6, 8, 10, 12], np.int32)
    colids = np.array([1, 2, 0, 2, 0, 1] * 2, np.int32)
    kw, code = {}, E_ARG
    if case == "eoff0":
        eoff[:] = [1, 7, 13]
        rowptr += 1
    elif case == "voff_empty_graph":
        voff[1] = 0
    elif case == "eoff_decreasing":
        eoff[1] = 13
    elif case == "rowptr_offset":
        rowptr[3] = 5
        code = E_GRAPH
    elif case == "colid_global":
        colids[7] = 4  # graph 1 has 3 vertices; colids are local to the graph
        code = E_GRAPH
    elif case == "rowptr_not_monotone":
        rowptr[1] = 5
        code = E_GRAPH
    elif case == "algo4":
        kw = dict(algo=4)
    xy = np.zeros((6, 2))
    p = bl.params(iters=1, batch=2, **kw)
    st = bl.bl_stats()
    rc = bl.lib().bl_run_batched(2, voff, eoff, rowptr, colids, xy, C.byref(p), np.zeros(2), np.zeros(2),
                                 C.byref(st))
    assert rc == code
