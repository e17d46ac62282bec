"""Host side of the drop-in (CPU only): the library's reader, preprocessing and
generators against the oracle; C-ABI exports; the tc CLI's error paths."""
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import wtc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wedge-parallel-triangle-counting_amd")

from test_oracle import READER_CASES, _write  # noqa: E402


def same_csr(a, b):
    return a.n == b.n and a.m == b.m and np.array_equal(np.asarray(a.row_ptr, np.uint64),
                                                       np.asarray(b.row_ptr, np.uint64)) \
        and np.array_equal(np.asarray(a.col_ind, np.uint32), np.asarray(b.col_ind, np.uint32))


def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "wtc.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(wtc_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 30
    import ctypes
    lib = ctypes.CDLL(wtc.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(wtc.EXPORTED) <= names


def test_golden_reader_and_preprocess_match_oracle(golden):
    for fx in golden:
        g = wtc.read_graph(fx["path"], fx["format"] == "mm", fx["zero_indexed"])
        og = O.read_graph(fx["path"], fx["format"] == "mm", fx["zero_indexed"])
        assert same_csr(g, og), fx["file"]
        p = wtc.preprocess(g)
        h = hashlib.sha256(np.asarray(p.row_ptr).astype("<u8").tobytes() +
                           np.asarray(p.col_ind).astype("<u4").tobytes()).hexdigest()
        assert h == fx["csr_sha256"], fx["file"]
        assert p.wedges() == fx["W"]
        assert p.validate_oriented()


@pytest.mark.parametrize("case", READER_CASES, ids=[c[0] for c in READER_CASES])
def test_reader_cases_match_oracle(tmp_path, case):
    name, text, mm, zero, n, expect = case
    path = _write(tmp_path, name, text)
    og = O.read_graph(path, mm, zero)
    if isinstance(og, int):
        with pytest.raises(wtc.WtcError) as e:
            wtc.read_graph(path, mm, zero)
        assert e.value.code == og
        return
    assert same_csr(wtc.read_graph(path, mm, zero), og)


def _junk_text(rng):
    """Edge-list-like text with stray characters, signs, blank and comment
    lines, and header lines longer than read_graph's 256-byte line buffer."""
    parts = []
    if rng.random() < 0.3:
        parts.append("%%MatrixMarket matrix coordinate pattern\n")
    if rng.random() < 0.3:
        parts.append("#" + "x" * int(rng.integers(200, 300)) + str(int(rng.integers(0, 9))) + " 7\n")
    for _ in range(int(rng.integers(0, 120))):
        r = rng.random()
        if r < 0.08:
            parts.append(str(rng.choice(list("%#x.e\t\r+\n "))))
        else:
            parts.append(("+" if rng.random() < 0.05 else "") + str(int(rng.integers(0, 400))))
            parts.append(" " if rng.random() < 0.6 else "\n")
    return "".join(parts)


def test_reader_junk_text_matches_oracle(tmp_path):
    """Seeded malformed inputs: the parallel reader must stop, fail or succeed
    exactly where the sequential restatement of tc.cu:561-685 does."""
    rng = np.random.default_rng(11)
    for i in range(150):
        path = _write(tmp_path, f"junk{i}.txt", _junk_text(rng))
        for mm in (False, True):
            for zero in (False, True):
                og = O.read_graph(path, mm, zero)
                if isinstance(og, int):
                    with pytest.raises(wtc.WtcError) as e:
                        wtc.read_graph(path, mm, zero)
                    assert e.value.code == og, (i, mm, zero)
                else:
                    assert same_csr(wtc.read_graph(path, mm, zero), og), (i, mm, zero)


def test_preprocess_style_checks():
    g = wtc.gen_rmat(11, 8, seed=3)
    ref = wtc.preprocess(g, 0)
    assert same_csr(wtc.preprocess(g), ref)  # the binding's default is the host builder
    with pytest.raises(wtc.WtcError):
        wtc.preprocess(g, 3)
    if wtc.device_count() == 0:  # -p 1/2 need the GPU: no silent host fallback
        with pytest.raises(wtc.WtcError) as e:
            wtc.preprocess(g, 2)
        assert e.value.code == -8


def test_random_graphs_match_oracle():
    for seed in range(4):
        pairs = wtc.gen_rmat_pairs(11, 8, seed=seed)
        g, og = wtc.graph_from_pairs(pairs), O.graph_from_pairs(pairs)
        assert same_csr(g, og)
        p, op = wtc.preprocess(g), O.preprocess(og)
        assert same_csr(p, op)
        assert p.wedges() == int(O.wedge_prefix(op)[-1])
        assert p.alg_bytes(4) == O.alg_bytes(op, 4)
        assert p.alg_bytes(8) == O.alg_bytes(op, 8)


def test_rmat_paths_agree(tmp_path):
    pairs = wtc.gen_rmat_pairs(10, 4, seed=5)
    path = tmp_path / "rmat.txt"
    np.savetxt(path, pairs, fmt="%d", header="rmat", comments="# ")
    g_file = wtc.read_graph(str(path), zero_indexed=True)
    assert same_csr(g_file, wtc.gen_rmat(10, 4, seed=5))
    assert same_csr(g_file, wtc.graph_from_pairs(pairs))


def test_generators_deterministic_and_sized():
    a = wtc.gen_rmat(9, 8, seed=1)
    assert same_csr(a, wtc.gen_rmat(9, 8, seed=1))
    assert not same_csr(a, wtc.gen_rmat(9, 8, seed=2))
    cl = wtc.gen_chung_lu(5000, 2.7, 1, 60, 8000, seed=4)
    assert cl.m == 2 * 8000
    assert same_csr(cl, wtc.gen_chung_lu(5000, 2.7, 1, 60, 8000, seed=4))
    ba = wtc.gen_barabasi_albert(1000, 5, seed=2)
    assert ba.m == 2 * (1000 - 5) * 5
    er = wtc.gen_erdos_renyi(2000, 0.01, seed=3)
    mean = er.m / 2 / (2000 * 1999 / 2)
    assert 0.008 < mean < 0.012


def test_generators_thread_count_independent():
    code = ("import sys; sys.path.insert(0, %r); import wtc, hashlib, numpy as np; "
            "g = wtc.preprocess(wtc.gen_rmat(12, 8, seed=9)); "
            "print(hashlib.sha256(np.asarray(g.row_ptr).tobytes() + np.asarray(g.col_ind).tobytes()).hexdigest())"
            % PKG)
    outs = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, WTC_HOST_THREADS=t)
        outs.add(subprocess.run(["python3", "-c", code], env=env, capture_output=True, text=True,
                                check=True).stdout.strip())
    assert len(outs) == 1


def test_validate_rejects_bad_csr():
    good = wtc.preprocess(wtc.gen_rmat(8, 8, seed=1))
    assert good.validate_oriented()
    bad_cases = []
    ci = np.array(good.col_ind).copy()
    r = int(np.argmax(np.diff(good.row_ptr.astype(np.int64)) >= 2))
    s = int(good.row_ptr[r])
    ci[s], ci[s + 1] = ci[s + 1], ci[s]  # unsorted row
    bad_cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci))
    ci2 = np.array(good.col_ind).copy()
    ci2[-1] = good.n  # id out of range
    bad_cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci2))
    rp = np.array(good.row_ptr).copy()
    rp[-1] -= 1  # rowPtr[n] != m
    bad_cases.append(wtc.CSR(good.n, good.m, rp, np.array(good.col_ind).copy()))
    sym = wtc.gen_rmat(8, 8, seed=1)  # symmetric: has ids below their row
    bad_cases.append(sym)
    for b in bad_cases:
        assert not b.validate_oriented()


def test_tc_gpu_drop_ins_reject_bad_csr_on_host():
    """Both tc_GPU drop-ins check the whole CSR on the host before any device
    call: a row out of order, an id not above its row, an id out of range and
    a decreasing row offset all return WTC_ERR_CSR (-10), with or without a
    GPU, and nothing reaches a kernel."""
    good = wtc.preprocess(wtc.gen_barabasi_albert(500, 3, seed=903))
    assert good.validate_oriented()
    d = np.diff(good.row_ptr.astype(np.int64))
    r = int(np.argmax(d >= 2))
    r0 = int(good.row_ptr[r])
    cases = []
    ci = good.col_ind.copy()
    ci[r0], ci[r0 + 1] = ci[r0 + 1], ci[r0]  # a row out of order
    cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci))
    ci = good.col_ind.copy()
    ci[r0] = r  # an id not above its row
    cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci))
    ci = good.col_ind.copy()
    ci[-1] = good.n  # an id out of range
    cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci))
    rp = good.row_ptr.copy()
    rp[r] = rp[r + 1] + 1  # a decreasing offset: rowPtr[r] > rowPtr[r+1]
    cases.append(wtc.CSR(good.n, good.m, rp, good.col_ind.copy()))
    for bad in cases:
        assert not bad.validate_oriented()
        for fn in (wtc.tc_gpu, wtc.tc_gpu_big):
            with pytest.raises(wtc.WtcError) as ei:
                fn(bad, 5, 64)
            assert ei.value.code == -10, fn.__name__


def test_autotune_table():
    # whatever table is in force (the committed sweep's or the built-in
    # rows): no bitmap for a graph without wedges, a bitmap length that is a
    # multiple of 32 and no larger than n rounded up, and a spread of at
    # least 1
    assert wtc.autotune(10, 0, 0) == (1, 0)
    for n, m, W in ((4083, 48254, 780919), (1048465, 15701962, 1226291937),
                    (16775427, 260385033, 41737375097), (268428511, 4236159419, 1357709474934)):
        s, a = wtc.autotune(n, m, W)
        assert s >= 1 and a % 32 == 0 and a <= ((n + 31) // 32) * 32
    for n in (64, 1000, 5000, 65536, 10 ** 6):
        for batch in (False, True):
            s, a = wtc.autotune(n, 20 * n, 600 * n, batch=batch)
            assert s >= 1 and a % 32 == 0


# ------------------------------------------------------------------- CLI

def _tc(*args, binary="tc"):
    return subprocess.run([os.path.join(PKG, binary), *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("binary", ["tc", "tc_big_graphs"])
def test_cli_usage_errors(tmp_path, binary):
    good = _write(tmp_path, "g.txt", "1 2\n2 3\n3 1\n")
    zero = _write(tmp_path, "z.txt", "0 1\n1 2\n")
    for args in ([], ["-e", good], ["-e", good, "-s", "0"], ["-e", good, "-s", "5", "-a", "33"],
                 ["-e", good, "-s", "5", "-p", "3"], ["-q"], ["-e"]):
        r = _tc(*args, binary=binary)
        assert r.returncode == 1, args
        assert "Usage:" in r.stdout and "-s <num>" in r.stdout
    r = _tc("-e", "/nonexistent.txt", "-s", "5", "-a", "0", binary=binary)
    assert r.returncode == 1 and "unable to open graph file" in r.stderr
    r = _tc("-e", zero, "-s", "5", "-a", "0", binary=binary)
    assert r.returncode == 1 and "zero vertex id detected but -z was not set" in r.stderr


@pytest.mark.parametrize("binary", ["tc_vertex", "tc_edge"])
def test_experiment_cli_usage_errors(tmp_path, binary):
    """tc_vertex / tc_edge: -w <0..2> instead of -s (experiments/tc_vertex.cu:383-400)."""
    good = _write(tmp_path, "g.txt", "1 2\n2 3\n3 1\n")
    for args in ([], ["-e", good, "-w", "3"], ["-e", good, "-s", "5"], ["-e", good, "-a", "33"], ["-e"]):
        r = _tc(*args, binary=binary)
        assert r.returncode == 1, args
        assert "Usage:" in r.stdout and "-w <num>" in r.stdout and "-s <num>" not in r.stdout
    r = _tc("-e", "/nonexistent.txt", "-a", "0", binary=binary)
    assert r.returncode == 1 and "unable to open graph file" in r.stderr


# ------------------------------------------------------------- ingestion

def _edge_text(rng, n_pairs, n_ids, one_indexed=True):
    """Edge list text with mixed separators (spaces, tabs, CRLF, several
    pairs per line, blank lines) so that the parallel reader's cuts land in
    every kind of whitespace run."""
    ids = rng.integers(1 if one_indexed else 0, n_ids, size=2 * n_pairs)
    seps = [" ", "\t", "  ", "\n", "\r\n", "\n\n", " \t "]
    pick = rng.integers(0, len(seps), size=2 * n_pairs)
    out = ["% header line\n# another\n"]
    out.extend(f"{x}{seps[k]}" for x, k in zip(ids.tolist(), pick.tolist()))
    return "".join(out)


def test_parallel_reader_matches_oracle_on_large_files(tmp_path):
    rng = np.random.default_rng(12)
    text = _edge_text(rng, 600_000, 200_000)  # several MiB: one piece per thread
    assert len(text) > 4 << 20
    path = _write(tmp_path, "big.txt", text)
    assert same_csr(wtc.read_graph(path), O.read_graph(path, False, False))
    # garbage at several depths, including right after a cut-sized prefix
    for frac in (0.0001, 0.3, 0.5, 0.77, 0.999):
        k = int(len(text) * frac)
        k = text.index(" ", k) + 1  # start of a token
        for junk in ("x", "1.5", "+ 3", "-"):
            bad = text[:k] + junk + " " + text[k:]
            p2 = _write(tmp_path, "junk.txt", bad)
            og = O.read_graph(p2, False, False)
            assert same_csr(wtc.read_graph(p2), og), (frac, junk)


def test_parallel_reader_zero_ids_and_odd_tails(tmp_path):
    rng = np.random.default_rng(13)
    text = _edge_text(rng, 400_000, 50_000)
    k = text.index("\n", len(text) * 2 // 3) + 1
    with_zero = text[:k] + "0 5\n" + text[k:]
    p = _write(tmp_path, "z.txt", with_zero)
    with pytest.raises(wtc.WtcError) as e:
        wtc.read_graph(p)
    assert e.value.code == O.read_graph(p, False, False) == -4
    # the zero id is never read when garbage comes first
    p = _write(tmp_path, "z2.txt", text[:k] + "stop\n0 5\n" + text[k:])
    assert same_csr(wtc.read_graph(p), O.read_graph(p, False, False))
    # zero-indexed file, odd number of tokens: the last one is dropped
    p = _write(tmp_path, "odd.txt", _edge_text(rng, 300_000, 90_000, one_indexed=False) + "17")
    assert same_csr(wtc.read_graph(p, zero_indexed=True), O.read_graph(p, False, True))


def test_reader_thread_count_independent(tmp_path):
    rng = np.random.default_rng(14)
    path = _write(tmp_path, "t.txt", _edge_text(rng, 500_000, 100_000))
    code = ("import sys; sys.path.insert(0, %r); import wtc, hashlib, numpy as np; "
            "g = wtc.read_graph(%r); "
            "print(hashlib.sha256(np.asarray(g.row_ptr).tobytes() + np.asarray(g.col_ind).tobytes()).hexdigest())"
            % (PKG, path))
    outs = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, WTC_HOST_THREADS=t)
        outs.add(subprocess.run(["python3", "-c", code], env=env, capture_output=True, text=True,
                                check=True).stdout.strip())
    assert len(outs) == 1


def test_csr_file_round_trip_and_corruption(tmp_path):
    g = wtc.preprocess(wtc.gen_rmat(12, 8, seed=2))
    path = str(tmp_path / "g.csr")
    wtc.write_csr(path, g, tag=0xDEADBEEF12345678)
    h, tag = wtc.read_csr(path)
    assert same_csr(h, g) and tag == 0xDEADBEEF12345678
    assert not os.path.exists(path + ".tmp")
    raw = bytearray(open(path, "rb").read())

    def expect_format(data, name):
        p = str(tmp_path / name)
        open(p, "wb").write(bytes(data))
        with pytest.raises(wtc.WtcError) as e:
            wtc.read_csr(p)
        assert e.value.code == -3, name

    flip = bytearray(raw)
    flip[64 + 8 * (g.n + 1) + 4 * 100] ^= 1  # one bit of colInd
    expect_format(flip, "flip.csr")
    expect_format(raw[:-4], "short.csr")
    expect_format(raw + b"\0\0\0\0", "long.csr")
    magic = bytearray(raw)
    magic[0] ^= 0xFF
    expect_format(magic, "magic.csr")
    expect_format(raw[:40], "tiny.csr")
    reserved = bytearray(raw)
    reserved[48] ^= 0x10  # header bytes 40..63 are reserved, written as zero
    expect_format(reserved, "reserved.csr")
    # a checksum-valid file whose ids run past n is still refused
    bad = wtc.CSR(g.n, g.m, g.row_ptr.copy(), np.array(g.col_ind).copy())
    bad.col_ind[-1] = g.n + 5
    wtc.write_csr(str(tmp_path / "ids.csr"), bad)
    with pytest.raises(wtc.WtcError) as e:
        wtc.read_csr(str(tmp_path / "ids.csr"))
    assert e.value.code == -3
    with pytest.raises(wtc.WtcError) as e:
        wtc.read_csr(str(tmp_path / "missing.csr"))
    assert e.value.code == -2
