"""GPU preprocessing (-p 1/2, device builders) against the host builders and
the golden fixtures: identical CSRs, identical counts."""
import hashlib

import numpy as np
import pytest

import oracle as O
import wtc

pytestmark = pytest.mark.gpu


def same_csr(a, b):
    return a.n == b.n and a.m == b.m and np.array_equal(np.asarray(a.row_ptr, np.uint64),
                                                       np.asarray(b.row_ptr, np.uint64)) \
        and np.array_equal(np.asarray(a.col_ind, np.uint32), np.asarray(b.col_ind, np.uint32))


def test_styles_identical_on_golden(golden):
    for fx in golden:
        mm, z = fx["format"] == "mm", fx["zero_indexed"]
        host = wtc.read_graph(fx["path"], mm, z, 0)
        for style in (1, 2):
            dev = wtc.read_graph(fx["path"], mm, z, style)
            assert same_csr(dev, host), fx["file"]
            p = wtc.preprocess(dev, style)
            h = hashlib.sha256(np.asarray(p.row_ptr).astype("<u8").tobytes() +
                               np.asarray(p.col_ind).astype("<u4").tobytes()).hexdigest()
            assert h == fx["csr_sha256"], fx["file"]


def test_preprocess_gpu_random():
    for seed in range(3):
        g = wtc.gen_rmat(13, 8, seed=seed)
        assert same_csr(wtc.preprocess(g, 2), wtc.preprocess(g, 0))
    empty = wtc.graph_from_pairs(np.zeros((0, 2), np.uint32))
    assert same_csr(wtc.preprocess(empty, 2), wtc.preprocess(empty, 0))


def test_device_builders(gpu_ctx):
    for scale, seed in ((6, 1), (10, 2), (14, 3), (16, 4)):
        host = wtc.preprocess(wtc.gen_rmat(scale, 16, seed=seed))
        dg = wtc.DeviceGraph.gen_rmat(gpu_ctx, scale, 16, seed=seed)
        assert same_csr(dg.download(), host), scale
        assert dg.count(5, 64) == O.tc_merge(host)
        dg.free()
        pairs = wtc.gen_rmat_pairs(scale, 8, seed=seed)
        for ob in (4, 8):
            dp = wtc.DeviceGraph.from_pairs(gpu_ctx, pairs, offset_bytes=ob)
            assert same_csr(dp.download(), wtc.preprocess(wtc.graph_from_pairs(pairs))), (scale, ob)
            dp.free()


def test_long_runs_of_empty_rows(gpu_ctx, tmp_path):
    """Runs of hundreds of isolated ids (empty rows) before, between and after
    the used ids, in the symmetric CSR and, ranked first, in the oriented one:
    the row-offset builder writes short runs from the entry after them and
    searches the rest."""
    k5 = [(i, j) for i in range(5) for j in range(i + 1, 5)]
    k4 = [(2000 + i, 2000 + j) for i in range(4) for j in range(i + 1, 4)]
    for extra in ([(1000, 1001)], [(1000, 1001), (5000, 5001)], []):
        pairs = np.array(k5 + k4 + extra, np.uint32)
        sym_host = wtc.graph_from_pairs(pairs)
        path = tmp_path / "runs.txt"
        np.savetxt(path, pairs, fmt="%d")
        for style in (1, 2):  # the symmetric CSR built on the GPU
            assert same_csr(wtc.read_graph(str(path), zero_indexed=True, preprocess_style=style), sym_host)
        host = wtc.preprocess(sym_host, 0)
        assert same_csr(wtc.preprocess(sym_host, 2), host)
        for ob in (4, 8):
            dp = wtc.DeviceGraph.from_pairs(gpu_ctx, pairs, offset_bytes=ob)
            assert same_csr(dp.download(), host), (extra, ob)
            assert dp.count(3, 0) == 10 + 4
            dp.free()
