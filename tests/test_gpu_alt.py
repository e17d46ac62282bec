"""GPU parity of the reference's alternative counts through the C ABI: the
arrow-wedge count (experiments/tc_arrow.cu; triangles and the arrow-wedge
total) and the vertex-/edge-parallel baselines in their three wedge styles
(experiments/tc_vertex.cu, experiments/tc_edge.cu), against the oracle on the
golden fixtures, RMAT graphs with 32- and 64-bit offsets, edge cases and the
per-call drop-ins."""
import numpy as np
import pytest

import oracle as O
import wtc

pytestmark = pytest.mark.gpu


def _rmat(scale, ef=16, seed=1):
    return wtc.preprocess(wtc.graph_from_pairs(wtc.gen_rmat_pairs(scale, ef, seed=seed)))


def test_arrow_golden_fixtures(gpu_ctx, golden):
    for fx in golden:
        g = wtc.preprocess(wtc.read_graph(fx["path"], fx["format"] == "mm", fx["zero_indexed"]))
        _, arrow = O.tc_arrow(g, 0)
        dg = gpu_ctx.upload(g)
        for s, a in ((1, 0), (4, 64), (9, 8192)):
            cnt, aw, st = dg.count_arrow(s, a, stats=True)
            assert (cnt, aw) == (fx["triangles"], arrow), (fx["file"], s, a)
            assert st.wedges == arrow
        dg.free()


@pytest.mark.parametrize("ob", [4, 8])
def test_arrow_rmat_sweep(gpu_ctx, ob):
    g = _rmat(14, seed=9)
    expect = O.tc_merge(g)
    o_cnt, o_arrow = O.tc_arrow(g, 4096)
    assert o_cnt == expect
    dg = gpu_ctx.upload(g, offset_bytes=ob)
    for s in (1, 2, 5, 8, 17, 64):
        for a in (0, 32, 4096, 16384):
            assert dg.count_arrow(s, a) == (expect, o_arrow), (s, a, ob)
    # the wedge-parallel count of the same resident graph agrees
    assert dg.count(8, 4096) == expect
    dg.free()


def test_arrow_edge_cases(gpu_ctx):
    cases = [
        np.zeros((0, 2), np.uint32),                                   # no edges
        np.array([[0, 1], [2, 3], [4, 5]], np.uint32),                 # rows of one edge: no arrow wedges
        np.array([[0, 1], [1, 2]], np.uint32),                         # path
        np.array([[0, i] for i in range(1, 3000)], np.uint32),        # star
        np.array([[i, j] for i in range(300) for j in range(i + 1, 300)], np.uint32),  # K_300
        np.array([[0, 1], [1, 2], [2, 0]], np.uint32),                 # one triangle
    ]
    for pairs in cases:
        g = wtc.preprocess(wtc.graph_from_pairs(pairs))
        expect = O.tc_arrow(g, 0)
        dg = gpu_ctx.upload(g)
        for s, a in ((1, 0), (5, 8192), (17, 64)):
            assert dg.count_arrow(s, a) == expect, (len(pairs), s, a)
        dg.free()


def test_arrow_reference_style_entry_point():
    g = _rmat(12, seed=6)
    expect = O.tc_merge(g)
    t = wtc.GpuTime()
    assert wtc.tc_arrow_gpu(g, 5, 8192, t) == expect
    c1, e1 = t.copy, t.exec
    assert c1 > 0 and e1 > 0
    assert wtc.tc_arrow_gpu_big(g, 1, 0, t) == expect
    assert t.copy > c1 and t.exec > e1
    with pytest.raises(wtc.WtcError):
        wtc.tc_arrow_gpu(g, 0, 0)
    with pytest.raises(wtc.WtcError):
        wtc.tc_arrow_gpu_big(g, 5, 33)


@pytest.mark.parametrize("ob", [4, 8])
def test_baselines_rmat(gpu_ctx, ob):
    g = _rmat(13, seed=5)
    expect = O.tc_merge(g)
    dg = gpu_ctx.upload(g, offset_bytes=ob)
    for par in (wtc.PARALLEL_VERTEX, wtc.PARALLEL_EDGE):
        for style in (wtc.WEDGE_OUTGOING, wtc.WEDGE_ARROW, wtc.WEDGE_MIXED):
            for a in (0, 32, 4096):
                assert dg.count_baseline(par, style, a) == expect, (par, style, a, ob)
    dg.free()


def test_baselines_golden_and_edge_cases(gpu_ctx, golden):
    graphs = [(fx["file"], wtc.preprocess(wtc.read_graph(fx["path"], fx["format"] == "mm", fx["zero_indexed"])),
               fx["triangles"]) for fx in golden]
    for pairs in (np.zeros((0, 2), np.uint32), np.array([[0, 1], [1, 2]], np.uint32),
                  np.array([[0, i] for i in range(1, 3000)], np.uint32),
                  np.array([[i, j] for i in range(200) for j in range(i + 1, 200)], np.uint32)):
        g = wtc.preprocess(wtc.graph_from_pairs(pairs))
        graphs.append((len(pairs), g, O.tc_merge(g)))
    for name, g, expect in graphs:
        dg = gpu_ctx.upload(g)
        for par in (0, 1):
            for style in (0, 1, 2):
                for a in (0, 64):
                    assert dg.count_baseline(par, style, a) == expect, (name, par, style, a)
                    assert O.tc_style(g, style, a) == expect
        dg.free()


def test_baseline_reference_style_entry_points():
    g = _rmat(11, seed=8)
    expect = O.tc_merge(g)
    t = wtc.GpuTime()
    for fn in (wtc.tc_vertex_gpu, wtc.tc_edge_gpu):
        for big in (False, True):
            assert fn(g, 8192, wtc.WEDGE_MIXED, t, big=big) == expect
    assert t.copy > 0 and t.exec > 0
    with pytest.raises(wtc.WtcError):
        wtc.tc_vertex_gpu(g, 0, 3)
    with pytest.raises(wtc.WtcError):
        wtc.tc_edge_gpu(g, 33, 0)
