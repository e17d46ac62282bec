"""GPU parity at the measured workloads' full sizes, through the C ABI:

  * the 64-bit path (experiments/tc_big_graphs.cu:173-293, 394-420): K_3000
    (W = C(3000,3) > 2^32, every wedge closed, so every wedge range counts
    exactly e - b), R-MAT-24 with 4- and 8-byte row offsets against the
    oracle's whole-graph count, and R-MAT-28 (8-byte offsets, built on the
    GPU) against wedge-range splits and oracle counts of row samples;
  * the batching workload (4,096 ER/BA graphs) through the host-batch and
    the resident, graph-replayed paths, every per-graph count against the
    oracle's;
  * the serving mix (R-MAT scales 10..22) under Poisson arrivals below and
    above saturation, every completed result against the oracle's count.

The oracle's counts of these workloads are committed under tests/golden/
(tests/golden/make_big_golden.py); the tests regenerate the same seeded
graphs with the library's generators and host builders.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import workloads
import wtc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def big_golden():
    with open(os.path.join(GOLD, "big_golden.json")) as f:
        return json.load(f)


def _complete(k):
    i, j = np.triu_indices(k, 1)
    return wtc.preprocess(wtc.graph_from_pairs(np.stack([i, j], 1).astype(np.uint32)))


def test_k3000_wedges_above_2_32(gpu_ctx, big_golden):
    """K_3000: W = C(3000,3) = 4,495,501,000 > 2^32; with 4- and 8-byte
    offsets, several (s, a) (a = 3008 puts every w in the bitmap), the tc_GPU
    drop-ins, and wedge ranges that straddle 2^32: all closed, so a range
    [b, e) has exactly e - b triangles."""
    ref = big_golden["k3000"]
    g = _complete(3000)
    W = g.wedges()
    assert (g.n, g.m, W) == (ref["n"], ref["m"], ref["W"]) and W > 2 ** 32
    for ob in (4, 8):
        dg = gpu_ctx.upload(g, offset_bytes=ob)
        try:
            for s, a in ((12, 0), (5, 1024), (16, 3008), (1, 64)):
                cnt, st = dg.count(s, a, stats=True)
                assert cnt == ref["triangles"] and st.wedges == W, (ob, s, a)
            cuts = [0, 2 ** 31 + 3, 2 ** 32 - 7, 2 ** 32, 2 ** 32 + 5, 2 ** 32 + 64 * 12 + 1, W - 1, W]
            for b, e in zip(cuts[:-1], cuts[1:]):
                for s, a in ((12, 0), (7, 2048)):
                    assert dg.count(s, a, b, e) == e - b, (ob, b, e, s, a)
        finally:
            dg.free()
    assert wtc.tc_gpu_big(g, 8, 1024) == ref["triangles"]
    assert wtc.tc_gpu(g, 8, 0) == ref["triangles"]


@pytest.mark.parametrize("ob", [4, 8])
def test_rmat24_full_graph_against_oracle(gpu_ctx, big_golden, ob):
    """R-MAT-24 (the benchmark graph) counted whole with 4- and 8-byte row
    offsets, at the autotuner's (s, a) and two more, against the oracle's
    whole-graph count; n, m and W against the oracle's preprocessing."""
    ref = big_golden["rmat24"]
    g = workloads.rmat(24, seed=ref["seed"])
    assert (g.n, g.m, g.wedges()) == (ref["n"], ref["m"], ref["W"])
    dg = gpu_ctx.upload(g, offset_bytes=ob)
    del g
    try:
        s0, a0 = wtc.autotune(ref["n"], ref["m"], ref["W"])
        for s, a in ((s0, a0), (8, 16384), (16, 0)):
            cnt, st = dg.count(s, a, stats=True)
            assert cnt == ref["triangles"] and st.wedges == ref["W"], (ob, s, a)
    finally:
        dg.free()


def test_rmat28_u64_offsets_range_split_and_samples(gpu_ctx):
    """R-MAT-28 with 8-byte offsets (m > 2^32 directed pairs before
    orientation), built on the GPU: the whole count equals the sum of five
    wedge-range counts (cut points straddling 2^32 and spread over the wedge
    space), and the GPU counts of three row samples equal the oracle's merge
    count of the same rows."""
    dg = wtc.DeviceGraph.gen_rmat(gpu_ctx, 28, 16, seed=1, offset_bytes=8)
    try:
        assert dg.offset_bytes == 8
        g = dg.download()
        ws = np.concatenate([[0], np.cumsum(g.degrees * (g.degrees - 1) // 2)]).astype(np.uint64)
        W = int(ws[-1])
        assert W > 2 ** 40
        s, a = wtc.autotune(g.n, g.m, W)
        total = dg.count(s, a)
        cuts = [0, 2 ** 32 - 3, W // 3, (2 * W) // 3, W - 12345, W]
        parts = [dg.count(s, a, b, e) for b, e in zip(cuts[:-1], cuts[1:])]
        assert sum(parts) == total
        for frac in (0.05, 0.5, 0.95):
            v0 = int(np.searchsorted(ws, int(W * frac), side="right")) - 1
            v1 = int(np.searchsorted(ws, ws[v0] + 10 ** 8, side="right"))
            v1 = min(max(v1, v0 + 1), g.n)
            expect = O.tc_merge_range(g, v0, v1, 16)
            assert dg.count(s, a, int(ws[v0]), int(ws[v1])) == expect, (v0, v1)
            assert dg.count(7, 0, int(ws[v0]), int(ws[v1])) == expect, (v0, v1)
    finally:
        dg.free()


def test_batch_4096_against_oracle(gpu_ctx):
    """The batching workload at full size: 4,096 seeded ER/BA graphs (n
    log-uniform in [64, 65536], average degree in [4, 64]), every per-graph
    count of the host-batch path and of the resident batch's graph replays
    (autotuned s and a; and with the small-graph kernel off) equal to the
    oracle's."""
    ref = np.load(os.path.join(GOLD, "batch4096.npz"))
    k = len(ref["triangles"])
    graphs = workloads.batch_graphs(k, int(ref["base_seed"]))
    assert [g.n for g in graphs] == ref["n"].tolist()
    assert [g.m for g in graphs] == ref["m"].tolist()
    assert [g.wedges() for g in graphs] == ref["W"].tolist()
    expect = ref["triangles"].astype(np.uint64)
    got = wtc.batch_count_host(gpu_ctx, graphs)
    assert np.array_equal(got, expect)
    default, _ = gpu_ctx.small_graphs()
    try:
        for limit in (default, 0):
            gpu_ctx.set_small_graphs(limit)
            b = wtc.Batch(gpu_ctx, graphs)
            try:
                for rep in range(3):
                    assert np.array_equal(b.count(), expect), (limit, rep)
            finally:
                b.free()
    finally:
        gpu_ctx.set_small_graphs(default)


def test_serving_mix_poisson_against_oracle(big_golden):
    """The serving mix (26 R-MAT graphs, scales 10..22) under Poisson
    arrivals at a rate the server keeps up with and one far above
    saturation (queue-full rejections allowed): every completed request's
    count equals the oracle's count of its graph."""
    ref = big_golden["serve_mix"]
    keys = workloads.serve_keys()
    assert [(e["scale"], e["seed"]) for e in ref] == keys
    graphs = workloads.serve_graphs()
    assert [(g.n, g.m, g.wedges()) for g in graphs] == [(e["n"], e["m"], e["W"]) for e in ref]
    expect = [e["triangles"] for e in ref]
    params = [wtc.autotune(g.n, g.m, g.wedges()) for g in graphs]
    srv = wtc.Server(0, hbm_budget_bytes=32 << 30, max_batch=256, slots=2, queue_capacity=256)
    try:
        gids = [srv.register(g) for g in graphs]
        del graphs
        for lam, seconds in ((100.0, 1.5), (4000.0, 1.0)):
            pick, results, _ = workloads.poisson_run(srv, gids, params, lam, seconds, seed=int(lam))
            done = [(p, r) for p, r in zip(pick, results) if r is not None]
            assert len(done) >= min(len(pick), 100), lam
            bad = [(int(p), r[0]) for p, r in done if r[0] != expect[p]]
            assert not bad, (lam, bad[:5])
        st = srv.stats()
        assert st["completed"] == st["submitted"]
    finally:
        srv.close()
