"""GPU parity: the HIP path (through the C ABI) against the oracle and the
golden fixtures.  Integer counts must match bit for bit."""
import numpy as np
import pytest

import oracle as O
import wtc

pytestmark = pytest.mark.gpu

SPREADS = (1, 2, 3, 5, 8, 16, 17, 40)
ALENS = (0, 32, 64, 4096, 8192)


def _rmat(scale, ef=16, seed=1):
    pairs = wtc.gen_rmat_pairs(scale, ef, seed=seed)
    g = wtc.preprocess(wtc.graph_from_pairs(pairs))
    return g


def test_golden_fixtures(gpu_ctx, golden):
    for fx in golden:
        g = wtc.preprocess(wtc.read_graph(fx["path"], fx["format"] == "mm", fx["zero_indexed"]))
        dg = gpu_ctx.upload(g)
        for s in (1, 5, 8):
            for a in (0, 32, 64, 8192):
                cnt, st = dg.count(s, a, stats=True)
                assert cnt == fx["triangles"], (fx["file"], s, a)
                assert st.wedges == fx["W"], fx["file"]
        dg.free()


def _square_bits(words, a):
    """Device bitmap words (a x a bits, row-major, 32-bit words) as a
    boolean a x a matrix."""
    bits = np.unpackbits(np.asarray(words, np.uint32).view(np.uint8), bitorder="little")
    return bits.reshape(a, a).astype(bool)


def test_device_prefix_tree_bitmap_match_oracle(gpu_ctx):
    """Device wedge prefix and search tree equal the oracle's word for word.
    The device keeps the top-a bitmap (tc.cu:329-346) as a square a x a bit
    matrix over the ranks [thr, thr + a): its bits must be exactly the
    upper-triangular adjacency of those ranks, and agree bit for bit with
    the oracle's packed bitmap through its index (tc.cu:263)."""
    for scale, seed in ((8, 1), (11, 2), (13, 3)):
        g = _rmat(scale, seed=seed)
        dg = gpu_ctx.upload(g)
        for a in (32, 64, 1024):
            dg.count(3, a)
            ws, tree = gpu_ctx.debug_prefix(g.n)
            ows = O.wedge_prefix(g)
            assert np.array_equal(ws, ows)
            assert np.array_equal(tree, O.search_tree(ows, g.n))
            sq = _square_bits(gpu_ctx.debug_bitmap(a * a // 32), a)
            thr = max(g.n, a) - a
            expect = np.zeros((a, a), bool)
            rp = g.row_ptr.astype(np.int64)
            for w in range(thr, g.n):
                expect[w - thr, g.col_ind[rp[w]:rp[w + 1]].astype(np.int64) - thr] = True
            assert np.array_equal(sq, expect), (scale, a)
            packed = np.unpackbits(O.bitmap_build(g, a).view(np.uint8), bitorder="little")
            rng = np.random.default_rng(scale)
            rows = min(g.n, a)
            for _ in range(400):
                w = thr + int(rng.integers(0, max(rows - 1, 1)))
                u = w + 1 + int(rng.integers(0, max(g.n - w - 1, 1)))
                if u >= g.n:
                    continue
                assert bool(packed[O.bitmap_index(g.n, a, w, u)]) == sq[w - thr, u - thr], (scale, a, w, u)
        dg.free()


def test_decode_full_domain(gpu_ctx):
    rng = np.random.default_rng(5)
    d_list, iv_list = [], []
    for d in [2, 3, 4, 5, 17, 255, 65536, 65537, 92682, 92683, 1 << 20, (1 << 28) + 3, (1 << 31) - 1,
              (1 << 31), (1 << 32) - 1]:
        C = d * (d - 1) // 2
        ivs = {0, C - 1, C // 2}
        for w in rng.integers(0, d - 1, size=20).tolist() + [0, d - 2]:
            start = w * d - w * (w + 1) // 2  # first wedge of row w
            ivs.update({start, max(start - 1, 0), min(start + 1, C - 1)})
        for iv in sorted(ivs):
            d_list.append(d)
            iv_list.append(iv)
    w, u = gpu_ctx.debug_decode(d_list, iv_list)
    for k, (d, iv) in enumerate(zip(d_list, iv_list)):
        assert (int(w[k]), int(u[k])) == O.decode(d, iv), (d, iv)


@pytest.mark.parametrize("ob", [4, 8])
def test_rmat_sweep(gpu_ctx, ob):
    g = _rmat(14, seed=9)
    expect = O.tc_merge(g)
    dg = gpu_ctx.upload(g, offset_bytes=ob)
    for s in SPREADS:
        for a in ALENS:
            assert dg.count(s, a) == expect, (s, a, ob)
    dg.free()


def test_wedge_ranges(gpu_ctx):
    g = _rmat(12, seed=4)
    W = g.wedges()
    dg = gpu_ctx.upload(g)
    total = dg.count(5, 64)
    cuts = sorted({0, W, *np.random.default_rng(1).integers(0, W, size=9).tolist()})
    parts = [dg.count(5, 64, b, e) for b, e in zip(cuts[:-1], cuts[1:])]
    assert sum(parts) == total
    for b, e in list(zip(cuts[:-1], cuts[1:]))[:4]:
        assert dg.count(3, 0, b, e) == O.tc_wedge(g, 3, 0, b, e)
    assert dg.count(5, 64, W, W + 10) == 0
    dg.free()


def test_reference_style_entry_points():
    g = _rmat(12, seed=6)
    expect = O.tc_merge(g)
    t = wtc.GpuTime()
    assert wtc.tc_gpu(g, 5, 8192, t) == expect
    c1, e1 = t.copy, t.exec
    assert c1 > 0 and e1 > 0
    assert wtc.tc_gpu_big(g, 5, 8192, t) == expect
    assert t.copy > c1 and t.exec > e1  # accumulated with +=, like GPU_time


def test_edge_cases(gpu_ctx):
    cases = [
        np.zeros((0, 2), np.uint32),                                   # no edges: n = 1
        np.array([[0, 1], [2, 3], [4, 5]], np.uint32),                 # no wedges
        np.array([[0, i] for i in range(1, 3000)], np.uint32),        # star: W = 0 after orientation
        np.array([[i, j] for i in range(300) for j in range(i + 1, 300)], np.uint32),  # K_300, d+ up to 299
        np.array([[0, 1], [1, 2], [2, 0]], np.uint32),                 # one triangle
    ]
    expect = [0, 0, 0, 300 * 299 * 298 // 6, 1]
    for pairs, ex in zip(cases, expect):
        g = wtc.preprocess(wtc.graph_from_pairs(pairs))
        dg = gpu_ctx.upload(g)
        for s, a in ((1, 0), (5, 8192), (17, 64)):
            assert dg.count(s, a) == ex
        dg.free()


def test_batch_exact(gpu_ctx):
    graphs, expect = [], []
    for i in range(40):
        if i % 2:
            g = wtc.preprocess(wtc.gen_erdos_renyi(50 + 37 * i, 8.0 / (50 + 37 * i), seed=i))
        else:
            g = wtc.preprocess(wtc.gen_barabasi_albert(60 + 29 * i, 2 + i % 7, seed=i))
        graphs.append(g)
        expect.append(O.tc_merge(g))
    b = wtc.Batch(gpu_ctx, graphs)
    for spread, a in ((None, None), (1, 0), (5, 64), ([1 + (i % 9) for i in range(40)], [32 * (i % 4) for i in range(40)])):
        got = b.count(spread, a)
        assert got.tolist() == expect, (spread, a)
    b.free()
    got = wtc.batch_count_host(gpu_ctx, graphs)
    assert got.tolist() == expect


@pytest.mark.parametrize("wpb", [1, 2, 4, 8])
@pytest.mark.parametrize("ways", [1, 2, 4, 8])
def test_probe_variants(gpu_ctx, ways, wpb):
    """Both counting kernels (the row walk and the LDS-staged kernel), every
    variant (probe_ways: steps or probes in flight) at every block shape
    (waves per block) and, for the staged kernel, every LDS chunk give the
    same counts, on single graphs (u32 and u64 offsets) and on batches.
    Chunks 8 and 16 (and the 4-wedge tail of s = 12 at chunk 8) divide 64,
    which takes the staged kernel's linear LDS read path; 3, 7 and 12 take
    the general one."""
    g = _rmat(13, seed=21)
    expect = O.tc_merge(g)
    k300 = wtc.preprocess(wtc.graph_from_pairs(
        np.array([[i, j] for i in range(300) for j in range(i + 1, 300)], np.uint32)))
    graphs = [wtc.preprocess(wtc.gen_barabasi_albert(70 + 13 * i, 1 + i % 6, seed=40 + i)) for i in range(17)]
    bexp = [O.tc_merge(x) for x in graphs]
    try:
        for kernel, chunk in [(wtc.KERNEL_WALK, 16)] + [(wtc.KERNEL_STAGED, c) for c in (3, 8, 12, 16)]:
            gpu_ctx.set_count_kernel(kernel)
            gpu_ctx.set_tuning(ways, chunk, wpb)
            for ob in (4, 8):
                dg = gpu_ctx.upload(g, offset_bytes=ob)
                for s, a in ((1, 0), (4, 0), (7, 64), (12, 4096), (16, 4096)):
                    assert dg.count(s, a) == expect, (ways, wpb, chunk, ob, s, a)
                dg.free()
            dk = gpu_ctx.upload(k300)
            assert dk.count(5, 0) == 300 * 299 * 298 // 6
            dk.free()
            b = wtc.Batch(gpu_ctx, graphs)
            assert b.count(3, 0).tolist() == bexp
            assert b.count(6, 32).tolist() == bexp
            b.free()
    finally:
        gpu_ctx.set_tuning()
        gpu_ctx.set_count_kernel(wtc.KERNEL_WALK)


def test_batch_graph_replays(gpu_ctx):
    """A resident batch replays its captured graph: every replay, with eager
    work and synchronisations between them, equals the eager count; a new
    (spread, a) or tuning captures anew and stays exact."""
    graphs = []
    for i in range(24):
        if i % 2:
            graphs.append(wtc.preprocess(wtc.gen_erdos_renyi(80 + 51 * i, 10.0 / (80 + 51 * i), seed=300 + i)))
        else:
            graphs.append(wtc.preprocess(wtc.gen_barabasi_albert(90 + 41 * i, 2 + i % 5, seed=300 + i)))
    expect = [O.tc_merge(g) for g in graphs]
    eager = wtc.batch_count_host(gpu_ctx, graphs)
    assert eager.tolist() == expect
    side = _rmat(11, seed=33)
    side_expect = O.tc_merge(side)
    b = wtc.Batch(gpu_ctx, graphs)
    try:
        for rep in range(5):
            got, st = b.count(stats=True)
            assert got.tolist() == expect, rep
            assert st.wedges == sum(g.wedges() for g in graphs)
            dg = gpu_ctx.upload(side)  # eager work on the context's own stream between replays
            assert dg.count(3, 64) == side_expect
            dg.free()
        for spread, a in ((2, 32), ([1 + i % 7 for i in range(24)], [32 * (i % 3) for i in range(24)])):
            for rep in range(3):
                assert b.count(spread, a).tolist() == expect, (spread, a, rep)
        gpu_ctx.set_tuning(2, 8, 2)
        for rep in range(3):
            assert b.count(4, 64).tolist() == expect, rep
    finally:
        gpu_ctx.set_tuning()
        b.free()


def test_chung_lu_2e18_full_size(gpu_ctx):
    """The size of the usage example of tc.cu (Amazon0302) as a Chung-Lu graph: n = 262,144, ~900k edges,
    power-law expected degrees (exponent 2.7 on [1, 420]); host preprocessing,
    the oracle's count, then s=5, a=8192 on the GPU (and the tc_GPU drop-ins)."""
    g = wtc.preprocess(wtc.gen_chung_lu(262144, 2.7, 1, 420, 900000, seed=1))
    assert g.m == 900000
    expect = O.tc_merge(g)
    assert expect > 0
    dg = gpu_ctx.upload(g)
    assert dg.count(5, 8192) == expect
    dg.free()
    assert wtc.tc_gpu(g, 5, 8192) == expect
    assert wtc.tc_gpu_big(g, 5, 8192) == expect


def test_rmat20_sa_sweep_full_size(gpu_ctx):
    """The reference (s, a) sweep at full size: R-MAT scale 20, edge factor 16, u32 ids and
    offsets, every (s, a) of the sweep exact against the oracle."""
    g = wtc.preprocess(wtc.gen_rmat(20, 16, seed=1))
    expect = O.tc_merge(g)
    dg = gpu_ctx.upload(g, offset_bytes=4)
    for s in (1, 3, 5, 8):
        for a in (0, 4096, 8192, 16384):
            assert dg.count(s, a) == expect, (s, a)
    dg.free()


def _small_set():
    """Small graphs of every shape the small-graph kernel must handle: empty,
    wedge-free, complete, star, sparse and dense random graphs."""
    k = lambda n: wtc.preprocess(wtc.graph_from_pairs(  # noqa: E731
        np.array([[i, j] for i in range(n) for j in range(i + 1, n)], np.uint32).reshape(-1, 2)))
    graphs = [
        k(3),
        wtc.preprocess(wtc.graph_from_pairs(np.zeros((0, 2), np.uint32))),           # n = 1, no edges
        wtc.preprocess(wtc.graph_from_pairs(np.array([[0, 1], [2, 3]], np.uint32))),  # no wedges
        k(140),                                                                       # W = C(140,3) on one block
        wtc.preprocess(wtc.graph_from_pairs(np.array([[0, i] for i in range(1, 900)], np.uint32))),  # star
    ]
    for i in range(30):
        n = 40 + 37 * i
        if i % 2:
            graphs.append(wtc.preprocess(wtc.gen_erdos_renyi(n, min(1.0, (4 + i % 12) / (n - 1)), seed=500 + i)))
        else:
            graphs.append(wtc.preprocess(wtc.gen_barabasi_albert(n, 1 + i % 9, seed=500 + i)))
    graphs.append(k(40))
    return graphs


def test_small_graph_kernel(gpu_ctx):
    """Kernel choice per set: a set of small graphs that all fit the limit is
    counted by the LDS small-graph kernel, any other set by the wedge-parallel
    kernel; both give the oracle's counts (graph replays and eager launches),
    at every limit from off to the device maximum.  The limit is inclusive;
    a set with one larger graph, and a single graph, never take it."""
    small = _small_set()
    words = [2 * (g.n + 1) + g.m for g in small]
    wmax = max(words)
    big = _rmat(12, seed=5)
    mixed = small[:10] + [big] + small[10:]
    expect = [O.tc_merge(g) if g.m else 0 for g in small]
    big_expect = O.tc_merge(big)
    mixed_expect = expect[:10] + [big_expect] + expect[10:]
    default, dev_max = gpu_ctx.small_graphs()
    assert 0 < default <= dev_max
    assert dev_max * 4 <= 160 * 1024
    assert wmax <= default < 2 * (big.n + 1) + big.m
    try:
        for limit in sorted({0, 64, 4096, wmax - 1, wmax, default, dev_max}):
            gpu_ctx.set_small_graphs(limit)
            assert gpu_ctx.small_graphs() == (limit, dev_max)
            n_small = len(small) if 0 < limit and wmax <= limit else 0
            for graphs, exp, ns in ((small, expect, n_small), (mixed, mixed_expect, 0)):
                b = wtc.Batch(gpu_ctx, graphs)
                for spread, a in ((None, None), (3, 64)):
                    for rep in range(2):
                        got, st = b.count(spread, a, stats=True)
                        assert got.tolist() == exp, (limit, spread, a, rep)
                        assert st.small_graphs == ns, (limit, st.small_graphs, ns)
                b.free()
                got, st = wtc.batch_count_host(gpu_ctx, graphs, stats=True)
                assert got.tolist() == exp, limit
                assert st.small_graphs == ns
        # a single graph never takes it (one block would hold the whole count)
        gpu_ctx.set_small_graphs(dev_max)
        got, st = wtc.batch_count_host(gpu_ctx, [small[3]], stats=True)
        assert got.tolist() == [expect[3]] and st.small_graphs == 0
        with pytest.raises(wtc.WtcError):
            gpu_ctx.set_small_graphs(dev_max + 1)
    finally:
        gpu_ctx.set_small_graphs(default)


def test_batch_upload_checks_and_reuse(gpu_ctx):
    """The batch upload checks every CSR while packing it (a graph that is not
    an oriented CSR fails the whole call with WTC_ERR_CSR, and the context
    stays usable), and repeated host-batch counts of different sizes reuse
    the context's device buffers with exact results."""
    graphs = [wtc.preprocess(wtc.gen_barabasi_albert(200 + 300 * i, 2 + i % 5, seed=900 + i)) for i in range(12)]
    expect = [O.tc_merge(g) for g in graphs]
    good = graphs[3]
    bad_cases = []
    ci = good.col_ind.copy()
    r0, r1 = int(good.row_ptr[5]), int(good.row_ptr[6])
    assert r1 - r0 >= 2
    ci[r0], ci[r0 + 1] = ci[r0 + 1], ci[r0]                    # a row out of order
    bad_cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci))
    ci = good.col_ind.copy()
    ci[r0] = 5                                                  # an id not above its row
    bad_cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci))
    ci = good.col_ind.copy()
    ci[-1] = good.n                                             # an id out of range
    bad_cases.append(wtc.CSR(good.n, good.m, good.row_ptr.copy(), ci))
    for bad in bad_cases:
        for where in (0, 6, 12):
            mix = graphs[:where] + [bad] + graphs[where:]
            with pytest.raises(wtc.WtcError) as ei:
                wtc.batch_count_host(gpu_ctx, mix)
            assert ei.value.code == -10
            with pytest.raises(wtc.WtcError):
                wtc.Batch(gpu_ctx, mix)
        assert wtc.batch_count_host(gpu_ctx, graphs).tolist() == expect
    for k in (12, 3, 12, 1, 7):
        assert wtc.batch_count_host(gpu_ctx, graphs[:k]).tolist() == expect[:k], k


def test_tiles_across_wedge_free_runs(gpu_ctx):
    """Tiles whose vertex range crosses runs of more than 64 wedge-free
    vertices (the tile setup's coalesced 64-vertex window then skips ahead
    or falls back to searches in global memory): 40 disjoint K4 blocks, each
    followed in rank order by 200 degree-3 claw centres that have no wedges,
    counted alone at several spreads and bitmap lengths and inside a batch."""
    B, C = 40, 200
    pairs, leaf = [], 10 ** 6
    for b in range(B):
        base = b * (4 + C)
        pairs += [(base + i, base + j) for i in range(4) for j in range(i + 1, 4)]
        for c in range(C):
            for _ in range(3):
                pairs.append((base + 4 + c, leaf))
                leaf += 1
    g = wtc.preprocess(wtc.graph_from_pairs(np.array(pairs, np.uint32)))
    d = np.diff(np.asarray(g.row_ptr, np.int64))
    gaps = np.diff(np.nonzero(d * (d - 1) // 2)[0])
    assert (gaps > 64).sum() >= 30  # the runs the test is about
    assert O.tc_merge(g) == 4 * B
    for ob in (4, 8):
        dg = gpu_ctx.upload(g, offset_bytes=ob)
        for s, a in ((1, 0), (2, 32), (3, 0), (5, 4096), (1, 64)):
            assert dg.count(s, a) == 4 * B, (ob, s, a)
        dg.free()
    other = wtc.preprocess(wtc.gen_barabasi_albert(500, 3, seed=77))
    got = wtc.batch_count_host(gpu_ctx, [other, g, other], [2, 1, 3], [0, 0, 32])
    assert got.tolist() == [O.tc_merge(other), 4 * B, O.tc_merge(other)]


def test_host_batch_chunked_pipeline(gpu_ctx, monkeypatch):
    """A host batch counted chunk by chunk (upload chunk forced down to a few
    KiB, so each chunk holds one or a few graphs and the two execution slots
    alternate): per-graph counts equal the oracle, stats sum over the chunks,
    and an invalid CSR in a late chunk still fails the whole call."""
    graphs = [wtc.preprocess(wtc.gen_erdos_renyi(100 + 90 * i, 8.0 / (100 + 90 * i), seed=1200 + i)) if i % 2 else
              wtc.preprocess(wtc.gen_barabasi_albert(120 + 70 * i, 2 + i % 4, seed=1200 + i)) for i in range(30)]
    expect = [O.tc_merge(g) for g in graphs]
    whole, st1 = wtc.batch_count_host(gpu_ctx, graphs, stats=True)
    assert whole.tolist() == expect and st1.launches == 1
    for chunk in (4096, 65536):
        monkeypatch.setenv("WTC_UPLOAD_CHUNK_BYTES", str(chunk))
        got, st = wtc.batch_count_host(gpu_ctx, graphs, stats=True)
        assert got.tolist() == expect, chunk
        assert st.launches > 2 and st.wedges == st1.wedges
        got = wtc.batch_count_host(gpu_ctx, graphs, [1 + i % 5 for i in range(30)], [32 * (i % 3) for i in range(30)])
        assert got.tolist() == expect
        bad = graphs[25]
        ci = bad.col_ind.copy()
        ci[-1] = bad.n  # an id out of range
        mix = graphs[:25] + [wtc.CSR(bad.n, bad.m, bad.row_ptr.copy(), ci)] + graphs[26:]
        with pytest.raises(wtc.WtcError) as ei:
            wtc.batch_count_host(gpu_ctx, mix)
        assert ei.value.code == -10
        assert wtc.batch_count_host(gpu_ctx, graphs).tolist() == expect
    monkeypatch.delenv("WTC_UPLOAD_CHUNK_BYTES")
