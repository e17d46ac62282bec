"""GPU parity of the split count (KERNEL_SPLIT): every in-edge grouped by its
middle vertex in a per-count transposition, the top-a rows' wedges counted
against an LDS bitmap row, the others against LDS windows of colInd (or HBM
for rows longer than a window holds).  Every count, whole graph or wedge
range, must equal the oracle's count of the same wedges (tc.cu:173-293 over
[w_begin, w_end))."""
import numpy as np
import pytest

import oracle as O
import wtc

pytestmark = pytest.mark.gpu


def _rmat(scale, ef=16, seed=1):
    return wtc.preprocess(wtc.graph_from_pairs(wtc.gen_rmat_pairs(scale, ef, seed=seed)))


def _complete(k):
    i, j = np.triu_indices(k, 1)
    return wtc.preprocess(wtc.graph_from_pairs(np.stack([i, j], axis=1).astype(np.uint32)))


@pytest.fixture
def split_ctx(gpu_ctx):
    gpu_ctx.set_count_kernel(wtc.KERNEL_SPLIT)
    try:
        yield gpu_ctx
    finally:
        gpu_ctx.set_count_kernel(wtc.KERNEL_WALK)


def _ranges(W, rng, k=6):
    """Whole order, halves and thirds, an empty range, a one-wedge range,
    ranges past the end and k random ranges."""
    out = [(0, 2 ** 64 - 1), (0, W), (0, W // 2), (W // 2, W), (W // 3, 2 * W // 3), (W // 7, W // 7),
           (W // 5, W // 5 + 1), (W - 1, W + 100), (W, 2 ** 64 - 1)]
    for _ in range(k):
        x, y = sorted(int(t) for t in rng.integers(0, W + 1, size=2))
        out.append((x, y))
    return out


@pytest.mark.parametrize("ob,narrow", [(4, True), (8, True), (8, False)])
def test_split_whole_and_ranges_match_oracle(split_ctx, monkeypatch, ob, narrow):
    """R-MAT, Barabasi-Albert and Erdos-Renyi graphs at several bitmap
    lengths (a below n, a at n, a above n so that every vertex is a top
    row), whole graph and wedge ranges straddling vertices, against the
    oracle's range count.  With 8-byte offsets the split count runs on a
    4-byte copy of them (every position fits 32 bits), or, with
    $WTC_SPLIT_NARROW=0, on 8-byte offsets and records throughout."""
    if not narrow:
        monkeypatch.setenv("WTC_SPLIT_NARROW", "0")
    rng = np.random.default_rng(5)
    cases = [(_rmat(12, seed=3), (32, 1024, 4096)),
             (_rmat(10, ef=32, seed=4), (64, 1024)),
             (wtc.preprocess(wtc.gen_barabasi_albert(3000, 7, seed=9)), (32, 2048, 3008)),
             (wtc.preprocess(wtc.gen_erdos_renyi(700, 0.05, seed=2)), (96, 704, 4096))]
    for g, alens in cases:
        W = g.wedges()
        expect = O.tc_merge(g)
        dg = split_ctx.upload(g, offset_bytes=ob)
        for a in alens:
            for s in (1, 4, 12):
                assert dg.count(s, a) == expect, (g.n, a, s)
            for lo, hi in _ranges(W, rng):
                got = dg.count(8, a, lo, hi)
                assert got == O.tc_wedge(g, 8, a, lo, hi), (g.n, a, lo, hi)
        dg.free()


def test_split_complete_graph_closed_form(split_ctx):
    """K_400: every wedge closes, so every range [lo, hi) holds hi - lo
    triangles (clipped tails at both ends included)."""
    g = _complete(400)
    W = g.wedges()
    assert W == 400 * 399 * 398 // 6
    dg = split_ctx.upload(g)
    rng = np.random.default_rng(11)
    for a in (32, 128, 416):
        assert dg.count(5, a) == W
        for lo, hi in _ranges(W, rng, k=10):
            lo_c, hi_c = min(lo, W), min(hi, W)
            assert dg.count(5, a, lo, hi) == max(0, hi_c - lo_c), (a, lo, hi)
    dg.free()


def test_split_sum_of_parts(split_ctx, monkeypatch):
    """A five-way range split of an R-MAT-14 graph sums to the whole count
    at u32 offsets and at u64 offsets (narrowed to a 4-byte copy, and kept
    8-byte with $WTC_SPLIT_NARROW=0), all equal to the oracle's count."""
    g = _rmat(14, seed=8)
    expect = O.tc_merge(g)
    W = g.wedges()
    cuts = [0, W // 9, W // 3, W // 2, (5 * W) // 6, W]
    for ob, narrow in ((4, "1"), (8, "1"), (8, "0")):
        monkeypatch.setenv("WTC_SPLIT_NARROW", narrow)
        dg = split_ctx.upload(g, offset_bytes=ob)
        for a in (512, 2048):
            assert dg.count(12, a) == expect
            parts = [dg.count(12, a, cuts[i], cuts[i + 1]) for i in range(5)]
            assert sum(parts) == expect, (ob, narrow, a, parts)
        dg.free()


def test_split_stats_report_the_full_order(split_ctx):
    """Stats of a split count report the wedges of the counted range of the
    full order (not of the list part alone)."""
    g = _rmat(13, seed=12)
    expect = O.tc_merge(g)
    W = g.wedges()
    dg = split_ctx.upload(g)
    cnt, st = dg.count(6, 2048, stats=True)
    assert cnt == expect
    assert st.wedges == W
    cnt, st = dg.count(6, 2048, W // 4, W // 2, stats=True)
    assert cnt == O.tc_wedge(g, 6, 2048, W // 4, W // 2)
    assert st.wedges == W // 2 - W // 4
    dg.free()


def _partition(g, a):
    """(top wedges, top in-edges, list in-edges) of the whole order: in-edges
    v -> w with a non-empty tail, split by w >= thr (thr = n - a, or 0 when
    a >= n), from the CSR alone."""
    rp = np.asarray(g.row_ptr, np.int64)
    ci = np.asarray(g.col_ind, np.int64)
    d = np.diff(rp)
    src = np.repeat(np.arange(g.n), d)
    tail = d[src] - (np.arange(g.m) - rp[src]) - 1
    thr = g.n - a if a < g.n else 0
    top = (ci >= thr) & (tail > 0)
    lst = (ci < thr) & (tail > 0)
    return int(tail[top].sum()), int(top.sum()), int(lst.sum())


def test_split_stats_partition_the_wedges(split_ctx):
    """The row kernel's wedges and both kernels' in-edges equal the CSR's own
    split of the in-edges at thr, for several bitmap lengths."""
    g = _rmat(13, seed=12)
    expect = O.tc_merge(g)
    W = g.wedges()
    dg = split_ctx.upload(g)
    for a in (32, 512, 2048, 8192):
        cnt, st = dg.count(6, a, stats=True)
        assert cnt == expect and st.split == 1 and st.wedges == W
        assert (st.top_wedges, st.in_edges, st.list_in_edges) == _partition(g, a), a
    dg.free()


def test_split_long_list_rows_closed_form(split_ctx):
    """K_4200 with a = 32 or 64: the list rows w < 103 have out-degree above
    the slab window's row limit (4096), so their ids are searched in HBM;
    every wedge closes, so any range [lo, hi) holds hi - lo triangles."""
    g = _complete(4200)
    W = g.wedges()
    dg = split_ctx.upload(g)
    rng = np.random.default_rng(17)
    for a in (32, 64):
        assert dg.count(8, a) == W
        for lo, hi in _ranges(W, rng, k=4):
            lo_c, hi_c = min(lo, W), min(hi, W)
            assert dg.count(8, a, lo, hi) == max(0, hi_c - lo_c), (a, lo, hi)
    dg.free()


def test_split_packed_layouts_match_oracle(split_ctx):
    """Every packing of the top tails' ids, at and around its row limit, on
    R-MAT-20, whole and a range, against the oracle: eight 16-bit ids per
    chunk (a = 2^16), seven 18-bit (a = 2^16 + 32, 2^18), six 21-bit fields
    holding the LDS word address and bit (a = 2^18 + 32, 2^19) and six
    21-bit ids (a = 786,432 with thr > 0, and a = 2^20 where every vertex is
    a top row)."""
    g = _rmat(20, seed=2)
    assert g.n > (1 << 19)
    expect = O.tc_merge(g)
    W = g.wedges()
    dg = split_ctx.upload(g)
    for a in (1 << 16, (1 << 16) + 32, 1 << 18, (1 << 18) + 32, 1 << 19, 786432, 1 << 20):
        cnt, st = dg.count(8, a, stats=True)
        assert cnt == expect and st.split == 1, a
        lo, hi = W // 3, W // 3 + W // 500
        assert dg.count(8, a, lo, hi) == O.tc_wedge(g, 8, a, lo, hi), a
    dg.free()
