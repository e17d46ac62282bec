"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

How each oracle function is pinned (none is pinned by retyping its formula):
  orc_read_graph / orc_graph_from_pairs  hand-derived CSRs of tiny files;
        golden fixtures (networkx graphs, hashes from an independent numpy
        restatement)
  orc_preprocess        golden CSR hashes + orientation invariants
  orc_wedge_prefix      brute-force enumeration of out-neighbour pairs
  orc_search_closest(_tree)  numpy.searchsorted on random non-decreasing arrays
  orc_search_tree       the cached search equals the plain search everywhere
  orc_decode            exhaustive row-by-row enumeration (d <= 700) and exact
        row-boundary identities up to d = 2^32-1; also equal to the reference's
        floating-point formula (tc.cu:211-212) wherever the survey found it exact
  orc_bitmap_*          bijection onto [0, C(a,2)) and membership == list probe
  orc_tc_wedge / orc_tc_merge  closed-form families, O(n^3) brute force,
        networkx/scipy triangle counts of the golden fixtures
  orc_alg_bytes         brute-force per-wedge sum on small graphs
"""
import itertools
import os
import random

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def csr_from_edges(n, edges):
    """Symmetric CSR built directly in Python (ids already 0-based)."""
    adj = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    rp = np.zeros(n + 1, np.uint64)
    ci = []
    for v in range(n):
        ci.extend(sorted(adj[v]))
        rp[v + 1] = len(ci)
    return O.OGraph(n, len(ci), rp, np.array(ci, dtype=np.uint32))


def brute_triangles(n, edges):
    adj = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    return sum(1 for a, b, c in itertools.combinations(range(n), 3) if b in adj[a] and c in adj[a] and c in adj[b])


def oriented(pairs):
    return O.preprocess(O.graph_from_pairs(np.asarray(pairs, np.uint32).reshape(-1, 2)))


def counts_all(g, combos=((1, 0), (3, 32), (5, 64), (8, 8192), (17, 0), (2, 4096))):
    return {(s, a): O.tc_wedge(g, s, a) for s, a in combos} | {"merge": O.tc_merge(g, 2)} | \
        {("arrow", a): O.tc_arrow(g, a)[0] for a in (0, 32, 4096)} | \
        {("style", k, a): O.tc_style(g, k, a) for k in (0, 1, 2) for a in (0, 64)}


# ----------------------------------------------------------------- families

def family_cases():
    cases = []
    for n in (3, 4, 7, 25):
        cases.append((f"K{n}", [(i, j) for i in range(n) for j in range(i + 1, n)], n * (n - 1) * (n - 2) // 6))
    for p, q in ((1, 5), (3, 4), (10, 12)):
        cases.append((f"K{p},{q}", [(i, p + j) for i in range(p) for j in range(q)], 0))
    for n in (3, 4, 9):
        cases.append((f"C{n}", [(i, (i + 1) % n) for i in range(n)], 1 if n == 3 else 0))
    for n in (4, 5, 12):  # wheel: hub 0 + cycle 1..n
        e = [(0, i) for i in range(1, n + 1)] + [(i, i % n + 1) for i in range(1, n + 1)]
        cases.append((f"W{n}", e, n))
    for k in (1, 3, 10):  # windmill / friendship graph F_k
        e = []
        for t in range(k):
            a, b = 1 + 2 * t, 2 + 2 * t
            e += [(0, a), (0, b), (a, b)]
        cases.append((f"F{k}", e, k))
    pet = [(i, (i + 1) % 5) for i in range(5)] + [(i, i + 5) for i in range(5)] + \
          [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    cases.append(("Petersen", pet, 0))
    grid = [(r * 8 + c, r * 8 + c + 1) for r in range(8) for c in range(7)] + \
           [(r * 8 + c, (r + 1) * 8 + c) for r in range(7) for c in range(8)]
    cases.append(("grid8x8", grid, 0))
    for t in (1, 2, 5):  # complete tripartite K_{t,t,t}: t^3
        parts = [range(0, t), range(t, 2 * t), range(2 * t, 3 * t)]
        e = [(a, b) for x, y in ((0, 1), (0, 2), (1, 2)) for a in parts[x] for b in parts[y]]
        cases.append((f"K{t},{t},{t}", e, t ** 3))
    return cases


@pytest.mark.parametrize("name,edges,expect", family_cases(), ids=[c[0] for c in family_cases()])
def test_closed_form_families(name, edges, expect):
    g = oriented(edges)
    for k, v in counts_all(g).items():
        assert v == expect, (name, k)


def test_brute_force_random_small():
    rnd = random.Random(7)
    for trial in range(25):
        n = rnd.randint(3, 40)
        p = rnd.random() * 0.5
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rnd.random() < p]
        edges += [(v, u) for u, v in edges[::3]] + [(u, u) for u in range(0, n, 5)]  # dups + loops
        if not edges:
            continue
        expect = brute_triangles(n, edges)
        g = oriented(edges)
        for k, v in counts_all(g).items():
            assert v == expect, (trial, k)


def test_golden_fixtures(golden):
    import hashlib
    for fx in golden:
        og = O.read_graph(fx["path"], fx["format"] == "mm", fx["zero_indexed"])
        assert og.n == fx["n"] and og.m == 2 * fx["m"], fx["file"]
        g = O.preprocess(og)
        h = hashlib.sha256(g.row_ptr.astype("<u8").tobytes() + g.col_ind.astype("<u4").tobytes()).hexdigest()
        assert h == fx["csr_sha256"], fx["file"]
        assert int(O.wedge_prefix(g)[-1]) == fx["W"]
        assert O.tc_merge(g) == fx["triangles"]
        assert O.tc_wedge(g, 5, 64) == fx["triangles"]
        assert O.tc_wedge(g, 1, 0) == fx["triangles"]
        assert O.tc_arrow(g, 64)[0] == fx["triangles"]
        assert [O.tc_style(g, k, 64) for k in (0, 1, 2)] == [fx["triangles"]] * 3


def _arrow_total_brute(g):
    """Sum over edges (v,w) with d+(v) >= 2 of d+(w), vectorised."""
    rp = np.asarray(g.row_ptr, np.int64)
    d = np.diff(rp)
    src = np.repeat(np.arange(g.n), d)
    dst = np.asarray(g.col_ind, np.int64)
    return int(np.sum(np.where(d[src] >= 2, d[dst], 0)))


def test_arrow_prefix_and_total():
    # K_n in any orientation is a transitive tournament: every path v->w->u
    # closes, so arrow wedges = triangles = C(n,3)
    for n in (3, 4, 9, 30):
        g = oriented([(i, j) for i in range(n) for j in range(i + 1, n)])
        assert O.tc_arrow(g, 0) == (n * (n - 1) * (n - 2) // 6,) * 2
    # a path a-b-c oriented by degree has one edge per row: no arrow wedges
    assert O.tc_arrow(oriented([(0, 1), (1, 2)]), 0) == (0, 0)
    rnd = np.random.default_rng(4)
    for trial in range(6):
        n = int(rnd.integers(20, 300))
        pairs = rnd.integers(0, n, size=(int(rnd.integers(n, 6 * n)), 2))
        g = oriented(pairs)
        vi = O.arrow_prefix(g)
        assert vi[0] == 0 and np.all(np.diff(vi.astype(np.int64)) >= 0)
        assert int(vi[-1]) == _arrow_total_brute(g) == O.tc_arrow(g, 0)[1]
        assert O.tc_arrow(g, 0)[0] == O.tc_merge(g, 1)


# ------------------------------------------------------------------- decode

def test_decode_exhaustive_small():
    assert O.decode_check(2, 700, fp=True) == (0, 0)


def test_decode_row_boundaries_large():
    rnd = random.Random(3)
    for d in [2, 3, 65536, 65537, 92682, 92683, 1 << 20, (1 << 27) - 1, 1 << 28, (1 << 31) + 5, (1 << 32) - 1]:
        C = d * (d - 1) // 2
        rows = {0, d - 2} | {rnd.randrange(0, d - 1) for _ in range(30)}
        for w in rows:
            start = w * d - w * (w + 1) // 2  # wedges in rows < w
            assert O.decode(d, start) == (w, w + 1)
            last = start + (d - 1 - w) - 1
            assert O.decode(d, last) == (w, d - 1)
        assert O.decode(d, C - 1) == (d - 2, d - 1)


def test_decode_matches_reference_fp_midrange():
    # The reference's FP formula is exact below d ~ 2^27 (survey probe); the
    # integer decode must agree there, at row starts and ends.
    rnd = random.Random(11)
    for d in [1000, 4097, 65537, 1 << 20, (1 << 26) + 7]:
        for _ in range(40):
            w = rnd.randrange(0, d - 1)
            start = w * d - w * (w + 1) // 2
            for iv in (start, start + (d - 2 - w)):
                assert O.decode(d, iv) == O.decode_ref_fp(d, iv), (d, iv)


# ------------------------------------------------------------------- search

def test_search_closest_vs_searchsorted():
    rng = np.random.default_rng(2)
    for n in (1, 2, 3, 7, 100, 5000, 9000):
        inc = rng.integers(0, 4, size=n) * (rng.random(n) < 0.6)
        ws = np.concatenate([[0], np.cumsum(inc)]).astype(np.uint64)  # non-decreasing, ws[0] = 0
        W = int(ws[-1])
        tree = O.search_tree(ws, n)
        targets = range(W) if W < 3000 else rng.integers(0, W, size=3000)
        for t in targets:
            t = int(t)
            expect = int(np.searchsorted(ws[:n], t, side="right")) - 1
            assert O.search_closest(ws, 0, n, t) == expect
            assert O.search_closest_tree(tree, ws, 0, n, t) == expect


# ------------------------------------------------------------ wedge prefix

def test_wedge_prefix_brute_force():
    rnd = random.Random(5)
    for _ in range(10):
        n = rnd.randint(2, 60)
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rnd.random() < 0.3]
        if not edges:
            continue
        g = oriented(edges)
        ws = O.wedge_prefix(g)
        acc = 0
        for v in range(g.n):
            assert int(ws[v]) == acc
            row = g.col_ind[g.row_ptr[v]:g.row_ptr[v + 1]]
            acc += sum(1 for _ in itertools.combinations(row.tolist(), 2))
        assert int(ws[g.n]) == acc


# ------------------------------------------------------------------- bitmap

@pytest.mark.parametrize("n,a", [(100, 64), (40, 64), (70, 32), (33, 96), (200, 160)])
def test_bitmap_bijection_and_membership(n, a):
    rnd = random.Random(n * 1000 + a)
    edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rnd.random() < 0.25]
    g = oriented(edges)
    n = g.n
    thr = max(n, a) - a
    idx = [O.bitmap_index(n, a, w, u) for w in range(thr, n) for u in range(w + 1, n)]
    assert len(set(idx)) == len(idx)
    assert all(0 <= i < a * (a - 1) // 2 for i in idx)
    bm = O.bitmap_build(g, a)
    assert bm.size == (a * (a - 1) // 2 + 31) // 32
    for w in range(thr, n):
        row = set(g.col_ind[g.row_ptr[w]:g.row_ptr[w + 1]].tolist())
        for u in range(w + 1, n):
            i = O.bitmap_index(n, a, w, u)
            assert bool((int(bm[i >> 5]) >> (i & 31)) & 1) == (u in row)


# ------------------------------------------------------------------ reader

def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


READER_CASES = [
    # (name, text, mm, zero, expected n, expected sorted symmetric adjacency or error code)
    ("basic", "# c\n1 2\n2 3\n3 1\n", False, False, 3, {0: [1, 2], 1: [0, 2], 2: [0, 1]}),
    ("loops_dups", "%x\n1 1\n1 2\n2 1\n1 2\n4 4\n", False, False, 2, {0: [1], 1: [0]}),
    ("zero_ok", "0 3\n3 1\n", False, True, 4, {0: [3], 1: [3], 3: [0, 1]}),
    ("mm", "%%MatrixMarket\n% c\n3 3 2\n1 3\n2 3\n", True, False, 3, {0: [2], 1: [2], 2: [0, 1]}),
    ("three_cols_first", "1 2 7\n2 3\n", False, False, 3, {0: [1], 1: [0, 2], 2: [1]}),
    ("stops_at_garbage", "1 2\n2 3\nx 9\n3 4\n", False, False, 3, {0: [1], 1: [0, 2], 2: [1]}),
    ("crlf", "1 2\r\n2 3\r\n", False, False, 3, {0: [1], 1: [0, 2], 2: [1]}),
    ("gap", "1 9\n", False, False, 9, {0: [8], 8: [0]}),
    ("single_number", "5\n", False, False, 1, {}),
    ("zero_id", "0 1\n", False, False, None, -4),
    ("no_data", "# only comments\n%\n", False, False, None, -3),
    ("mm_size_only", "%%MM\n3 3 0\n", True, False, None, -3),
    ("long_header", "#" + "a" * 300 + "\n1 2\n", False, False, None, "long"),
]


@pytest.mark.parametrize("case", READER_CASES, ids=[c[0] for c in READER_CASES])
def test_reader_cases(tmp_path, case):
    name, text, mm, zero, n, expect = case
    path = _write(tmp_path, name, text)
    g = O.read_graph(path, mm, zero)
    if expect == "long":
        # fgets reads 255 chars of the header, then continues the same line:
        # 'aaa...' (no digit) is skipped, so the first data line is "1 2"
        assert g.n == 2 and g.m == 2
        return
    if isinstance(expect, int):
        assert g == expect
        return
    assert g.n == n
    adj = {v: g.col_ind[g.row_ptr[v]:g.row_ptr[v + 1]].tolist() for v in range(g.n)}
    assert {v: r for v, r in adj.items() if r} == expect


def test_reader_missing_file():
    assert O.read_graph("/nonexistent/graph.txt") == -2


# --------------------------------------------------------------- alg bytes

def test_alg_bytes_brute():
    import math
    rnd = random.Random(9)
    edges = [(u, v) for u in range(50) for v in range(u + 1, 50) if rnd.random() < 0.2]
    g = oriented(edges)
    d = np.diff(g.row_ptr.astype(np.int64))
    probes, W = 0, 0
    for v in range(g.n):
        row = g.col_ind[g.row_ptr[v]:g.row_ptr[v + 1]].tolist()
        for w, u in itertools.combinations(row, 2):
            probes += 4 + 4 * math.ceil(math.log2(d[w] + 1))
            W += 1
    base = 4 * (g.n + 1) + 4 * g.m + 8 * (g.n + 1)
    a, b = O.alg_bytes(g, 4)
    assert a == base + probes and b == base + 8 * W
