"""Writes the golden fixtures under tests/golden/.

Each fixture is an input graph file in the reference's formats (SNAP-style
edge list for `-e`, Matrix Market for `-m`) plus the expected results, all
computed here WITHOUT the code under test:

  triangles  networkx.triangles (sum/3), cross-checked with scipy trace(A^3)/6
  n, m       a numpy restatement of read_graph/preprocess (tc.cu:561-787):
             n = max id + 1, m = undirected edges after dedup / loop removal
  W          sum of C(d+,2) after the stable degree order of tc.cu:697-716
  csr_sha256 hash of the oriented CSR (rowPtr as uint64 LE, colInd as uint32 LE)

The reference publishes no tests or vectors of its own; its one known answer
(Amazon0302 -> 717,719 triangles, README.md:32-39) needs an input file that is
not in the reference tree.  Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os

import networkx as nx
import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def oriented_csr(n, und_edges):
    """numpy restatement: symmetric dedup CSR -> stable degree rank -> oriented."""
    e = np.asarray(und_edges, dtype=np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    both = np.concatenate([e, e[:, ::-1]]) if len(e) else np.zeros((0, 2), np.int64)
    both = np.unique(both, axis=0)
    deg = np.bincount(both[:, 0], minlength=n) if len(both) else np.zeros(n, np.int64)
    order = np.argsort(deg, kind="stable")
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    r = rank[both[:, 0]]
    c = rank[both[:, 1]]
    keep = c > r
    r, c = r[keep], c[keep]
    idx = np.lexsort((c, r))
    r, c = r[idx], c[idx]
    rowptr = np.zeros(n + 1, np.uint64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.uint64)
    return rowptr, c.astype(np.uint32)


def wedges(rowptr):
    d = np.diff(rowptr.astype(np.int64))
    return int(np.sum(d * (d - 1) // 2))


def csr_hash(rowptr, colind):
    h = hashlib.sha256()
    h.update(rowptr.astype("<u8").tobytes())
    h.update(colind.astype("<u4").tobytes())
    return h.hexdigest()


def triangles(G):
    G = nx.Graph(G)
    G.remove_edges_from(nx.selfloop_edges(G))
    t = sum(nx.triangles(G).values()) // 3
    if G.number_of_nodes():
        A = nx.to_scipy_sparse_array(G, format="csr", dtype=np.int64, weight=None)
        A = sp.csr_matrix(A)
        t2 = int((A @ A).multiply(A).sum()) // 6
        assert t == t2, (t, t2)
    return t


def write_fixture(name, G, fmt, one_indexed=True, header=None, extra_lines=None):
    """G: networkx graph on ids 0..N-1.  The file uses ids+1 unless zero-indexed."""
    und = [(u, v) for u, v in G.edges() if u != v]
    shift = 1 if one_indexed else 0
    lines = []
    if fmt == "mm":
        lines.append("%%MatrixMarket matrix coordinate pattern symmetric")
        lines.append("% generated fixture")
        nn = max(G.nodes()) + 1 if G.number_of_nodes() else 0
        lines.append(f"{nn} {nn} {len(und)}")
    else:
        lines.append(f"# {name}")
        lines.append("# FromNodeId\tToNodeId")
    for u, v in und:
        lines.append(f"{u + shift}\t{v + shift}" if fmt == "el" else f"{u + shift} {v + shift}")
    if extra_lines:
        lines.extend(extra_lines)
    path = os.path.join(HERE, f"{name}.{'mtx' if fmt == 'mm' else 'txt'}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    ids = [x for e in und for x in e]
    n = max(ids) + 1 if ids else 1
    rp, ci = oriented_csr(n, und)
    return {
        "file": os.path.basename(path), "format": fmt, "zero_indexed": not one_indexed,
        "n": n, "m": int(len(ci)), "W": wedges(rp), "triangles": triangles(G),
        "csr_sha256": csr_hash(rp, ci),
    }


def main():
    fx = []
    rng = np.random.default_rng(20240601)
    fx.append(write_fixture("karate", nx.karate_club_graph(), "el"))
    fx.append(write_fixture("les_miserables", nx.convert_node_labels_to_integers(nx.les_miserables_graph()), "mm"))
    fx.append(write_fixture("florentine", nx.convert_node_labels_to_integers(nx.florentine_families_graph()), "el"))
    fx.append(write_fixture("davis_bipartite", nx.convert_node_labels_to_integers(nx.davis_southern_women_graph()),
                            "el"))
    fx.append(write_fixture("gnp_300", nx.gnp_random_graph(300, 0.06, seed=11), "el", one_indexed=False))
    fx.append(write_fixture("ba_800", nx.barabasi_albert_graph(800, 6, seed=12), "mm"))
    fx.append(write_fixture("plc_1000", nx.powerlaw_cluster_graph(1000, 5, 0.6, seed=13), "el"))
    fx.append(write_fixture("ws_600", nx.watts_strogatz_graph(600, 10, 0.1, seed=14), "el"))
    fx.append(write_fixture("complete_60", nx.complete_graph(60), "el"))
    fx.append(write_fixture("complete_bipartite_20_30", nx.complete_bipartite_graph(20, 30), "mm"))
    fx.append(write_fixture("grid_30x30", nx.convert_node_labels_to_integers(nx.grid_2d_graph(30, 30)), "el"))
    fx.append(write_fixture("petersen", nx.petersen_graph(), "el"))
    fx.append(write_fixture("wheel_50", nx.wheel_graph(50), "el"))
    # scale-free with hubs and a long tail of degree-1 vertices
    fx.append(write_fixture("sf_2000", nx.Graph(nx.scale_free_graph(2000, seed=15).to_undirected()), "el"))
    # gaps in the id space: n = max id + 1 with isolated ids
    G = nx.Graph()
    G.add_edges_from([(0, 5), (5, 9), (9, 0), (9, 40), (40, 5), (2, 3)])
    fx.append(write_fixture("gapped_ids", G, "el"))
    # duplicates, reversed duplicates and self-loops in the file
    G = nx.gnp_random_graph(120, 0.1, seed=16)
    e = list(G.edges())
    extra = [f"{u + 1}\t{v + 1}" for u, v in e[::7]] + [f"{v + 1}\t{u + 1}" for u, v in e[::5]]
    extra += [f"{u + 1}\t{u + 1}" for u in range(0, 120, 9)]
    fx.append(write_fixture("dups_loops", G, "el", extra_lines=extra))
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(fx, f, indent=1)
    print(f"wrote {len(fx)} fixtures")
    del rng


if __name__ == "__main__":
    main()
