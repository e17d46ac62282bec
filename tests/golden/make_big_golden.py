"""Writes the full-size golden counts of the measured workloads, computed with
the CPU oracle only (oracle/: graph_from_pairs -> preprocess -> tc_merge), so
that the GPU tests and the benchmark compare the HIP path with an independent
whole-graph count at the sizes they run:

  big_golden.json   R-MAT-24 (edge factor 16, seed 1): n, m, W, triangles;
                    K_3000 (closed form, W = C(3000,3) > 2^32); and the
                    26 graphs of the serving mix (R-MAT scales 10..22)
  batch4096.npz     the 4,096 Erdos-Renyi / Barabasi-Albert graphs of the
                    batching workload: n, m, W and triangles per graph

The generators (seeded, in the library) only provide the input: R-MAT graphs
enter the oracle as raw (src, dst) pairs, the ER/BA graphs as the symmetric
CSR that read_graph would build; the oracle does the preprocessing (degree
order, orientation) and the merge-intersection count itself.

Run from the repository root: python tests/golden/make_big_golden.py
(about 15 minutes on 8 cores).
"""
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "wedge-parallel-triangle-counting_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
import wtc  # noqa: E402
import workloads  # noqa: E402


def wedges(g):
    d = np.diff(g.row_ptr.astype(np.int64))
    return int(np.sum(d * (d - 1) // 2))


def rmat_entry(scale, seed, threads=0):
    t0 = time.perf_counter()
    pairs = wtc.gen_rmat_pairs(scale, 16, seed=seed)
    g = O.preprocess(O.graph_from_pairs(pairs))
    del pairs
    t1 = time.perf_counter()
    tri = O.tc_merge(g, threads)
    t2 = time.perf_counter()
    return {"scale": scale, "edge_factor": 16, "seed": seed, "n": g.n, "m": g.m, "W": wedges(g),
            "triangles": tri, "build_s": round(t1 - t0, 1), "count_s": round(t2 - t1, 1)}


def main():
    out = {}
    k = 3000
    out["k3000"] = {"n": k, "m": k * (k - 1) // 2, "W": k * (k - 1) * (k - 2) // 6,
                    "triangles": k * (k - 1) * (k - 2) // 6, "source": "closed form C(3000,3)"}
    serve = []
    for sc, seed in workloads.serve_keys():
        e = rmat_entry(sc, seed)
        print(f"serve rmat{sc} seed {seed}: {e}", flush=True)
        serve.append(e)
    out["serve_mix"] = serve
    e = rmat_entry(24, 1)
    print(f"rmat24: {e}", flush=True)
    out["rmat24"] = e
    with open(os.path.join(HERE, "big_golden.json"), "w") as f:
        json.dump(out, f, indent=1)

    kb = 4096

    def one(i):
        g = O.preprocess(workloads.batch_symmetric(i))
        return g.n, g.m, wedges(g), O.tc_merge(g, 1)

    t0 = time.perf_counter()
    workers = min(8, os.cpu_count() or 8)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        rows = list(ex.map(one, range(kb)))
    arr = np.asarray(rows, dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "batch4096.npz"), n=arr[:, 0], m=arr[:, 1], W=arr[:, 2],
                        triangles=arr[:, 3], base_seed=np.uint64(workloads.BATCH_BASE_SEED))
    print(f"batch4096: {kb} graphs, W {int(arr[:, 2].sum())}, triangles {int(arr[:, 3].sum())}, "
          f"{time.perf_counter() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
