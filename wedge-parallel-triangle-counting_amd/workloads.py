"""Seeded synthetic workloads of the measured configurations, shared by the
benchmark and the parity tests so that both count exactly the same graphs.

  rmat(scale)        R-MAT (a,b,c,d = .57,.19,.19,.05), edge factor 16, seed 1,
                     symmetrised, deduplicated, self-loops dropped, oriented
                     by degree (the reference's preprocess, tc.cu:687-787)
  batch_graphs(k)    the batching workload: k graphs, each Erdos-Renyi or
                     Barabasi-Albert (50/50), n log-uniform in [64, 65536],
                     average degree uniform in [4, 64]
  serve_keys()       the serving mix: R-MAT scales 10..22, edge factor 16,
                     two seeds per scale

Every generator is deterministic for any thread count.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import wtc

SERVE_SCALES = tuple(range(10, 23))
SERVE_SEEDS_PER_SCALE = 2
BATCH_BASE_SEED = 1000


def _workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 4
    return max(1, min(16, n))


def batch_params(i, base_seed=BATCH_BASE_SEED):
    """(family, n, generator argument) of graph i of the batching workload:
    family 'er' with edge probability p, or 'ba' with m0 attachments."""
    r = np.random.default_rng(base_seed + i)
    er = r.random() < 0.5
    n = int(round(np.exp(r.uniform(np.log(64), np.log(65536)))))
    kdeg = r.uniform(4, 64)
    if er:
        return "er", n, min(1.0, kdeg / (n - 1))
    return "ba", n, max(1, min(n - 1, int(np.ceil(kdeg / 2))))


def batch_symmetric(i, base_seed=BATCH_BASE_SEED):
    """Graph i of the batching workload as the symmetric CSR read_graph would
    build (not yet oriented)."""
    fam, n, x = batch_params(i, base_seed)
    if fam == "er":
        return wtc.gen_erdos_renyi(n, x, seed=base_seed + i)
    return wtc.gen_barabasi_albert(n, x, seed=base_seed + i)


def batch_graphs(k, base_seed=BATCH_BASE_SEED):
    """The k oriented graphs of the batching workload."""
    with ThreadPoolExecutor(max_workers=_workers()) as ex:
        return list(ex.map(lambda i: wtc.preprocess(batch_symmetric(i, base_seed)), range(k)))


def serve_keys():
    """(scale, seed) of every graph of the serving mix, in registration order."""
    return [(sc, 7000 + 31 * sc + k) for sc in SERVE_SCALES for k in range(SERVE_SEEDS_PER_SCALE)]


def serve_graphs():
    """The oriented graphs of the serving mix, in serve_keys() order."""
    with ThreadPoolExecutor(max_workers=4) as ex:
        return list(ex.map(lambda key: wtc.preprocess(wtc.gen_rmat(key[0], 16, seed=key[1])), serve_keys()))


def rmat(scale, seed=1, edge_factor=16):
    """The oriented R-MAT graph of the given scale (host builders)."""
    return wtc.preprocess(wtc.gen_rmat(scale, edge_factor, seed=seed))


def poisson_run(srv, gids, params, lam, seconds, seed):
    """Poisson arrivals at rate lam (requests/s) for `seconds` on a server:
    request i names graph pick[i] (uniform over gids) with that graph's
    (spread, a) from params.  A submitter thread sends each request at its
    arrival time; a request refused because the queue is full is counted as
    rejected.  Returns (pick, results, span_s): results[i] is None for a
    rejected request, else (triangles, latency_ms)."""
    import threading
    import time

    rng = np.random.default_rng(seed)
    n_req = max(1, int(lam * seconds))
    arrive = np.cumsum(rng.exponential(1.0 / lam, size=n_req))
    pick = rng.integers(0, len(gids), size=n_req)
    rids = [None] * n_req
    t0 = time.perf_counter()

    def submitter():
        for i in range(n_req):
            while True:
                left = arrive[i] - (time.perf_counter() - t0)
                if left <= 0:
                    break
                time.sleep(left - 2e-4 if left > 1e-3 else 0)
            try:
                rids[i] = srv.submit(gids[pick[i]], *params[pick[i]])
            except wtc.WtcError:
                rids[i] = -1  # rejected: queue full

    th = threading.Thread(target=submitter)
    th.start()
    results = [None] * n_req
    t_last = t0
    for i in range(n_req):
        while rids[i] is None:
            time.sleep(0.0002)
        if rids[i] == -1:
            continue
        cnt, st = srv.wait(rids[i])
        t_last = time.perf_counter()
        results[i] = (cnt, st.latency_ms)
    th.join()
    return pick, results, max(t_last - t0, 1e-9)
