"""Benchmark: exact triangle counting of a seeded R-MAT scale-24 graph on one GPU.

A step is one full device count of the resident oriented CSR: wedge-prefix
scan, search tree, top-a bitmap, the wedge-parallel counting kernel and the
8-byte count back on the host.  value = wedge checks/s = W / step time.

Prints one JSON line (rank 0, N=1).  --gpus N>1 is not supported: this build
runs on one GPU.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "wedge-parallel-triangle-counting_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import wtc  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_name(workload, spread, a):
    return f"k_count_{workload}_s{spread}_a{a}_traffic.json"


def load_traffic(workload, spread, a):
    """HBM bytes per counting-kernel launch from the committed rocprofv3 PMC
    summary, if one exists for this workload and configuration."""
    path = os.path.join(ROOT, "profiles", traffic_name(workload, spread, a))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(g, dg, spread, a, target_s, threads):
    """Oracle merge-intersection over a contiguous vertex sample, timed on the
    host; the GPU counts the same sample's wedge range for a parity check."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    d = np.diff(g.row_ptr.astype(np.int64))
    ws = np.concatenate([[0], np.cumsum(d * (d - 1) // 2)]).astype(np.uint64)
    W = int(ws[-1])
    v0 = g.n // 2
    # calibrate on a small sample, then size the sample to ~target_s seconds
    frac = 1e-4
    for _ in range(6):
        v1 = int(np.searchsorted(ws, ws[v0] + int(W * frac), side="right"))
        v1 = min(max(v1, v0 + 1), g.n)
        t0 = time.perf_counter()
        O.tc_merge_range(g, v0, v1, threads)
        dt = time.perf_counter() - t0
        if dt > 0.5 or v1 >= g.n:
            break
        frac *= 8
    frac = min(frac * target_s / max(dt, 1e-3), 1.0)
    v1 = int(np.searchsorted(ws, ws[v0] + int(W * frac), side="right"))
    v1 = min(max(v1, v0 + 1), g.n)
    sample_w = int(ws[v1] - ws[v0])
    t0 = time.perf_counter()
    cpu_count = O.tc_merge_range(g, v0, v1, threads)
    dt = time.perf_counter() - t0
    gpu_count = dg.count(spread, a, int(ws[v0]), int(ws[v1]))
    # the same oracle on one thread, over the first eighth of the sample's rows
    v1s = int(np.searchsorted(ws, ws[v0] + sample_w // 8, side="right"))
    v1s = min(max(v1s, v0 + 1), v1)
    t0 = time.perf_counter()
    O.tc_merge_range(g, v0, v1s, 1)
    dt1 = time.perf_counter() - t0
    return {
        "value": sample_w / dt, "unit": "wedge checks/s", "cores": threads, "kind": "port",
        "value_1thread": int(ws[v1s] - ws[v0]) / dt1,
        "sample": f"oracle merge-intersection over rows [{v0},{v1}) = {sample_w} of {W} wedges "
                  f"({100.0 * sample_w / W:.2f}%), {dt:.1f} s on {threads} host threads",
        "seconds": dt, "sample_triangles": cpu_count, "gpu_sample_triangles": gpu_count,
        "sample_parity": cpu_count == gpu_count,
    }


def batch_graphs(k, base_seed):
    """The batching workload: k seeded graphs, each Erdos-Renyi or
    Barabasi-Albert (50/50), n log-uniform in [64, 65536], average degree
    uniform in [4, 64]; returned oriented (degree order)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        r = np.random.default_rng(base_seed + i)
        er = r.random() < 0.5
        n = int(round(np.exp(r.uniform(np.log(64), np.log(65536)))))
        kdeg = r.uniform(4, 64)
        if er:
            g = wtc.gen_erdos_renyi(n, min(1.0, kdeg / (n - 1)), seed=base_seed + i)
        else:
            g = wtc.gen_barabasi_albert(n, max(1, min(n - 1, int(np.ceil(kdeg / 2)))), seed=base_seed + i)
        return wtc.preprocess(g)

    with ThreadPoolExecutor(max_workers=min(16, len(os.sched_getaffinity(0)))) as ex:
        return list(ex.map(one, range(k)))


def batch_bench(ctx, k, base_seed):
    """Batched vs plain graphs/s on the same host-resident oriented CSRs.
    plain: one reference-style tc_GPU call per graph (alloc, copy, count, copy
    back, free); batched: one upload of all graphs and one persistent counting
    launch.  Both use the autotuner's spread and bitmap length per graph."""
    t0 = time.perf_counter()
    graphs = batch_graphs(k, base_seed)
    tgen = time.perf_counter() - t0
    tune = [wtc.autotune(g.n, g.m, g.wedges(), batch=True) for g in graphs]
    spread = [s for s, _ in tune]
    alen = [a for _, a in tune]
    log(f"batch: {k} graphs, {sum(g.m for g in graphs)} edges, {sum(g.wedges() for g in graphs)} wedges, "
        f"generated in {tgen:.1f}s")
    # warm-up: the plain path on a few graphs; the batched path once at full
    # size, so that its reusable pinned staging buffer and device workspace
    # exist before the timed run (a serving/batching process keeps them)
    for g, s, a in list(zip(graphs, spread, alen))[:4]:
        wtc.tc_gpu_big(g, s, a)
    wtc.batch_count_host(ctx, graphs, spread, alen)
    t0 = time.perf_counter()
    plain = [wtc.tc_gpu_big(g, s, a) for g, s, a in zip(graphs, spread, alen)]
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched, st = wtc.batch_count_host(ctx, graphs, spread, alen, stats=True)
    t_batch = time.perf_counter() - t0
    b = wtc.Batch(ctx, graphs)
    b.count(spread, alen)
    t0 = time.perf_counter()
    resident, st_res = b.count(spread, alen, stats=True)
    t_res = time.perf_counter() - t0
    b.free()
    exact = bool(np.array_equal(np.asarray(plain, np.uint64), batched)) and bool(np.array_equal(batched, resident))
    return {"graphs": k, "edges": int(sum(g.m for g in graphs)), "wedges": int(sum(g.wedges() for g in graphs)),
            "plain_graphs_per_s": k / t_plain, "batched_graphs_per_s": k / t_batch,
            "resident_batched_graphs_per_s": k / t_res, "batched_over_plain": t_plain / t_batch,
            "plain_s": t_plain, "batched_s": t_batch, "resident_batched_s": t_res,
            "batched_count_kernel_ms": st.count_ms, "resident_count_kernel_ms": st_res.count_ms,
            "per_graph_counts_equal": exact}


def small_batch_graphs(k, base_seed, n_max):
    """A batch of small graphs: k seeded graphs, Erdos-Renyi or
    Barabasi-Albert (50/50), n log-uniform in [32, n_max], average degree
    uniform in [4, 32]; returned oriented."""
    def one(i):
        r = np.random.default_rng(base_seed + i)
        n = int(round(np.exp(r.uniform(np.log(32), np.log(n_max)))))
        kdeg = r.uniform(4, 32)
        if i % 2:
            g = wtc.gen_erdos_renyi(n, min(1.0, kdeg / (n - 1)), seed=base_seed + i)
        else:
            g = wtc.gen_barabasi_albert(n, max(1, min(n - 1, int(np.ceil(kdeg / 2)))), seed=base_seed + i)
        return wtc.preprocess(g)
    return [one(i) for i in range(k)]


def small_kernel_bench(ctx, k, base_seed, reps=7):
    """Kernel choice for sets of small graphs: graph-replay time of one
    resident batch counted by the small-graph kernel (CSR in LDS, one block
    per slice of a graph) and by the wedge-parallel kernel, for batches of
    graphs up to n = 512 and n = 1024."""
    default, _ = ctx.small_graphs()
    res = {}
    try:
        for n_max in (512, 1024):
            graphs = small_batch_graphs(k, base_seed, n_max)
            tune = [wtc.autotune(g.n, g.m, g.wedges(), batch=True) for g in graphs]
            spread = [s for s, _ in tune]
            alen = [a for _, a in tune]
            row = {"graphs": k, "n_max": n_max, "wedges": int(sum(g.wedges() for g in graphs))}
            counts = {}
            for name, limit in (("small_kernel", default), ("wedge_kernel", 0)):
                ctx.set_small_graphs(limit)
                b = wtc.Batch(ctx, graphs)
                b.count(spread, alen)  # capture
                ms = []
                for _ in range(reps):
                    got, st = b.count(spread, alen, stats=True)
                    ms.append(st.count_ms)
                b.free()
                counts[name] = got
                row[f"{name}_replay_ms"] = float(np.median(ms))
                row[f"{name}_graphs"] = int(st.small_graphs)
            row["counts_equal"] = bool(np.array_equal(counts["small_kernel"], counts["wedge_kernel"]))
            row["speedup"] = row["wedge_kernel_replay_ms"] / row["small_kernel_replay_ms"]
            res[f"n_max_{n_max}"] = row
    finally:
        ctx.set_small_graphs(default)
    return res


def arrow_bench(scale, seed, spreads=(8, 16, 32), a=131072, iters=3, baselines=False):
    """The arrow-wedge count (experiments/tc_arrow.cu) against the
    wedge-parallel count on the same resident R-MAT graph: per spread, the
    median count-kernel time over `iters` counts after a warm-up, the arrow
    wedges checked and the rate; all counts must agree.  With `baselines`,
    also one timed count of each vertex-/edge-parallel kernel and wedge style
    (experiments/tc_vertex.cu, tc_edge.cu)."""
    ctx = wtc.Context(0)
    dg = wtc.DeviceGraph.gen_rmat(ctx, scale, 16, seed=seed)
    n, m, a_w = dg.n, dg.m, a
    dg.count(12, a_w)
    times, tri, W = [], 0, 0
    for _ in range(iters):
        tri, st = dg.count(12, a_w, stats=True)
        times.append(st.count_ms)
        W = st.wedges
    res = {"workload": f"rmat{scale}", "n": n, "m": m, "triangles": tri,
           "wedge_parallel": {"spread": 12, "a": a_w, "wedges": W, "count_ms": statistics.median(times),
                              "checks_per_s": W / statistics.median(times) * 1e3}, "arrow": []}
    for s in spreads:
        dg.count_arrow(s, a)
        times = []
        for _ in range(iters):
            t, A, st = dg.count_arrow(s, a, stats=True)
            if t != tri:
                raise SystemExit(f"arrow count {t} != wedge-parallel count {tri} (s={s}, a={a})")
            times.append(st.count_ms)
        ms = statistics.median(times)
        res["arrow"].append({"spread": s, "a": a, "arrow_wedges": A, "count_ms": ms, "prefix_ms": st.prefix_ms,
                             "checks_per_s": A / ms * 1e3})
        log(f"arrow s={s} a={a}: {A} arrow wedges, count {ms:.1f} ms")
    if baselines:  # experiments/tc_vertex.cu and tc_edge.cu, every wedge style
        res["baselines"] = []
        for par, pname in ((wtc.PARALLEL_VERTEX, "vertex"), (wtc.PARALLEL_EDGE, "edge")):
            for style, sname in ((wtc.WEDGE_OUTGOING, "outgoing"), (wtc.WEDGE_ARROW, "arrow"),
                                 (wtc.WEDGE_MIXED, "mixed")):
                t, st = dg.count_baseline(par, style, a, stats=True)
                if t != tri:
                    raise SystemExit(f"{pname}/{sname} count {t} != wedge-parallel count {tri}")
                res["baselines"].append({"parallelism": pname, "style": sname, "a": a, "count_ms": st.count_ms})
                log(f"{pname}-parallel {sname}: count {st.count_ms:.1f} ms")
    dg.free()
    ctx.close()
    return res


def reference_configs(iters=10):
    """The reference's own measurement cases through the tc_GPU drop-in (per
    call: allocate, copy, count, copy back, free): the size of its usage example
    (Chung-Lu n = 2^18, ~900k edges, s = 5, a = 8192) and its (s, a) sweep
    (R-MAT-20, s in {1,3,5,8} x a in {0,4096,8192,16384}).  Per (s, a): the
    median over `iters` calls of GPU_time's copy and exec (ms, tc.cu:76-79)
    and of the call's wall time, and the device time of the same count on a
    resident copy of the graph."""
    ctx = wtc.Context(0)

    def run(g, params):
        rows = []
        dg = ctx.upload(g)
        for s_, a_ in params:
            ref = wtc.tc_gpu(g, s_, a_)  # warm-up
            cp, ex, wall, res = [], [], [], []
            for _ in range(iters):
                t = wtc.GpuTime()
                t0 = time.perf_counter()
                c = wtc.tc_gpu(g, s_, a_, t)
                wall.append((time.perf_counter() - t0) * 1e3)
                assert c == ref
                cp.append(t.copy)
                ex.append(t.exec)
                c2, st = dg.count(s_, a_, stats=True)
                assert c2 == ref
                res.append(st.prefix_ms + st.bitmap_ms + st.count_ms)
            rows.append({"spread": s_, "a": a_, "triangles": ref, "copy_ms": float(np.median(cp)),
                         "exec_ms": float(np.median(ex)), "call_ms": float(np.median(wall)),
                         "resident_ms": float(np.median(res))})
        dg.free()
        return {"n": g.n, "m": g.m, "wedges": g.wedges(), "runs": rows}

    def prep(raw):
        """The preprocessing drop-in (tc.cu:687-787): median of 3 wall times
        for -p 0 (host builder) and -p 2 (GPU builder); both give one CSR."""
        res = {}
        for style in (0, 2):
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                g = wtc.preprocess(raw, style)
                ts.append((time.perf_counter() - t0) * 1e3)
            res[f"prepro_p{style}_ms"] = float(np.median(ts))
        return g, res

    g1, p1 = prep(wtc.gen_chung_lu())
    g2, p2 = prep(wtc.gen_rmat(20, 16, seed=1))
    out = {"chung_lu_2e18": dict(run(g1, [(5, 8192)]), **p1),
           "rmat20_sweep": dict(run(g2, [(s_, a_) for s_ in (1, 3, 5, 8) for a_ in (0, 4096, 8192, 16384)]), **p2)}
    ctx.close()
    return out


def serve_bench(lams, seconds, seeds_per_scale, queue_capacity, ways, chunk, waves):
    """Serving workload: Poisson arrivals of count requests on seeded R-MAT
    graphs (scale uniform in [10, 22], edge factor 16), each with the
    autotuner's spread and bitmap length.  Graphs are registered once and stay
    resident; requests are admitted by HBM footprint and batched continuously.
    Per arrival rate: p50/p99 latency, completed graphs/s, rejected requests,
    and exactness against each graph's count served alone.  The plain path on
    the same mix is one reference-style tc_GPU call per request (alloc, copy,
    count, copy back, free)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    scales = list(range(10, 23))
    keys = [(sc, k) for sc in scales for k in range(seeds_per_scale)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        graphs = list(ex.map(lambda key: wtc.preprocess(wtc.gen_rmat(key[0], 16, seed=7000 + 31 * key[0] + key[1])),
                             keys))
    tune = [wtc.autotune(g.n, g.m, g.wedges()) for g in graphs]
    # plain path on the same mix: mean time of one reference-style call per graph
    plain_ms = []
    for g, (s_g, a_g) in zip(graphs, tune):
        wtc.tc_gpu_big(g, s_g, a_g)
        t0 = time.perf_counter()
        for _ in range(2):
            wtc.tc_gpu_big(g, s_g, a_g)
        plain_ms.append((time.perf_counter() - t0) / 2 * 1e3)
    plain_gps = 1e3 / float(np.mean(plain_ms))
    log(f"serve: plain path {plain_gps:.1f} graphs/s on the uniform mix")
    srv = wtc.Server(0, hbm_budget_bytes=32 << 30, max_batch=256, slots=2, queue_capacity=queue_capacity,
                     probe_ways=ways, lds_chunk=chunk, waves_per_block=waves)
    gids = [srv.register(g) for g in graphs]
    for gid, (s_g, a_g) in zip(gids, tune):  # each graph alone once: resident, server warmed up
        srv.wait(srv.submit(gid, s_g, a_g))
    alone, alone_ms = [], []
    for gid, (s_g, a_g) in zip(gids, tune):  # then timed, each graph alone
        cnt, st = srv.wait(srv.submit(gid, s_g, a_g))
        alone.append(cnt)
        alone_ms.append(st.service_ms)
    log(f"serve: {len(graphs)} graphs resident, alone service ms by scale: "
        + ", ".join(f"{sc}:{alone_ms[i]:.2f}" for i, (sc, k) in enumerate(keys) if k == 0))
    results = []
    for lam in lams:
        rng = np.random.default_rng(int(lam))
        n_req = int(lam * seconds)
        gaps = rng.exponential(1.0 / lam, size=n_req)
        arrive = np.cumsum(gaps)
        pick = [rng.integers(0, len(scales)) * seeds_per_scale + rng.integers(0, seeds_per_scale)
                for _ in range(n_req)]
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
                    rids[i] = srv.submit(gids[pick[i]], *tune[pick[i]])
                except wtc.WtcError:
                    rids[i] = -1  # rejected: queue full

        th = threading.Thread(target=submitter)
        th.start()
        lat, exact, done = [], True, 0
        t_last = t0
        for i in range(n_req):
            while rids[i] is None:
                time.sleep(0.0002)
            if rids[i] == -1:
                continue
            cnt, st = srv.wait(rids[i])
            t_last = time.perf_counter()
            exact &= cnt == alone[pick[i]]
            lat.append(st.latency_ms)
            done += 1
        th.join()
        span = max(t_last - t0, 1e-9)
        lat = np.asarray(lat) if lat else np.zeros(1)
        rec = {"lambda": lam, "requests": n_req, "completed": done, "rejected": n_req - done,
               "graphs_per_s": done / span, "p50_ms": float(np.percentile(lat, 50)),
               "p99_ms": float(np.percentile(lat, 99)), "exact": bool(exact)}
        log(f"serve lambda={lam}: {rec}")
        results.append(rec)
    stats = srv.stats()
    srv.close()
    best = max(r["graphs_per_s"] for r in results) if results else 0.0
    return {"graphs": len(graphs), "scales": [scales[0], scales[-1]], "seconds_per_rate": seconds,
            "queue_capacity": queue_capacity, "params": "autotuned per graph", "rates": results,
            "plain_graphs_per_s": plain_gps, "served_max_graphs_per_s": best,
            "served_over_plain": best / plain_gps, "server": stats}


def ingest_bench(scale, seed):
    """Ingestion: an R-MAT edge list as text -> read_graph (the parallel
    reader, with every host thread and with one) -> preprocess -> binary CSR
    file written and read back.  Host only."""
    import subprocess
    import tempfile

    import pandas as pd

    pairs = wtc.gen_rmat_pairs(scale, 16, seed=seed)
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        txt = os.path.join(d, "g.txt")
        with open(txt, "w") as f:
            f.write("% R-MAT edge list\n")
        pd.DataFrame(pairs).to_csv(txt, sep=" ", header=False, index=False, mode="a")
        size = os.path.getsize(txt)
        threads = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        g = wtc.read_graph(txt, zero_indexed=True)
        t_read = time.perf_counter() - t0
        code = ("import sys, time; sys.path.insert(0, %r); import wtc; t = time.perf_counter(); "
                "g = wtc.read_graph(%r, zero_indexed=True); print(time.perf_counter() - t, g.m)" % (PKG, txt))
        r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, WTC_HOST_THREADS="1"),
                           capture_output=True, text=True, check=True)
        t_read1, m1 = r.stdout.split()
        assert int(m1) == g.m
        t0 = time.perf_counter()
        og = wtc.preprocess(g)
        t_pre = time.perf_counter() - t0
        csr = os.path.join(d, "g.csr")
        t0 = time.perf_counter()
        wtc.write_csr(csr, og, tag=1)
        t_w = time.perf_counter() - t0
        t0 = time.perf_counter()
        back, tag = wtc.read_csr(csr)
        t_r = time.perf_counter() - t0
        same = tag == 1 and back.m == og.m and np.array_equal(back.col_ind, og.col_ind) and \
            np.array_equal(back.row_ptr, og.row_ptr)
        csr_size = os.path.getsize(csr)
    return {"workload": f"rmat{scale}-ef16 edge list", "text_bytes": size, "pairs": int(len(pairs)),
            "host_threads": threads, "read_graph_s": t_read, "read_graph_1thread_s": float(t_read1),
            "read_speedup": float(t_read1) / t_read, "text_MB_per_s": size / t_read / 1e6,
            "preprocess_s": t_pre, "csr_bytes": csr_size, "write_csr_s": t_w, "read_csr_s": t_r,
            "read_csr_GB_per_s": csr_size / t_r / 1e9, "csr_round_trip_equal": bool(same),
            "text_to_oriented_s": t_read + t_pre}


def _rocprof_rows(directory, pattern):
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(directory, "**", pattern), recursive=True):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    return rows


def _per_dispatch(directory, kernel_sub, counter):
    """Mean per-dispatch value of `counter` for kernels whose name contains
    kernel_sub, from rocprofv3 counter_collection.csv files."""
    vals = {}
    for r in _rocprof_rows(directory, "*counter_collection.csv"):
        if kernel_sub in r.get("Kernel_Name", "") and r.get("Counter_Name") == counter:
            key = r.get("Dispatch_Id") or r.get("Correlation_Id")
            vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
    if not vals:
        raise SystemExit(f"no {counter} rows for {kernel_sub} under {directory}")
    return sum(vals.values()) / len(vals), len(vals)


def pmc_summary(args):
    """Turns four rocprofv3 --pmc runs (FETCH_SIZE and WRITE_SIZE, each on the
    counting kernel and on the calibration copy) into HBM bytes per counting
    launch, each counter scaled by what it reads for the copy of known size."""
    kdir_f, kdir_w, cdir_f, cdir_w = args.pmc_summary
    copy_bytes = args.copy_calibration_mb * 2 ** 20
    kf, nk = _per_dispatch(kdir_f, "k_count", "FETCH_SIZE")
    kw, _ = _per_dispatch(kdir_w, "k_count", "WRITE_SIZE")
    cf, _ = _per_dispatch(cdir_f, "k_copy", "FETCH_SIZE")
    cw, _ = _per_dispatch(cdir_w, "k_copy", "WRITE_SIZE")
    f_scale = copy_bytes / (cf * 1024.0)  # true bytes per counted byte
    w_scale = copy_bytes / (cw * 1024.0)
    fetch = kf * 1024.0 * f_scale
    write = kw * 1024.0 * w_scale
    out = {"workload": args.pmc_workload, "kernel": "k_count", "dispatches": nk,
           "fetch_kib_per_dispatch": kf, "write_kib_per_dispatch": kw,
           "copy_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
           "fetch_scale": f_scale, "write_scale": w_scale,
           "fetch_bytes": fetch, "write_bytes": write, "bytes_per_launch": fetch + write}
    out["config"] = {"spread": args.spread, "a": args.a}
    path = os.path.join(ROOT, "profiles", traffic_name(args.pmc_workload, args.spread, args.a))
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--spread", type=int, default=0, help="0: the autotuner's choice")
    ap.add_argument("--a", type=int, default=-1, help="-1: the autotuner's choice")
    ap.add_argument("--ways", type=int, default=4, help="probes advanced together per lane")
    ap.add_argument("--chunk", type=int, default=12, help="wedges per lane staged in LDS per phase")
    ap.add_argument("--waves", type=int, default=4, help="waves per counting block (1, 2, 4 or 8)")
    ap.add_argument("--builder", choices=("auto", "host", "device"), default="auto",
                    help="build the oriented CSR on the host or on the GPU (auto: GPU from scale 26)")
    ap.add_argument("--offset-bytes", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-range-check", action="store_true",
                    help="skip the four partial-range counts (keeps profiler runs to whole-graph launches)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run the calibration copy of this many MiB (for profiler counters)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("KFETCH", "KWRITE", "CFETCH", "CWRITE"),
                    help="summarise rocprofv3 --pmc output dirs into profiles/ (no GPU use)")
    ap.add_argument("--pmc-workload", default="rmat24")
    ap.add_argument("--batch-graphs", type=int, default=4096, help="batched vs plain on this many graphs (0: skip)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the RMAT count: no batched-vs-plain and no serving legs (profiler runs)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched vs plain comparison")
    ap.add_argument("--reference-configs", action="store_true",
                    help="run only the reference's Chung-Lu 2^18 case and R-MAT-20 (s, a) sweep through the "
                         "tc_GPU drop-in")
    ap.add_argument("--arrow-only", action="store_true",
                    help="run only the arrow-wedge count against the wedge-parallel count (R-MAT --scale)")
    ap.add_argument("--baselines", action="store_true",
                    help="with --arrow-only: also time the vertex-/edge-parallel baselines")
    ap.add_argument("--small-batches-only", action="store_true",
                    help="run only the small-graph kernel vs wedge kernel comparison on batches of small graphs")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving (Poisson arrivals) workload")
    ap.add_argument("--serve-rates", default="100,1000",
                    help="Poisson arrival rates of the serving leg (empty: skip it)")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--ingest-only", action="store_true", help="run only the text/CSR-file ingestion timing")
    args = ap.parse_args()

    if args.pmc_summary:
        pmc_summary(args)
        return
    if args.ingest_only:
        print(json.dumps(ingest_bench(args.scale, args.seed)), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if wtc.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration_mb:
        ctx = wtc.Context(0)
        nbytes = args.copy_calibration_mb * 2 ** 20
        ms = ctx.debug_copy(nbytes, reps=1)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms, "GB_per_s": 2 * nbytes / ms / 1e6}), flush=True)
        ctx.close()
        return

    if args.serve_only:
        res = serve_bench([float(x) for x in args.serve_rates.split(",")], args.serve_seconds, 2, 512,
                          args.ways, args.chunk, args.waves)
        print(json.dumps(res), flush=True)
        return
    if args.reference_configs:
        print(json.dumps(reference_configs()), flush=True)
        return
    if args.arrow_only:
        print(json.dumps(arrow_bench(args.scale, args.seed, baselines=args.baselines)), flush=True)
        return
    if args.small_batches_only:
        ctx = wtc.Context(0)
        ctx.set_tuning(args.ways, args.chunk, args.waves)
        res = small_kernel_bench(ctx, args.batch_graphs or 4096, 7000)
        ctx.close()
        print(json.dumps(res), flush=True)
        return
    if args.batch_only:
        ctx = wtc.Context(0)
        ctx.set_tuning(args.ways, args.chunk, args.waves)
        res = batch_bench(ctx, args.batch_graphs or 4096, 1000)
        res["small_graph_batches"] = small_kernel_bench(ctx, args.batch_graphs or 4096, 7000)
        ctx.close()
        print(json.dumps(res), flush=True)
        return

    workload = f"rmat{args.scale}"
    ctx = wtc.Context(0)
    ctx.set_tuning(args.ways, args.chunk, args.waves)
    device_build = args.builder == "device" or (args.builder == "auto" and args.scale >= 26)
    t0 = time.perf_counter()
    if device_build:  # generate, sort, deduplicate and orient on the GPU; keep it resident
        dg = wtc.DeviceGraph.gen_rmat(ctx, args.scale, args.edge_factor, seed=args.seed,
                                      offset_bytes=args.offset_bytes)
        t1 = t2 = time.perf_counter()
        log(f"{workload}: built on the GPU in {t2 - t0:.1f}s ({dg.nbytes / 1e9:.2f} GB resident)")
        g = dg.download()  # host copy for W, B_alg and the CPU baseline
        ob = dg.offset_bytes
    else:
        sym = wtc.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
        t1 = time.perf_counter()
        g = wtc.preprocess(sym)
        t2 = time.perf_counter()
        del sym
        ob = args.offset_bytes or (4 if g.m < 2 ** 32 - 1 else 8)
        dg = ctx.upload(g, offset_bytes=ob)
        log(f"uploaded {dg.nbytes / 1e9:.2f} GB")
    W = g.wedges()
    log(f"{workload}: n={g.n} m={g.m} W={W} gen {t1 - t0:.1f}s prepro {t2 - t1:.1f}s")
    b_alg, b_min = g.alg_bytes(ob)
    auto_s, auto_a = wtc.autotune(g.n, g.m, W)
    args.spread = args.spread or auto_s
    args.a = auto_a if args.a < 0 else args.a
    log(f"spread={args.spread} a={args.a} ways={args.ways} chunk={args.chunk} waves={args.waves}")
    for _ in range(args.warmup):
        dg.count(args.spread, args.a)
    counts, stats, wall = [], [], []
    tstart = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        c, st = dg.count(args.spread, args.a, stats=True)
        wall.append(time.perf_counter() - s0)
        counts.append(c)
        stats.append(st.as_dict())
    elapsed = time.perf_counter() - tstart
    assert len(set(counts)) == 1, counts
    log(f"triangles={counts[0]} step={1e3 * elapsed / args.steps:.2f} ms")

    ms_per_step = 1e3 * elapsed / args.steps
    count_ms = statistics.median(s["count_ms"] for s in stats)
    achieved = b_alg / (count_ms / 1e3)
    traffic = load_traffic(workload, args.spread, args.a)
    out = {
        "metric": "wedge checks/s & edges/s at RMAT-24 (1 GPU, exact); batched vs plain graphs/s",
        "value": W / (ms_per_step / 1e3),
        "unit": "wedge checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"u32 ids, u{8 * ob} offsets, u64 counts (integer)",
        "data": "synthetic: seeded R-MAT (a,b,c,d=.57,.19,.19,.05), symmetrised, deduplicated, degree-oriented",
        "config": {"workload": f"rmat{args.scale}-ef{args.edge_factor}", "scale": args.scale,
                   "edge_factor": args.edge_factor, "seed": args.seed, "n": g.n, "m": g.m, "wedges": W,
                   "spread": args.spread, "a": args.a, "offset_bytes": ob},
        "edges_per_s": g.m / (ms_per_step / 1e3),
        "triangles": counts[0],
        "kernel_ms": {"count": count_ms,
                      "prefix": statistics.median(s["prefix_ms"] for s in stats),
                      "bitmap": statistics.median(s["bitmap_ms"] for s in stats),
                      "grid": stats[0]["blocks"]},
        "wedge_checks_per_s_kernel": W / (count_ms / 1e3),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "b_alg": b_alg, "b_min": b_min,
                     # measured HBM bytes per launch over the same launch time: the rate the memory
                     # actually ran at (achieved/frac count every probe as HBM bytes, most hit caches)
                     "traffic_GBps": traffic["bytes_per_launch"] / (count_ms / 1e3) / 1e9 if traffic else None,
                     "traffic_frac": traffic["bytes_per_launch"] / (count_ms / 1e3) / HBM_PEAK if traffic else None},
        "prepro_s": t2 - t1,
        "builder": "device" if device_build else "host",
        "tuning": {"probe_ways": args.ways, "lds_chunk": args.chunk, "waves_per_block": args.waves},
    }
    # size-independent check at full size: four wedge ranges sum to the whole count
    if not args.no_range_check:
        cuts = [0, W // 5, W // 2, (4 * W) // 5, W]
        parts = [dg.count(args.spread, args.a, cuts[i], cuts[i + 1]) for i in range(4)]
        out["range_split"] = {"cuts": cuts, "parts": parts, "sum_equals_total": sum(parts) == counts[0]}
        log(f"range split: {parts} sum ok = {sum(parts) == counts[0]}")
    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        out["cpu_baseline"] = cpu_baseline(g, dg, args.spread, args.a, args.cpu_seconds, threads)
        log(f"cpu baseline: {out['cpu_baseline']['sample']}")
    dg.free()
    if args.batch_graphs and not args.no_extras:
        out["batched_vs_plain"] = batch_bench(ctx, args.batch_graphs, 1000)
        log(f"batched/plain = {out['batched_vs_plain']['batched_over_plain']:.2f}x")
        out["small_graph_batches"] = small_kernel_bench(ctx, args.batch_graphs, 7000)
        log("small-graph kernel speedup: " + ", ".join(
            f"{k} {v['speedup']:.2f}x" for k, v in out["small_graph_batches"].items()))
    ctx.close()
    if args.serve_rates and not args.no_extras:
        out["serving"] = serve_bench([float(x) for x in args.serve_rates.split(",")], args.serve_seconds, 2, 512,
                                     args.ways, args.chunk, args.waves)
        log(f"served/plain = {out['serving']['served_over_plain']:.2f}x")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
