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


def split_id_bytes(rows):
    """Bytes the split count's row kernel reads per tail id: its packed copy
    of colInd holds P ids per 16-byte chunk, P = 8, 7 or 6 for graphs of at
    most 2^16, 2^18 or 2^21 top rows (the library's packing layouts)."""
    return 16.0 / (8 if rows <= 1 << 16 else 7 if rows <= 1 << 18 else 6)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_name(workload, spread, a, kernel="k_walk"):
    # the split count's kernels do not depend on the spread
    if kernel.startswith("k_split"):
        return f"{kernel}_{workload}_a{a}_traffic.json"
    return f"{kernel}_{workload}_s{spread}_a{a}_traffic.json"


def load_traffic(workload, spread, a, kernel="k_walk"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary, if one exists for this workload and configuration."""
    path = os.path.join(ROOT, "profiles", traffic_name(workload, spread, a, kernel))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count() or 0


def stratified_rows(ws, frac, strata):
    """`strata` runs of rows spread evenly over the wedge order, each starting
    at the first row at or after k/strata of the wedges and covering about
    frac/strata of them: a sample whose per-wedge cost follows the whole graph
    (hub rows at the front, low-degree rows at the back)."""
    W = int(ws[-1])
    n = len(ws) - 1
    runs = []
    for k in range(strata):
        v0 = int(np.searchsorted(ws, (W * k) // strata, side="left"))
        v0 = min(v0, n - 1)
        v1 = int(np.searchsorted(ws, int(ws[v0]) + max(1, int(W * frac / strata)), side="left"))
        v1 = min(max(v1, v0 + 1), n)
        if runs and v0 < runs[-1][1]:
            v0 = runs[-1][1]
        if v0 < v1:
            runs.append((v0, v1))
    return runs


def cpu_baseline(g, dg, spread, a, target_s, threads, strata=32, reps=5):
    """The oracle's merge intersection (the CPU restatement of the count) over
    a stratified sample of the graph's rows, timed on the host: one warm-up
    and the median of `reps` timed runs, each sized to about target_s/(reps+1)
    seconds.  The GPU counts the same rows' wedge ranges for a parity check."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    d = np.diff(g.row_ptr.astype(np.int64))
    ws = np.concatenate([[0], np.cumsum(d * (d - 1) // 2)]).astype(np.uint64)
    W = int(ws[-1])

    def run(runs):
        t0 = time.perf_counter()
        c = sum(O.tc_merge_range(g, v0, v1, threads) for v0, v1 in runs)
        return c, time.perf_counter() - t0

    per_rep = target_s / (reps + 1)
    frac = 2e-4
    for _ in range(8):  # calibrate: grow the sample until one run takes a noticeable time
        runs = stratified_rows(ws, frac, strata)
        _, dt = run(runs)
        if dt > 0.3 or frac >= 1.0:
            break
        frac = min(1.0, frac * 4)
    frac = min(1.0, frac * per_rep / max(dt, 1e-3))
    runs = stratified_rows(ws, frac, strata)
    sample_w = sum(int(ws[v1] - ws[v0]) for v0, v1 in runs)
    cpu_count, _ = run(runs)  # warm-up
    times = []
    for _ in range(reps):
        c, dt = run(runs)
        assert c == cpu_count
        times.append(dt)
    dt = statistics.median(times)
    gpu_count = sum(dg.count(spread, a, int(ws[v0]), int(ws[v1])) for v0, v1 in runs)
    model, logical = cpu_model()
    return {
        "value": sample_w / dt, "unit": "wedge checks/s", "cores": threads, "kind": "port",
        "cpu_model": model, "host_logical_cpus": logical,
        "sample": f"oracle merge intersection over {len(runs)} row runs spread evenly over the wedge order "
                  f"({sample_w} of {W} wedges, {100.0 * sample_w / W:.2f}%), {threads} host threads, "
                  f"one warm-up then median of {reps}: {dt:.2f} s",
        "seconds": dt, "seconds_all": times, "sample_wedges": sample_w,
        "sample_triangles": cpu_count, "gpu_sample_triangles": gpu_count,
        "sample_parity": cpu_count == gpu_count,
    }


def batch_graphs(k, base_seed):
    """The batching workload (Erdos-Renyi / Barabasi-Albert, n log-uniform in
    [64, 65536], average degree uniform in [4, 64]), oriented."""
    import workloads as WL
    return WL.batch_graphs(k, base_seed)


def batch_bench(ctx, k, base_seed):
    """Batched vs plain graphs/s on the same host-resident oriented CSRs.
    plain: one reference-style tc_GPU call per graph (alloc, copy, count, copy
    back, free); batched: one upload of all graphs and one persistent counting
    launch.  Both use the autotuner's spread and bitmap length per graph;
    the batched counts take the table's batch row's kernel and block shape."""
    t0 = time.perf_counter()
    graphs = batch_graphs(k, base_seed)
    tgen = time.perf_counter() - t0
    ctx.use_tuned(wtc.autotune_shape(graphs[0].n, graphs[0].m, graphs[0].wedges(), batch=True))
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
    resident copy of the graph (with the autotuner's kernel and block shape
    for the graph, as the drop-in takes them)."""
    ctx = wtc.Context(0)

    def run(g, params):
        rows = []
        ctx.use_tuned(wtc.autotune_shape(g.n, g.m, g.wedges()))
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


def serve_bench(lams, seconds, queue_capacity, ways, chunk, waves, slots=2):
    """Serving workload: Poisson arrivals of count requests on seeded R-MAT
    graphs (scales 10..22, edge factor 16, two seeds per scale), each with the
    autotuner's spread and bitmap length.  Graphs are registered once and stay
    resident; requests are admitted by HBM footprint and batched continuously.
    Per arrival rate: p50/p99 latency, completed graphs/s, rejected requests,
    and exactness of every completed request against the oracle's count of
    its graph (tests/golden) and against the graph served alone.  The plain
    path on the same mix is one reference-style tc_GPU call per request
    (alloc, copy, count, copy back, free)."""
    import workloads as WL

    keys = WL.serve_keys()
    graphs = WL.serve_graphs()
    oracle_counts = load_golden_serve(keys)
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
    srv = wtc.Server(0, hbm_budget_bytes=32 << 30, max_batch=256, slots=slots, queue_capacity=queue_capacity,
                     probe_ways=ways, lds_chunk=chunk, waves_per_block=waves)
    gids = [srv.register(g) for g in graphs]
    for gid, (s_g, a_g) in zip(gids, tune):  # each graph alone once: resident, server warmed up
        srv.wait(srv.submit(gid, s_g, a_g))
    alone, alone_ms = [], []
    for gid, (s_g, a_g) in zip(gids, tune):  # then timed, each graph alone
        cnt, st = srv.wait(srv.submit(gid, s_g, a_g))
        alone.append(cnt)
        alone_ms.append(st.service_ms)
    alone_oracle = oracle_counts is not None and alone == oracle_counts
    log(f"serve: {len(graphs)} graphs resident, alone service ms by scale: "
        + ", ".join(f"{sc}:{alone_ms[i]:.2f}" for i, (sc, _) in enumerate(keys) if i % WL.SERVE_SEEDS_PER_SCALE == 0))
    expect = oracle_counts if oracle_counts is not None else alone
    results = []
    for lam in lams:
        pick, res, span = WL.poisson_run(srv, gids, tune, lam, seconds, seed=int(lam))
        done = [(i, r) for i, r in enumerate(res) if r is not None]
        lat = np.asarray([r[1] for _, r in done]) if done else np.zeros(1)
        exact = all(r[0] == expect[pick[i]] for i, r in done)
        rec = {"lambda": lam, "requests": len(res), "completed": len(done), "rejected": len(res) - len(done),
               "graphs_per_s": len(done) / span, "p50_ms": float(np.percentile(lat, 50)),
               "p99_ms": float(np.percentile(lat, 99)), "exact": bool(exact)}
        log(f"serve lambda={lam}: {rec}")
        results.append(rec)
    stats = srv.stats()
    srv.close()
    best = max(r["graphs_per_s"] for r in results) if results else 0.0
    return {"graphs": len(graphs), "scales": [WL.SERVE_SCALES[0], WL.SERVE_SCALES[-1]], "seconds_per_rate": seconds,
            "slots": slots, "queue_capacity": queue_capacity, "params": "autotuned per graph",
            "exact_against": "oracle counts (tests/golden)" if oracle_counts is not None else "each graph served alone",
            "alone_equals_oracle": alone_oracle, "alone_service_ms": alone_ms, "rates": results,
            "plain_graphs_per_s": plain_gps, "served_max_graphs_per_s": best,
            "served_over_plain": best / plain_gps, "server": stats}


def load_golden_serve(keys):
    """The oracle's triangle count of every serving-mix graph, in `keys`
    order, from tests/golden (None when absent or for another mix)."""
    path = os.path.join(ROOT, "tests", "golden", "big_golden.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        mix = json.load(f).get("serve_mix", [])
    by_key = {(r["scale"], r["seed"]): int(r["triangles"]) for r in mix if r.get("edge_factor", 16) == 16}
    if not all(k in by_key for k in keys):
        return None
    return [by_key[k] for k in keys]


def _time_count(dg, s, a, reps=3):
    """Median device time (prefix + bitmap/transposition + count, ms) of a
    count, after one untimed count, and the count itself."""
    c0 = dg.count(s, a)
    ts = []
    for _ in range(reps):
        c, st = dg.count(s, a, stats=True)
        assert c == c0
        ts.append(st.prefix_ms + st.bitmap_ms + st.count_ms)
    return statistics.median(ts), c0


def tune_sweep(args):
    """The autotuner's sweep: per size bucket, the fastest (spread, a,
    kernel, block waves, probe ways, row-kernel waves) measured on this GPU,
    written as the table the library loads (tuning/mi355x.tsv beside it).
    Skewed buckets: R-MAT graphs (edge factor 16, seed 1) at --tune-scales;
    flat buckets: Erdos-Renyi graphs of average degree 16; batch: the
    4096-graph batching workload.  Stage 1 takes (kernel, s, a) at the
    default block shape, stage 2 the block shape for that choice."""
    import math

    import workloads as WL

    ctx = wtc.Context(0)
    rows = []

    def sweep_alone(g, family_skew, workload, dg=None, big=False):
        # big: a graph built on the GPU, where only the split count is swept
        # (the walk's top-a bitmap in HBM takes a*a/8 bytes)
        if dg is None:
            n, m, W = g.n, g.m, g.wedges()
            dg = ctx.upload(g)
        else:  # W from a count's own wedge total (no host copy)
            n, m = dg.n, dg.m
            ctx.use_tuned(wtc.Tuned(12, 262144, wtc.KERNEL_SPLIT, 4, 4, 8))
            W = dg.count(12, min(262144, ((n + 31) // 32) * 32), stats=True)[1].wedges
        n32 = ((n + 31) // 32) * 32
        alens = [0] + [x for x in (2048, 8192, 32768, 131072, 262144) if x <= n32 and W > 0]
        # bitmap rows past 2^18 only for the split count (one row in LDS, no a*a/8 bytes in HBM);
        # 65,536 is the largest a whose packed ids take 16 bits
        split_alens = sorted(set([x for x in alens if x] + [x for x in (65536, 524288, 1048576) if x <= n32 and W > 0]))
        best = best_walk = None
        base = wtc.Tuned(0, 0, wtc.KERNEL_WALK, 4, 4, 8)
        kernels = (wtc.KERNEL_SPLIT,) if big else (wtc.KERNEL_WALK, wtc.KERNEL_STAGED, wtc.KERNEL_SPLIT)
        spreads = {wtc.KERNEL_WALK: (2, 4, 8, 16, 32, 64, 128), wtc.KERNEL_STAGED: (2, 4, 8, 16, 32),
                   wtc.KERNEL_SPLIT: (12,)}
        for kernel in kernels:
            for a in (split_alens if kernel == wtc.KERNEL_SPLIT else alens):
                if kernel == wtc.KERNEL_SPLIT and (a == 0 or (big and a < 131072)):
                    continue
                # (the split count has no spread: one value stands in its row of the table)
                log(f"tune {workload}: kernel {kernel} a {a}")
                for sp in spreads[kernel]:
                    if sp > 8 and W < (1 << 20):
                        continue
                    for sw0 in ((8, 16) if kernel == wtc.KERNEL_SPLIT and a > 262144 else (base.split_waves,)):
                        t = wtc.Tuned(sp, a, kernel, base.waves_per_block, base.probe_ways, sw0)
                        ctx.use_tuned(t)
                        ms, _ = _time_count(dg, sp, a, reps=2 if big else 3)
                        if best is None or ms < best[0]:
                            best = (ms, t)
                        if kernel == wtc.KERNEL_WALK and (best_walk is None or ms < best_walk[0]):
                            best_walk = (ms, t)
        ms0, t0 = best
        log(f"tune {workload}: stage 1 best {ms0:.3f} ms {t0.as_dict()}")
        split = t0.kernel == wtc.KERNEL_SPLIT
        # stage 2: the block shape of the winner; a split row also keeps the
        # walk's best spread and block shape, which the walk takes where the
        # split count does not apply (a = 0, a set of graphs)
        walk0 = (best_walk[1] if best_walk else None) if split else t0
        walk_best = None
        if walk0 is not None and not big:  # (graphs built on the GPU: the split count only)
            for waves in (2, 4, 8):
                for ways in (2, 4, 8):
                    t = wtc.Tuned(walk0.spread, walk0.adjacency_matrix_len, walk0.kernel, waves, ways, 8)
                    ctx.use_tuned(t)
                    ms, _ = _time_count(dg, t.spread, t.adjacency_matrix_len)
                    if walk_best is None or ms < walk_best[0]:
                        walk_best = (ms, t)
        if split:
            for sw in ((8, 16) if big else (4, 8, 16)):
                t = wtc.Tuned(t0.spread, t0.adjacency_matrix_len, t0.kernel, 4, 4, sw)
                ctx.use_tuned(t)
                ms, _ = _time_count(dg, t.spread, t.adjacency_matrix_len)
                if ms < best[0]:
                    best = (ms, t)
            if walk_best:  # the walk's spread and block shape for the fallback
                w = walk_best[1]
                best = (best[0], wtc.Tuned(w.spread, best[1].adjacency_matrix_len, wtc.KERNEL_SPLIT, w.waves_per_block,
                                           w.probe_ways, best[1].split_waves))
        elif walk_best and walk_best[0] < best[0]:
            best = walk_best
        dg.free()
        ms, t = best
        a = t.adjacency_matrix_len
        kname = {wtc.KERNEL_WALK: "walk", wtc.KERNEL_STAGED: "staged", wtc.KERNEL_SPLIT: "split"}[t.kernel]
        row = ("alone", family_skew, int(math.floor(math.log2(W))) if W else 0, t.spread, 1 if a else 0, a, kname,
               t.waves_per_block, t.probe_ways, t.split_waves, ms, workload)
        log(f"tune {workload}: n={n} m={m} W={W} -> {row}")
        rows.append(row)

    for scale in [int(x) for x in args.tune_scales.split(",") if x]:
        sweep_alone(WL.rmat(scale), "skew", f"rmat{scale}")
    for scale in [int(x) for x in args.tune_device_scales.split(",") if x]:
        sweep_alone(None, "skew", f"rmat{scale}", dg=wtc.DeviceGraph.gen_rmat(ctx, scale), big=True)
    for lg in [int(x) for x in args.tune_flat.split(",") if x]:
        n = 1 << lg
        g = wtc.preprocess(wtc.gen_erdos_renyi(n, 16.0 / (n - 1), seed=lg))
        sweep_alone(g, "flat", f"er{lg}-deg16")
    if args.tune_batch:
        graphs = WL.batch_graphs(args.tune_batch)
        b = wtc.Batch(ctx, graphs)
        best = None
        for kernel in (wtc.KERNEL_WALK, wtc.KERNEL_STAGED):
            for sp in (4, 8, 12, 16):
                for waves in (2, 4, 8):
                    for ways in (2, 4, 8):
                        ctx.use_tuned(wtc.Tuned(sp, 0, kernel, waves, ways, 8))
                        spreads, alens = [sp] * len(graphs), [0] * len(graphs)
                        b.count(spreads, alens)
                        ts = []
                        for _ in range(3):
                            t0 = time.perf_counter()
                            b.count(spreads, alens)
                            ts.append((time.perf_counter() - t0) * 1e3)
                        ms = statistics.median(ts)
                        if best is None or ms < best[0]:
                            best = (ms, sp, waves, ways, kernel)
        b.free()
        ms, sp, waves, ways, kernel = best
        kname = "staged" if kernel == wtc.KERNEL_STAGED else "walk"
        rows.append(("batch", "any", 0, sp, 0, 0, kname, waves, ways, 8, ms, f"batch{args.tune_batch}"))
        log(f"tune batch: {rows[-1]}")
    ctx.close()
    # buckets below the smallest measured graph of a family take its row
    path = args.tune_out or os.path.join(PKG, "tuning", "mi355x.tsv")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("# Autotuner table for MI355X, written by: python bench.py --tune --tune-scales "
                f"{args.tune_scales} --tune-device-scales {args.tune_device_scales} --tune-flat {args.tune_flat} "
                f"--tune-batch {args.tune_batch}\n")
        f.write("# family skew log2w spread a_div a_cap kernel waves ways split_waves  measured_ms workload\n")
        by_fam = {}
        for r in rows:
            by_fam.setdefault((r[0], r[1]), []).append(r)
        for key, rs in by_fam.items():
            rs.sort(key=lambda r: r[2])
            rs[0] = (rs[0][0], rs[0][1], 0) + tuple(rs[0][3:])  # the smallest graph's row covers the buckets below it
            for r in rs:
                f.write(" ".join(str(x) for x in r[:10]) + f"  {r[10]:.3f} {r[11]}\n")
    print(json.dumps({"tuning_table": path, "rows": rows}), flush=True)


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
    """Mean per-dispatch value of `counter` for the kernel named kernel_sub
    (template arguments aside), from rocprofv3 counter_collection.csv files."""
    import re
    pat = re.compile(r"\b" + re.escape(kernel_sub) + r"[<(]")  # the kernel itself, not a longer name
    vals = {}
    for r in _rocprof_rows(directory, "*counter_collection.csv"):
        if pat.search(r.get("Kernel_Name", "")) and r.get("Counter_Name") == counter:
            key = r.get("Dispatch_Id") or r.get("Correlation_Id")
            vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
    if not vals:
        raise SystemExit(f"no {counter} rows for {kernel_sub} under {directory}")
    return sum(vals.values()) / len(vals), len(vals)


def pmc_summary(args):
    """Turns four rocprofv3 --pmc runs (FETCH_SIZE and WRITE_SIZE, each on the
    dominant kernel and on the calibration copy) into HBM bytes per launch of
    that kernel (--pmc-kernel), each counter scaled by what it reads for the
    copy of known size."""
    kdir_f, kdir_w, cdir_f, cdir_w = args.pmc_summary
    kern = args.pmc_kernel
    copy_bytes = args.copy_calibration_mb * 2 ** 20
    kf, nk = _per_dispatch(kdir_f, kern, "FETCH_SIZE")
    kw, _ = _per_dispatch(kdir_w, kern, "WRITE_SIZE")
    cf, _ = _per_dispatch(cdir_f, "k_copy_u32x4", "FETCH_SIZE")
    cw, _ = _per_dispatch(cdir_w, "k_copy_u32x4", "WRITE_SIZE")
    f_scale = copy_bytes / (cf * 1024.0)  # true bytes per counted byte
    w_scale = copy_bytes / (cw * 1024.0)
    fetch = kf * 1024.0 * f_scale
    write = kw * 1024.0 * w_scale
    out = {"workload": args.pmc_workload, "kernel": kern, "dispatches": nk,
           "fetch_kib_per_dispatch": kf, "write_kib_per_dispatch": kw,
           "copy_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
           "fetch_scale": f_scale, "write_scale": w_scale,
           "fetch_bytes": fetch, "write_bytes": write, "bytes_per_launch": fetch + write}
    out["config"] = {"spread": args.spread, "a": args.a}
    path = os.path.join(ROOT, "profiles", traffic_name(args.pmc_workload, args.spread, args.a, kern))
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def trace_summary(args):
    """Markdown summary of a rocprofv3 --kernel-trace --stats -f csv run: per
    kernel calls, average and total duration (from the kernel trace, so one
    row per dispatched kernel name), written to --summary-out with the
    command that made the trace."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(args.trace_summary, "**", "*kernel_trace.csv"), recursive=True):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no kernel_trace.csv under {args.trace_summary}")
    agg = {}
    for r in rows:
        name = r["Kernel_Name"]
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        a = agg.setdefault(name, [0, 0.0, float("inf"), 0.0])
        a[0] += 1
        a[1] += dur
        a[2] = min(a[2], dur)
        a[3] = max(a[3], dur)
    total = sum(a[1] for a in agg.values())
    lines = [f"# {args.summary_title}", "", "Command (from the repository root, on the MI355X):", "",
             f"    {args.summary_cmd}", "",
             "| kernel | calls | avg ms | min ms | max ms | total ms | share |", "|---|---|---|---|---|---|---|"]
    for name, (k, tot, lo, hi) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        short = name.split("(")[0].replace("void ", "")
        lines.append(f"| `{short}` | {k} | {tot / k:.4f} | {lo:.4f} | {hi:.4f} | {tot:.3f} | {100 * tot / total:.1f}% |")
    text = "\n".join(lines) + "\n"
    with open(args.summary_out, "w") as f:
        f.write(text)
    print(text)


def load_golden_rmat(scale, edge_factor, seed):
    """The oracle's whole-graph triangle count of a seeded R-MAT graph, where
    tests/golden holds one (R-MAT-24, edge factor 16, seed 1)."""
    path = os.path.join(ROOT, "tests", "golden", "big_golden.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        gold = json.load(f)
    rec = gold.get(f"rmat{scale}")
    if rec and rec.get("edge_factor", 16) == edge_factor and rec.get("seed", 1) == seed:
        return int(rec["triangles"])
    return None


def offsets_u64_leg(ctx, g, W, spread, a, steps, warmup, expect):
    """The headline graph uploaded again with 8-byte row offsets and counted
    the same way: the 64-bit CSR layout at the same size."""
    dg8 = ctx.upload(g, offset_bytes=8)
    try:
        for _ in range(warmup):
            dg8.count(spread, a)
        counts, stats = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            c, st = dg8.count(spread, a, stats=True)
            counts.append(c)
            stats.append(st.as_dict())
        el = time.perf_counter() - t0
    finally:
        dg8.free()
    _, b_min8 = g.alg_bytes(8)
    count_ms = statistics.median(s["count_ms"] for s in stats)
    out = {"offset_bytes": 8, "value": W / (el / steps), "unit": "wedge checks/s", "ms_per_step": 1e3 * el / steps,
           "kernel_ms": {"count": count_ms, "prefix": statistics.median(s["prefix_ms"] for s in stats),
                         "bitmap_or_transposition": statistics.median(s["bitmap_ms"] for s in stats)},
           "count_frac_b_min": b_min8 / (count_ms / 1e3) / HBM_PEAK,
           "triangles": counts[0], "counts_equal_u32_offsets": all(c == expect for c in counts)}
    if stats[-1]["split"]:  # the row kernel's roofline, as for the headline
        rows_ms = statistics.median(s["rows_ms"] for s in stats)
        # in-edge records of 2*4 B when every position fits 32 bits (the split
        # count then runs on a 4-byte copy of the offsets), else 2*8 B
        rec_bytes = 8.0 if g.m < 0xFFFFFFFF else 16.0
        dom_bytes = split_id_bytes(min(a, g.n)) * stats[-1]["top_wedges"] + rec_bytes * stats[-1]["in_edges"]
        out["kernel_ms"]["split_rows"] = rows_ms
        out["kernel_ms"]["split_slabs"] = statistics.median(s["list_ms"] for s in stats)
        out["roofline"] = {"kernel": "k_split_rows", "achieved": dom_bytes / (rows_ms / 1e3) / 1e9, "unit": "GB/s",
                           "frac": dom_bytes / (rows_ms / 1e3) / HBM_PEAK, "algorithmic_bytes": dom_bytes}
    return out


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
    ap.add_argument("--ways", type=int, default=None, help="probes advanced together per lane (default: autotuner)")
    ap.add_argument("--chunk", type=int, default=12, help="wedges per lane staged in LDS per phase (staged kernel)")
    ap.add_argument("--waves", type=int, default=None, help="waves per counting block, 1/2/4/8 (default: autotuner)")
    ap.add_argument("--kernel", type=int, default=None, help="0 walk, 1 staged, 2 split (default: autotuner)")
    ap.add_argument("--split-waves", type=int, default=None, help="waves per row-kernel block, 4/8/16 (default: autotuner)")
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
    ap.add_argument("--pmc-kernel", default="k_split_rows", help="kernel name the PMC summary covers")
    ap.add_argument("--trace-summary", default="", help="summarise a rocprofv3 kernel-trace CSV dir (no GPU use)")
    ap.add_argument("--summary-out", default="", help="markdown file --trace-summary writes")
    ap.add_argument("--summary-title", default="rocprofv3 kernel trace")
    ap.add_argument("--summary-cmd", default="", help="the profiled command, quoted in the summary")
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
    ap.add_argument("--serve-rates", default="100,250,500,1000,2000,3500,5000",
                    help="Poisson arrival rates of the serving leg (empty: skip it)")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-slots", type=int, default=2, help="execution slots (streams) of the server")
    ap.add_argument("--no-u64-leg", action="store_true", help="skip the 8-byte-offset leg of the headline graph")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (at most the CPUs this process may use)")
    ap.add_argument("--ingest-only", action="store_true", help="run only the text/CSR-file ingestion timing")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner's sweep and write the tuning table the library loads")
    ap.add_argument("--tune-scales", default="12,14,16,18,20,22,24", help="R-MAT scales of the skewed buckets")
    ap.add_argument("--tune-device-scales", default="26,28",
                    help="R-MAT scales of the skewed buckets generated on the GPU (split count only)")
    ap.add_argument("--tune-flat", default="14,18,22", help="log2 n of the Erdos-Renyi (flat) buckets")
    ap.add_argument("--tune-batch", type=int, default=4096, help="graphs of the batch bucket (0: skip)")
    ap.add_argument("--tune-out", default="", help="where --tune writes the table (default: tuning/mi355x.tsv "
                                                   "beside the library)")
    args = ap.parse_args()

    if args.pmc_summary:
        pmc_summary(args)
        return
    if args.trace_summary:
        trace_summary(args)
        return
    if args.ingest_only:
        print(json.dumps(ingest_bench(args.scale, args.seed)), flush=True)
        return
    if args.tune:
        tune_sweep(args)
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
        res = serve_bench([float(x) for x in args.serve_rates.split(",")], args.serve_seconds, 512,
                          args.ways or 4, args.chunk, args.waves or 4, args.serve_slots)
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
        ctx.set_tuning(args.ways or 4, args.chunk, args.waves or 4)
        res = small_kernel_bench(ctx, args.batch_graphs or 4096, 7000)
        ctx.close()
        print(json.dumps(res), flush=True)
        return
    if args.batch_only:
        ctx = wtc.Context(0)
        ctx.set_tuning(args.ways or 4, args.chunk, args.waves or 4)
        res = batch_bench(ctx, args.batch_graphs or 4096, 1000)
        res["small_graph_batches"] = small_kernel_bench(ctx, args.batch_graphs or 4096, 7000)
        ctx.close()
        print(json.dumps(res), flush=True)
        return

    workload = f"rmat{args.scale}"
    ctx = wtc.Context(0)
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
    # the autotuner's choice for this graph; flags given on the command line override it
    tuned = wtc.autotune_shape(g.n, g.m, W)
    args.spread = args.spread or tuned.spread
    args.a = tuned.adjacency_matrix_len if args.a < 0 else args.a
    for k, v in (("kernel", args.kernel), ("waves_per_block", args.waves), ("probe_ways", args.ways),
                 ("split_waves", args.split_waves)):
        if v is not None:
            setattr(tuned, k, v)
    ctx.use_tuned(tuned)
    ctx.set_tuning(tuned.probe_ways, args.chunk, tuned.waves_per_block)
    log(f"spread={args.spread} a={args.a} tuned={tuned.as_dict()} table={wtc.tuning_source()}")
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
    split = bool(stats[-1]["split"])
    if split:
        # dominant kernel: the split count's row kernel, timed by its own HIP
        # events on the count's stream.  Its algorithmic bytes: the packed id
        # of every wedge it checks (16/P bytes; each tail id read once; the
        # probe itself is an LDS read) and every in-edge's (lo, hi) record of
        # 2*ob bytes
        dom, dom_ms = "k_split_rows", statistics.median(s["rows_ms"] for s in stats)
        top_w, in_e = stats[-1]["top_wedges"], stats[-1]["in_edges"]
        id_bytes = split_id_bytes(min(args.a, dg.n))
        dom_bytes = id_bytes * top_w + 2.0 * ob * in_e
        dom_units = {"wedges_checked": top_w, "in_edges": in_e, "bytes_per_wedge": id_bytes,
                     "bytes_per_in_edge": 2 * ob}
    else:
        # dominant kernel: the counting kernel.  Algorithmic bytes: the CSR once
        # plus, per wedge, the 4-byte id of its far end and one 4-byte probe word
        dom, dom_ms = "k_walk", count_ms
        dom_bytes = b_min
        dom_units = {"wedges_checked": W, "bytes_per_wedge": 8}
    achieved = dom_bytes / (dom_ms / 1e3)
    traffic = load_traffic(workload, args.spread, args.a, dom)
    golden = load_golden_rmat(args.scale, args.edge_factor, args.seed)
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
                      "bitmap_or_transposition": statistics.median(s["bitmap_ms"] for s in stats),
                      "split_rows": statistics.median(s["rows_ms"] for s in stats) if split else None,
                      "split_slabs": statistics.median(s["list_ms"] for s in stats) if split else None,
                      "grid": stats[0]["blocks"]},
        "wedge_checks_per_s_kernel": W / (count_ms / 1e3),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "algorithmic_bytes": dom_bytes, "units": dom_units, "kernel_ms": dom_ms,
                     # the whole count (every counting kernel) against the reference algorithm's
                     # floor of one id and one probe word per wedge from memory, B_min; above 1 when
                     # the split count answers the top-a probes from LDS
                     "b_min": b_min, "count_frac_b_min": b_min / (count_ms / 1e3) / HBM_PEAK,
                     # the reference's own probe words (every binary-search step one word): an upper
                     # figure, most of whose words a bitmap or the caches never fetch from HBM
                     "b_ref_alg": b_alg,
                     # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated on
                     # a copy of known size) over the same launch time: the rate the memory ran at
                     "traffic_GBps": traffic["bytes_per_launch"] / (dom_ms / 1e3) / 1e9 if traffic else None,
                     "traffic_frac": traffic["bytes_per_launch"] / (dom_ms / 1e3) / HBM_PEAK if traffic else None,
                     "traffic_profile": traffic_name(workload, args.spread, args.a, dom) if traffic else None},
        "prepro_s": t2 - t1,
        "builder": "device" if device_build else "host",
        "tuning": dict(tuned.as_dict(), table=wtc.tuning_source()),
    }
    if golden is not None:
        out["triangles_expected"] = golden
        out["triangles_match_expected"] = counts[0] == golden
        log(f"triangles {'match' if counts[0] == golden else 'DIFFER FROM'} the oracle's whole-graph count {golden}")
    # the same graph with 8-byte row offsets (the 64-bit build's CSR layout)
    if not args.no_u64_leg and ob == 4:
        out["offsets_u64"] = offsets_u64_leg(ctx, g, W, args.spread, args.a, args.steps, args.warmup,
                                             counts[0])
        log(f"u64 offsets: {out['offsets_u64']['ms_per_step']:.2f} ms/step, "
            f"count kernel {out['offsets_u64']['kernel_ms']['count']:.2f} ms")
    # size-independent check at full size: four wedge ranges sum to the whole count
    if not args.no_range_check:
        cuts = [0, W // 5, W // 2, (4 * W) // 5, W]
        parts = [dg.count(args.spread, args.a, cuts[i], cuts[i + 1]) for i in range(4)]
        out["range_split"] = {"cuts": cuts, "parts": parts, "sum_equals_total": sum(parts) == counts[0]}
        log(f"range split: {parts} sum ok = {sum(parts) == counts[0]}")
    if not args.no_cpu:
        # every host CPU this process may use, up to the 16 a one-GPU share of the host holds
        threads = min(len(os.sched_getaffinity(0)), args.cpu_threads)
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
        out["serving"] = serve_bench([float(x) for x in args.serve_rates.split(",")], args.serve_seconds, 512,
                                     args.ways or 4, args.chunk, args.waves or 4, args.serve_slots)
        log(f"served/plain = {out['serving']['served_over_plain']:.2f}x")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
