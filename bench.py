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


def load_traffic(workload):
    """HBM bytes per counting-kernel launch from the committed rocprofv3 PMC
    summary, if one exists for this workload."""
    path = os.path.join(ROOT, "profiles", f"k_count_{workload}_traffic.json")
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
    return {
        "value": sample_w / dt, "unit": "wedge checks/s", "cores": threads, "kind": "port",
        "sample": f"oracle merge-intersection over rows [{v0},{v1}) = {sample_w} of {W} wedges "
                  f"({100.0 * sample_w / W:.2f}%), {dt:.1f} s on {threads} host threads",
        "seconds": dt, "sample_triangles": cpu_count, "gpu_sample_triangles": gpu_count,
        "sample_parity": cpu_count == gpu_count,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--spread", type=int, default=5)
    ap.add_argument("--a", type=int, default=8192)
    ap.add_argument("--offset-bytes", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if wtc.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)

    workload = f"rmat{args.scale}"
    t0 = time.perf_counter()
    sym = wtc.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    t1 = time.perf_counter()
    g = wtc.preprocess(sym)
    t2 = time.perf_counter()
    del sym
    W = g.wedges()
    log(f"{workload}: n={g.n} m={g.m} W={W} gen {t1 - t0:.1f}s prepro {t2 - t1:.1f}s")
    ob = args.offset_bytes or (4 if g.m < 2 ** 32 - 1 else 8)
    b_alg, b_min = g.alg_bytes(ob)

    ctx = wtc.Context(0)
    dg = ctx.upload(g, offset_bytes=ob)
    log(f"uploaded {dg.nbytes / 1e9:.2f} GB")
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
    traffic = load_traffic(workload)
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
                     "b_alg": b_alg, "b_min": b_min},
        "prepro_s": t2 - t1,
    }
    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        out["cpu_baseline"] = cpu_baseline(g, dg, args.spread, args.a, args.cpu_seconds, threads)
        log(f"cpu baseline: {out['cpu_baseline']['sample']}")
    dg.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
