"""Benchmark of the hot path on one MI355X.

Workload (the full-size configuration the headline metric is quoted on):
a uint8 text of 6*10^9 symbols drawn uniformly (counter-based hash, seed 4)
from sigma = 256 symbols, resident in HBM.  One step = one wavelet-tree
build over the resident text + Q queries of each kind (access, rank,
select; uniform positions/symbols, drawn on the device before the timed
region) through the device-resident entry points, plus the same 3Q queries
as one mixed-kind batch in a single launch.

value = aggregate device-resident query throughput of the per-kind calls
(access+rank+select queries / their kernel time).  build_gbps = text bytes /
build time.  The roofline object is for the dominant kernel (the construction
pass, which writes the bits of every level from one read of the text), with
per-launch time from HIP events inside the library.
The CPU baseline runs the oracle's OpenMP direct scans on a bounded sample
of the same text and queries.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("access/rank/select queries/s + build GB/s at 6GB σ=256; "
          "batched vs plain, p99")
HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=6_000_000_000)
    ap.add_argument("--sigma", type=int, default=256)
    ap.add_argument("--queries", type=int, default=10_000_000,
                    help="queries of each kind per step")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-calls", type=int, default=200)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def level_bytes(n, sigma):
    """Per-level algorithmic bytes of the level-at-a-time build (read n_l
    symbols, write n_{l+1} symbols except on the last level, write n_l/8
    bytes of bits) for a uniform text over a power-of-two alphabet."""
    L = int(np.ceil(np.log2(sigma)))
    out = []
    for l in range(L):
        b = n + n / 8 + (n if l + 1 < L else 0)
        out.append(b)
    return out


def pass_bytes(n, sigma):
    """Compulsory bytes of the single construction pass over a uint8 text:
    read the n symbols once, write the n/8 bytes of bits of every level."""
    L = int(np.ceil(np.log2(sigma)))
    return n + L * n / 8


def cpu_baseline(n, sigma, seed, args):
    import oracle as O
    t0 = time.time()
    text = O.gen_uniform(seed, n, sigma, np.uint8)
    log(f"cpu: generated host text in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(123)
    threads = O.max_threads()
    # access: O(1) per query
    na = 2_000_000
    idx = rng.integers(0, n, na).astype(np.uint64)
    t0 = time.time()
    O.scan_access_batch(text, idx)
    ta = (time.time() - t0) / na
    # rank / select: O(n) scans per query; sample sized to the budget
    per = max(threads, 8)
    c = rng.integers(0, sigma, per).astype(np.uint64)
    i = rng.integers(0, n, per).astype(np.uint64)
    t0 = time.time()
    O.scan_rank_batch(text, i, c)
    tr = (time.time() - t0) / per
    k = np.array([1 + rng.integers(0, max(1, n // sigma)) for _ in range(per)], np.uint64)
    t0 = time.time()
    O.scan_select_batch(text, k, c)
    ts = (time.time() - t0) / per
    spent = ta * na + (tr + ts) * per
    # extend the rank/select sample while inside the budget
    extra = int(max(0, args.cpu_seconds - spent) / max(tr + ts, 1e-9) / 2) // per * per
    if extra:
        c2 = rng.integers(0, sigma, extra).astype(np.uint64)
        i2 = rng.integers(0, n, extra).astype(np.uint64)
        t0 = time.time()
        O.scan_rank_batch(text, i2, c2)
        tr = (tr * per + (time.time() - t0)) / (per + extra)
        k2 = np.array([1 + rng.integers(0, max(1, n // sigma)) for _ in range(extra)], np.uint64)
        t0 = time.time()
        O.scan_select_batch(text, k2, c2)
        ts = (ts * per + (time.time() - t0)) / (per + extra)
    del text
    mixed = 3.0 / (ta + tr + ts)
    return {
        "value": mixed, "unit": "queries/s", "cores": threads, "kind": "port",
        "sample": (f"oracle direct scans (OpenMP, {threads} threads) over the same "
                   f"6e9-symbol host text: {na} access, {per + extra} rank, "
                   f"{per + extra} select queries; value = 1:1:1 mixed rate"),
        "access_qps": 1.0 / ta, "rank_qps": 1.0 / tr, "select_qps": 1.0 / ts,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU per process; "
                          "multi-GPU runs are not part of it", "n_gpus": args.gpus}))
        sys.exit(2)
    import ecl_wavelet_amd as E
    from ecl_wavelet_amd import _lib

    if _lib.device_count() < 1:
        raise SystemExit("bench.py needs a GPU")
    n, sigma, Q = args.n, args.sigma, args.queries
    seed = 4
    log(f"device: {E.device_name()}  free/total: {E.mem_info()}")
    t0 = time.time()
    d_text = E.generate_text(n, sigma, np.uint8, seed=seed)
    E.synchronize()
    log(f"generated {n} symbols on the device in {time.time() - t0:.2f}s")
    alphabet = np.arange(sigma, dtype=np.uint8)

    d_idx = E.DeviceBuffer(Q * 8)
    d_rq = E.DeviceBuffer(Q * 16)
    d_sq = E.DeviceBuffer(Q * 16)
    d_mq = E.DeviceBuffer(3 * Q * 16)
    d_out = E.DeviceBuffer(3 * Q * 8)
    timer = E.Timer()

    build_ms, q_ms, mixed_ms = [], [], []
    kind_ms = {"access": [], "rank": [], "select": []}
    wt = None
    pass_launch_ms = []
    build_kernels = {}
    for step in range(args.warmup + args.steps):
        timed = step >= args.warmup
        wt = None  # free the previous tree before building the next one
        E.Profiler.reset()
        E.Profiler.enable(timed)
        timer.start()
        wt = E.WaveletTree(d_text, alphabet, dtype=np.uint8, n=n)
        b = timer.stop()
        E.Profiler.enable(False)
        if timed:
            cnt, tot = E.Profiler.get("wt_pass")
            pass_launch_ms.append(tot / max(cnt, 1))
            for kname in ("tile_hist", "tile_scan", "wt_pass", "rs_build", "fill"):
                kc, kt = E.Profiler.get(kname)
                build_kernels.setdefault(kname, []).append((kc, kt))
        wt.generateQueries(0, 100 + step, Q, d_idx)
        wt.generateQueries(1, 200 + step, Q, d_rq)
        wt.generateQueries(2, 300 + step, Q, d_sq)
        wt.generateQueries(3, 400 + step, 3 * Q, d_mq)
        E.synchronize()
        times = {}
        timer.start()
        wt.accessDevice(d_idx, Q, d_out, flags=2)
        times["access"] = timer.stop()
        timer.start()
        wt.rankDevice(d_rq, Q, d_out, flags=1)
        times["rank"] = timer.stop()
        timer.start()
        wt.selectDevice(d_sq, Q, d_out, flags=1)
        times["select"] = timer.stop()
        timer.start()
        wt.queryBatchDevice(d_mq, 3 * Q, d_out)
        mx = timer.stop()
        log(f"step {step}{'' if timed else ' (warmup)'}: build {b:.1f} ms, "
            f"access {times['access']:.1f} ms, rank {times['rank']:.1f} ms, "
            f"select {times['select']:.1f} ms, mixed {mx:.1f} ms")
        if timed:
            build_ms.append(b)
            q_ms.append(sum(times.values()))
            mixed_ms.append(mx)
            for k, v in times.items():
                kind_ms[k].append(v)

    # per-call latency of a plain small batch vs the same work batched
    lat = []
    small = 1000
    for j in range(args.latency_calls):
        timer.start()
        wt.rankDevice(d_rq, small, d_out, flags=1)
        lat.append(timer.stop())
    lat = np.array(lat)

    # kernel-level profile of the query kernels (one extra round)
    E.Profiler.reset()
    E.Profiler.enable(True)
    wt.rankDevice(d_rq, Q, d_out, flags=1)
    wt.selectDevice(d_sq, Q, d_out, flags=1)
    wt.accessDevice(d_idx, Q, d_out, flags=2)
    E.Profiler.enable(False)
    rank_k = E.Profiler.get("wt_rank")

    L = wt.getNumLevels()
    del wt
    build = float(np.median(build_ms))
    qtime = float(np.mean(q_ms)) / 1e3
    value = 3 * Q / qtime
    pb = pass_bytes(n, sigma)
    pass_ms = float(np.median(pass_launch_ms))
    achieved = pb / (pass_ms / 1e3) / 1e9
    # algorithmic bytes of the level-at-a-time construction (histogram, level
    # reads and partition writes, bits, RS directory read and write)
    build_bytes = n + sum(level_bytes(n, sigma)) + L * (n / 8 + n / 256)
    kern = {k: {"launches": int(np.median([c for c, _ in v])),
                "ms": float(np.median([t for _, t in v]))}
            for k, v in build_kernels.items()}
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(np.array(build_ms) + np.array(q_ms))),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": (f"wavelet tree over a uint8 text of n={n} uniform over "
                         f"sigma={sigma} (device-resident); per step: 1 build + "
                         f"{Q} each of access/rank/select + one mixed batch of {3 * Q}"),
            "n": n, "sigma": sigma, "queries_per_kind": Q, "levels": L,
        },
        "build_ms": build,
        "build_gbps": n / (build / 1e3) / 1e9,
        "build_algorithmic_gbps": build_bytes / (build / 1e3) / 1e9,
        "build_kernels_ms": kern,
        "access_qps": Q / (np.mean(kind_ms["access"]) / 1e3),
        "rank_qps": Q / (np.mean(kind_ms["rank"]) / 1e3),
        "select_qps": Q / (np.mean(kind_ms["select"]) / 1e3),
        "mixed_batch_qps": 3 * Q / (np.mean(mixed_ms) / 1e3),
        "small_call_rank_p50_ms": float(np.percentile(lat, 50)),
        "small_call_rank_p99_ms": float(np.percentile(lat, 99)),
        "rank_kernel_ms": rank_k[1] / max(rank_k[0], 1),
        "roofline": {
            "kernel": "wt_pass (all levels of the build in one pass)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "bytes_per_launch": pb,
            "launch_ms": pass_ms,
        },
    }
    if not args.no_cpu:
        d_text.free()
        res["cpu_baseline"] = cpu_baseline(n, sigma, seed, args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
