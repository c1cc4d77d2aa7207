"""Benchmark of the hot path on one MI355X.

Workload (the full-size configuration the headline metric is quoted on):
a uint8 text of 6*10^9 symbols drawn uniformly (counter-based hash, seed 4)
from sigma = 256 symbols, resident in HBM.  One step = one wavelet-tree
build over the resident text; Q queries of each kind (access, rank, select;
uniform positions/symbols, drawn on the device before the timed region)
through the device-resident entry points, plus the same 3Q queries as one
mixed-kind batch in a single launch; the query lines and the select lines;
the same queries again over them.

value = aggregate device-resident query throughput of the per-kind calls
over the tree with its query structures (access+rank+select queries / their
kernel time).  The query lines (the level bits in 128-byte lines headed by
their rank, with select hints) and the select lines (a 128-byte line every r
ones or zeros: its position and the 960 bits from there) are built on the
finished tree inside each step and timed on their own
(query_lines_build_ms, select_lines_build_ms; ms_per_step includes them);
the same queries over the plain directory, timed in the same step before
those exist, are reported under plain_directory.  build_gbps = text bytes /
build time (the tree and its directory, without the lines).  query_roofline
prices the per-kind rates against the device's measured rate of random,
dependent 128-byte line reads (one line per level and query).

The roofline object is for the dominant kernel (the construction pass,
which writes the bits of every level from one read of the text), with
per-launch time from HIP events inside the library; `achieved` counts the
pass's share of the build's algorithmic bytes (the level-at-a-time build's
level reads, partition writes and bit writes), and the compulsory bytes of
the one pass (text read once, bits written once) are reported beside it.
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
    ap.add_argument("--cpu-seconds", type=float, default=25.0,
                    help="budget of the CPU-baseline sample")
    ap.add_argument("--cpu-check-queries", type=int, default=1_000_000,
                    help="rank and select queries (each) whose GPU answers the CPU "
                         "baseline's variant B checks (variant A takes a prefix)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-calls", type=int, default=200)
    ap.add_argument("--bt-texts", type=int, default=10_000,
                    help="texts of the batched configuration (0 skips it)")
    ap.add_argument("--bt-queries", type=int, default=100_000_000)
    ap.add_argument("--sv-trees", type=int, default=64,
                    help="resident trees of the serving configuration (0 skips it)")
    ap.add_argument("--sv-max-log2", type=float, default=32.0,
                    help="log2 of the largest serving text")
    ap.add_argument("--bin-log2", type=int, default=34,
                    help="binary RankSelect section: 2^k bits (0: skip)")
    ap.add_argument("--bin-queries", type=int, default=100_000_000)
    ap.add_argument("--full-sweep", type=int, default=1,
                    help="builds and queries at the other full sizes: u8 sigma 4/16/64 "
                         "(n=6e9) and u16 sigma 2^16 Zipf(1.1) (n=2e9); 0 skips")
    ap.add_argument("--host-queries", type=int, default=100_000_000,
                    help="host-to-host section: queries per kind from host memory (0 skips)")
    ap.add_argument("--pmc-queries", action="store_true",
                    help="counter workload of the queries: the 1 GiB copy calibration, one "
                         "build, then the bench's access, rank and select batches once each")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="short run for hardware-counter passes: one 1 GiB "
                         "device-to-device copy (the calibration kernel) and "
                         "two builds of the benchmark tree; prints no JSON")
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


def pass_algorithmic_bytes(n, sigma):
    """The construction pass's share of the build's algorithmic bytes (the
    level-at-a-time build's level reads, stable-partition writes and bit
    writes: every term of the build's byte count except the histogram read,
    the directory's read of the bits and its index writes, which other
    kernels do)."""
    return sum(level_bytes(n, sigma))


def pass_traffic(n, sigma):
    """HBM bytes per construction-pass launch from the committed counter
    profile (FETCH_SIZE and WRITE_SIZE scaled by a 1 GiB copy's counts), for
    the workload that profile covers; None for any other workload."""
    path = os.path.join(ROOT, "profiles", "wt_pass_counters_u8_n6e9_sigma256.txt")
    if (n, sigma) != (6_000_000_000, 256) or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(n, sigma, seed, args, sample):
    """The oracle on the host cores, over the same 6e9-symbol text (the same
    counter-based hash, generated on the host) and the first queries of the
    GPU's own device-generated query arrays (`sample`, with the GPU's answers
    to them, which both variants check).

    Variant A is the literal oracle of the reference's tests: access reads
    text[i]; rank counts c over [0, i); select walks to the k-th c (n when
    absent); OpenMP over queries.  Rank and select cost O(n) per query, so
    they run on a bounded sample sized to --cpu-seconds.  Variant B answers
    the same queries from per-symbol counts before every 2^16 symbols plus a
    scan of one block (a realistic CPU index; the table build is timed
    separately).  `value` is variant A's 1:1:1 mixed rate."""
    import oracle as O
    t0 = time.time()
    text = O.gen_uniform(seed, n, sigma, np.uint8)
    log(f"cpu: generated host text in {time.time() - t0:.1f}s")
    threads = O.max_threads()
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    idx, a_gpu = sample["access_idx"], sample["access_gpu"]
    rq, r_gpu = sample["rank_q"], sample["rank_gpu"]
    sq, s_gpu = sample["select_q"], sample["select_gpu"]
    # ---- variant A
    t0 = time.time()
    a_cpu = O.scan_access_batch(text, idx)
    ta = (time.time() - t0) / len(idx)
    # rank/select: a first batch of 2 per thread estimates the per-query
    # cost; the sample then grows to the time budget (at most the download)
    per = 2 * threads
    t0 = time.time()
    r_a = O.scan_rank_batch(text, rq["index"][:per], rq["symbol"][:per])
    s_a = O.scan_select_batch(text, sq["index"][:per], sq["symbol"][:per])
    first = time.time() - t0
    budget = max(args.cpu_seconds - ta * len(idx), 0.0)
    m_a = int(min(len(rq), max(per, per * budget / max(first, 1e-9)))) // per * per
    t_rank = t_sel = 0.0
    if m_a > per:
        t0 = time.time()
        r_a = np.concatenate([r_a, O.scan_rank_batch(text, rq["index"][per:m_a],
                                                     rq["symbol"][per:m_a])])
        t_rank = time.time() - t0
        t0 = time.time()
        s_a = np.concatenate([s_a, O.scan_select_batch(text, sq["index"][per:m_a],
                                                       sq["symbol"][per:m_a])])
        t_sel = time.time() - t0
        tr, ts = t_rank / (m_a - per), t_sel / (m_a - per)
    else:
        tr = ts = first / (2 * per)
    # ---- variant B
    t0 = time.time()
    B = O.BlockedOracle(text, sigma)
    table_s = time.time() - t0
    t0 = time.time()
    r_b = B.rank_batch(rq["index"], rq["symbol"])
    trb = (time.time() - t0) / len(rq)
    t0 = time.time()
    s_b = B.select_batch(sq["index"], sq["symbol"])
    tsb = (time.time() - t0) / len(sq)
    ok_a = (np.array_equal(a_cpu.astype(np.uint8), a_gpu) and np.array_equal(r_a, r_gpu[:len(r_a)])
            and np.array_equal(s_a, s_gpu[:len(s_a)]))
    ok_b = np.array_equal(r_b, r_gpu) and np.array_equal(s_b, s_gpu)
    del text, B
    return {
        "value": 3.0 / (ta + tr + ts), "unit": "queries/s", "cores": threads, "kind": "port",
        "sample": (f"variant A, the oracle's direct scans (OpenMP, {threads} threads) over the "
                   f"same 6e9-symbol text, generated on the host: the first {len(idx)} access, "
                   f"{len(r_a)} rank and {len(s_a)} select queries of the GPU's query arrays; "
                   f"value = 1:1:1 mixed rate"),
        "omp_threads": threads, "nproc": os.cpu_count(), "affinity_cpus": affinity,
        "access_qps": 1.0 / ta, "rank_qps": 1.0 / tr, "select_qps": 1.0 / ts,
        "rank_select_sample": len(r_a),
        "variant_b": {
            "kind": "port (blocked: per-symbol counts every 2^16 symbols + one-block scan)",
            "table_build_s": table_s, "table_bytes": int(B_bytes(n, sigma)),
            "access_qps": 1.0 / ta, "rank_qps": 1.0 / trb, "select_qps": 1.0 / tsb,
            "mixed_qps": 3.0 / (ta + trb + tsb), "queries_per_kind": len(rq),
        },
        "gpu_answers_equal_variant_a": bool(ok_a),
        "gpu_answers_equal_variant_b": bool(ok_b),
    }


def B_bytes(n, sigma):
    """Bytes of variant B's count table: one row of sigma u64 per 2^16 symbols."""
    return ((n + 65535) // 65536 + 1) * sigma * 8


def bench_batched(args):
    """Batched configuration: args.bt_texts uint8 texts, lengths log-uniform in
    [2^10, 2^24], sigma uniform over {4, 20, 64, 256}, uniform or Zipf(1.1)
    symbols.  Batched form: one forest build, then all queries (per-text
    count proportional to length, mixed 1:1:1) as one launch replayed from a
    HIP graph.  Plain form: one build per text through the reference-shaped
    API and three device calls (access, rank, select) per tree.  Both over the
    plain directory; the batched graph is then replayed over the forest's
    query lines and select hints, then with its select lines too, and a graph
    captured after the quad layouts and hex lines exist is replayed over them
    (reported beside)."""
    import ecl_wavelet_amd as E
    rng = np.random.default_rng(9)
    N = args.bt_texts
    lens = np.floor(2.0 ** rng.uniform(10, 24, N)).astype(np.uint64)
    sig = rng.choice([4, 20, 64, 256], N).astype(np.uint64)
    zipf = rng.random(N) < 0.5
    starts = np.zeros(N, np.uint64)
    pos = 0
    for i in range(N):
        starts[i] = pos
        pos += int((int(lens[i]) + 15) // 16 * 16)
    buf = E.DeviceBuffer(pos)
    thr = {int(x): E.zipf_thresholds(int(x), 1.1) for x in (4, 20, 64, 256)}
    for i in range(N):
        E.generate_text_into(buf, int(starts[i]), int(lens[i]), int(sig[i]), np.uint8,
                             seed=100_000 + i, thresholds=thr[int(sig[i])] if zipf[i] else None)
    E.synchronize()
    total = int(lens.sum())
    log(f"batched: {N} texts, {total} symbols")
    timer = E.Timer()
    # ---- batched build (one forest), warm once
    forest = E.Forest(buf, starts, lens, sig)
    del forest
    timer.start()
    forest = E.Forest(buf, starts, lens, sig)
    forest_build_ms = timer.stop()
    # the construction as one graph (tile size autotuned), replayed
    tile_rows, tuned_ms = forest.autotuneBuild()
    gb = forest.buildGraph()
    gb.launch()
    E.synchronize()
    runs = []
    for _ in range(3):
        timer.start()
        gb.launch()
        runs.append(timer.stop())
    build_graph_ms = float(np.median(runs))
    del gb
    log(f"batched: build {forest_build_ms:.1f} ms (create), {build_graph_ms:.1f} ms as one graph "
        f"(tile rows {tile_rows})")
    # ---- batched queries
    per = np.maximum(1, np.round(args.bt_queries * lens.astype(np.float64) / total)).astype(np.uint64)
    qoff = np.concatenate([[0], np.cumsum(per)]).astype(np.uint64)
    m = int(qoff[-1])
    d_q = E.DeviceBuffer(m * E.FOREST_QUERY_DTYPE.itemsize)
    d_out = E.DeviceBuffer(m * 8)
    forest.generateQueries(qoff, 77, d_q)
    E.synchronize()
    block, grid = forest.autotune(d_q, m, d_out)
    g = forest.graph(d_q, m, d_out)
    g.launch()
    E.synchronize()
    runs = []
    for _ in range(3):
        timer.start()
        g.launch()
        runs.append(timer.stop())
    batched_ms = float(np.median(runs))
    log(f"batched: build {forest_build_ms:.1f} ms, {m} queries in {batched_ms:.2f} ms "
        f"(block {block}, grid {grid})")
    # the same graph over the forest's query lines and select hints
    timer.start()
    forest.buildQueryLines()
    lines_ms = timer.stop()
    g.launch()
    E.synchronize()
    runs = []
    for _ in range(3):
        timer.start()
        g.launch()
        runs.append(timer.stop())
    batched_lines_ms = float(np.median(runs))
    log(f"batched: query lines built in {lines_ms:.1f} ms, {m} queries in "
        f"{batched_lines_ms:.2f} ms over them")
    # and with the select lines too
    timer.start()
    forest.buildSelectLines()
    sel_lines_ms = timer.stop()
    g.launch()
    E.synchronize()
    runs = []
    for _ in range(3):
        timer.start()
        g.launch()
        runs.append(timer.stop())
    batched_all_ms = float(np.median(runs))
    log(f"batched: select lines built in {sel_lines_ms:.1f} ms, {m} queries in "
        f"{batched_all_ms:.2f} ms over all lines")
    # and over the quad layouts of the trees over 4, 64 and 256 symbols (a
    # graph captured after they exist runs the pair-lane kernel)
    timer.start()
    forest.buildQuadLayouts()
    quad_ms = timer.stop()
    qblock, qgrid = forest.autotune(d_q, m, d_out)  # now the pair-lane kernel's grid
    gq = forest.graph(d_q, m, d_out)
    gq.launch()
    E.synchronize()
    runs = []
    for _ in range(3):
        timer.start()
        gq.launch()
        runs.append(timer.stop())
    batched_quad_ms = float(np.median(runs))
    log(f"batched: quad layouts built in {quad_ms:.1f} ms, {m} queries in "
        f"{batched_quad_ms:.2f} ms over them (grid {qgrid})")
    del g, gq, forest
    # ---- plain: one tree per text, per-tree per-kind calls
    alph = {int(x): np.arange(int(x), dtype=np.uint8) for x in (4, 20, 64, 256)}
    timer.start()
    trees = [E.WaveletTree(buf.view(int(starts[i]), int(lens[i])), alph[int(sig[i])],
                           dtype=np.uint8, n=int(lens[i])) for i in range(N)]
    plain_build_ms = timer.stop()
    kinds = [(per + 2 - k) // 3 for k in range(3)]  # access, rank, select counts per tree
    offs = [np.concatenate([[0], np.cumsum(kc)]).astype(np.uint64) for kc in kinds]
    d_idx = E.DeviceBuffer(int(offs[0][-1]) * 8)
    d_rq = E.DeviceBuffer(int(offs[1][-1]) * 16)
    d_sq = E.DeviceBuffer(int(offs[2][-1]) * 16)
    d_po = E.DeviceBuffer(max(int(o[-1]) for o in offs) * 8)
    for i, wt in enumerate(trees):
        for k, (d, w) in enumerate(((d_idx, 8), (d_rq, 16), (d_sq, 16))):
            c = int(kinds[k][i])
            if c:
                wt.generateQueries(k, 500 + 100 * k + i, c, d.view(int(offs[k][i]) * w, c * w))
    E.synchronize()

    def plain_calls(i, wt):
        ka, kr, ks = int(kinds[0][i]), int(kinds[1][i]), int(kinds[2][i])
        if ka:
            wt.accessDevice(d_idx.view(int(offs[0][i]) * 8, ka * 8), ka, d_po.view(0, ka * 8),
                            flags=2)
        if kr:
            wt.rankDevice(d_rq.view(int(offs[1][i]) * 16, kr * 16), kr, d_po.view(0, kr * 8),
                          flags=1)
        if ks:
            wt.selectDevice(d_sq.view(int(offs[2][i]) * 16, ks * 16), ks, d_po.view(0, ks * 8),
                            flags=1)

    for i, wt in enumerate(trees[:64]):  # warm
        plain_calls(i, wt)
    E.synchronize()
    timer.start()
    for i, wt in enumerate(trees):
        plain_calls(i, wt)
    plain_ms = timer.stop()
    # per-tree latency of the plain calls (each tree's three calls, synchronised)
    lat = []
    for i in rng.choice(N, min(N, 300), replace=False):
        timer.start()
        plain_calls(int(i), trees[int(i)])
        lat.append(timer.stop())
    del trees
    log(f"plain: build {plain_build_ms:.1f} ms, queries {plain_ms:.2f} ms")
    return {
        "config": (f"{N} uint8 texts, lengths log-uniform in [2^10, 2^24] (total {total}), "
                   f"sigma in {{4, 20, 64, 256}}, uniform or Zipf(1.1); {m} mixed queries "
                   f"(per-text count proportional to length)"),
        "batched_build_ms": forest_build_ms,
        "batched_build_graph_ms": build_graph_ms,
        "build_tile_rows": tile_rows,
        "plain_build_ms": plain_build_ms,
        "build_speedup": plain_build_ms / forest_build_ms,
        "build_graph_speedup": plain_build_ms / build_graph_ms,
        "batched_qps": m / (batched_ms / 1e3),
        "plain_qps": m / (plain_ms / 1e3),
        "query_speedup": plain_ms / batched_ms,
        "batched_launch_ms": batched_ms,
        "query_lines_build_ms": lines_ms,
        "batched_qps_query_lines": m / (batched_lines_ms / 1e3),
        "batched_launch_ms_query_lines": batched_lines_ms,
        "select_lines_build_ms": sel_lines_ms,
        "batched_qps_all_lines": m / (batched_all_ms / 1e3),
        "batched_launch_ms_all_lines": batched_all_ms,
        "quad_layouts_build_ms": quad_ms,
        "batched_qps_quad_layouts": m / (batched_quad_ms / 1e3),
        "batched_launch_ms_quad_layouts": batched_quad_ms,
        "plain_per_tree_p50_ms": float(np.percentile(lat, 50)),
        "plain_per_tree_p99_ms": float(np.percentile(lat, 99)),
        "autotuned": {"block": block, "grid": grid, "build_tile_rows": tile_rows},
    }


def bench_host_path(args, d_text, n, sigma):
    """Host-to-host figures beside the device-resident ones: a build from a
    host copy of the text (its upload included), then host arrays of Q
    queries per kind answered through the pinned, chunked upload / kernel /
    download pipeline.  Wall-clock times around each call."""
    import ecl_wavelet_amd as E
    Q = args.host_queries
    text = d_text.download(np.uint8, n)
    alphabet = np.arange(sigma, dtype=np.uint8)
    # untimed: the pinned staging of host uploads is allocated once per process
    E.WaveletTree(text[: 1 << 28], alphabet)
    t0 = time.perf_counter()
    wt = E.WaveletTree(text, alphabet)
    build_s = time.perf_counter() - t0
    del text
    d_q = E.DeviceBuffer(Q * 16)
    wt.generateQueries(0, 501, Q, d_q)
    idx = d_q.download(np.uint64, Q)
    wt.generateQueries(1, 502, Q, d_q)
    rq = d_q.download(E.QUERY_DTYPE, Q)
    wt.generateQueries(2, 503, Q, d_q)
    sq = d_q.download(E.QUERY_DTYPE, Q)
    del d_q
    wt.access(idx[:1000], flags=2)
    wt.rank(rq[:1000], flags=1)
    wt.select(sq[:1000], flags=1)
    t = {}
    for kind, fn in (("access", lambda: wt.access(idx, flags=2)),
                     ("rank", lambda: wt.rank(rq, flags=1)),
                     ("select", lambda: wt.select(sq, flags=1))):
        t0 = time.perf_counter()
        fn()
        t[kind] = time.perf_counter() - t0
    res = {
        "config": (f"host uint8 text of {n} symbols (sigma={sigma}) and host arrays of {Q} "
                   "queries per kind; wall clock, uploads and downloads included"),
        "build_ms": build_s * 1e3,
        "build_gbps": n / build_s / 1e9,
        "access_qps": Q / t["access"],
        "rank_qps": Q / t["rank"],
        "select_qps": Q / t["select"],
        # bytes crossing PCIe per query: 8 B index + 1 B symbol (access),
        # 16 B query + 8 B result (rank, select)
        "access_pcie_gbps": Q * 9 / t["access"] / 1e9,
        "rank_pcie_gbps": Q * 24 / t["rank"] / 1e9,
        "select_pcie_gbps": Q * 24 / t["select"] / 1e9,
    }
    log(f"host path: build {res['build_ms']:.0f} ms, access {res['access_qps']:.3g}, "
        f"rank {res['rank_qps']:.3g}, select {res['select_qps']:.3g} q/s")
    return res


def bench_binary(args):
    """Binary RankSelect configuration: 2^bin_log2 i.i.d. Bernoulli(p) bits
    drawn on the device, p in {0.5, 0.05}; directory build time, then
    bin_queries rank1(i) with uniform i and select1(k) with uniform k over
    the popcount, generated on the device (timed with their generation);
    then the same queries over the query lines and select hints, and select
    over select lines."""
    import ecl_wavelet_amd as E
    m = 1 << args.bin_log2
    Q = args.bin_queries
    timer = E.Timer()
    out = E.DeviceBuffer(Q * 8)
    res = {"config": (f"{m} bits i.i.d. Bernoulli(p), drawn on the device; {Q} rank1 and "
                      f"{Q} select1 queries generated on the device"),
           "bits": m, "queries": Q}
    for p, seed in ((0.5, 2), (0.05, 3)):
        builds = []
        for rep in range(3):  # median of three builds (the array is consumed)
            rs = None
            ba = E.BitArray([m])
            ba.fillRandom(0, seed, min(int(p * 2.0 ** 64), 2 ** 64 - 1))
            E.synchronize()
            timer.start()
            rs = E.RankSelect(ba)
            builds.append(timer.stop())
        build_ms = float(np.median(builds))
        rs.rankGenerated(0, 1, Q, out, 1)  # warm-up
        rs.selectGenerated(0, 1, Q, out, 1)
        E.synchronize()
        timer.start()
        rs.rankGenerated(0, 11, Q, out, 1)
        rank_ms = timer.stop()
        timer.start()
        rs.selectGenerated(0, 12, Q, out, 1)
        select_ms = timer.stop()
        # the same queries over the query lines and select hints
        timer.start()
        rs.buildQueryLines()
        lines_ms = timer.stop()
        rs.rankGenerated(0, 1, Q, out, 1)  # warm-up
        rs.selectGenerated(0, 1, Q, out, 1)
        E.synchronize()
        timer.start()
        rs.rankGenerated(0, 11, Q, out, 1)
        rank_l_ms = timer.stop()
        timer.start()
        rs.selectGenerated(0, 12, Q, out, 1)
        select_l_ms = timer.stop()
        r = {
            "build_ms": build_ms,
            "build_gbps": m / 8 / (build_ms / 1e3) / 1e9,
            "rank1_qps": Q / (rank_ms / 1e3),
            "select1_qps": Q / (select_ms / 1e3),
            # one bitvector line + directory per query (128+16 B rank1,
            # 256+16 B select1)
            "rank1_line_gbps": Q * 144 / (rank_ms / 1e3) / 1e9,
            "select1_line_gbps": Q * 272 / (select_ms / 1e3) / 1e9,
            "query_lines_build_ms": lines_ms,
            "rank1_qps_query_lines": Q / (rank_l_ms / 1e3),
            "select1_qps_query_lines": Q / (select_l_ms / 1e3),
        }
        # and select over select lines
        timer.start()
        rs.buildSelectLines()
        r["select_lines_build_ms"] = timer.stop()
        rs.selectGenerated(0, 1, Q, out, 1)  # warm-up
        E.synchronize()
        timer.start()
        rs.selectGenerated(0, 12, Q, out, 1)
        r["select1_qps_select_lines"] = Q / (timer.stop() / 1e3)
        res[f"p={p}"] = r
        log(f"binary p={p}: build {build_ms:.2f} ms, rank1 {r['rank1_qps']:.3g} q/s, "
            f"select1 {r['select1_qps']:.3g} q/s; lines {lines_ms:.2f} ms, rank1 "
            f"{r['rank1_qps_query_lines']:.3g} q/s, select1 {r['select1_qps_query_lines']:.3g} q/s; "
            f"select lines {r['select_lines_build_ms']:.2f} ms, select1 "
            f"{r['select1_qps_select_lines']:.3g} q/s")
        del rs
    del out
    return res


def bench_full_sweep(args):
    """The full-size configurations beside the headline one: uint8 texts of
    6e9 symbols uniform over sigma in {4, 16, 64}, and a uint16 text of 2e9
    symbols Zipf(1.1) over sigma = 2^16.  Per text: median of two timed
    builds after one untimed, then 10^7 device-generated queries per kind,
    over the plain directory, again over the query and select lines, access
    and rank over the quad lines, select over the quad select lines, and
    access and rank over the hex lines where the level count is a multiple
    of 4."""
    import ecl_wavelet_amd as E
    timer = E.Timer()
    Q = 10_000_000
    d_idx, d_rq, d_sq = E.DeviceBuffer(Q * 8), E.DeviceBuffer(Q * 16), E.DeviceBuffer(Q * 16)
    d_out = E.DeviceBuffer(Q * 8)
    cases = [(np.uint8, 6_000_000_000, s, None, 5 + j) for j, s in enumerate((4, 16, 64))]
    cases.append((np.uint16, 2_000_000_000, 65536, E.zipf_thresholds(65536, 1.1), 8))
    out = []
    for dt, n, sigma, thr, seed in cases:
        d_text = E.generate_text(n, sigma, dt, seed=seed, thresholds=thr)
        E.synchronize()
        alphabet = np.arange(sigma, dtype=dt)
        builds = []
        for rep in range(3):
            wt = None
            timer.start()
            wt = E.WaveletTree(d_text, alphabet, dtype=dt, n=n)
            b = timer.stop()
            if rep:
                builds.append(b)
        wt.generateQueries(0, 31, Q, d_idx)
        wt.generateQueries(1, 32, Q, d_rq)
        wt.generateQueries(2, 33, Q, d_sq)
        E.synchronize()
        calls = (("access", lambda: wt.accessDevice(d_idx, Q, d_out, flags=2)),
                 ("rank", lambda: wt.rankDevice(d_rq, Q, d_out, flags=1)),
                 ("select", lambda: wt.selectDevice(d_sq, Q, d_out, flags=1)))
        t = {}
        for kind, call_ in calls:
            timer.start()
            call_()
            t[kind] = timer.stop()
        # the same queries over query lines, select hints and select lines
        timer.start()
        wt.buildQueryLines()
        wt.buildSelectLines()
        lines_ms = timer.stop()
        tl = {}
        for kind, call_ in calls:
            timer.start()
            call_()
            tl[kind] = timer.stop()
        # access and rank over the quad lines (every sweep alphabet is a power
        # of two with an even number of levels)
        timer.start()
        wt.buildQuadLines()
        quad_ms = timer.stop()
        tq = {}
        for kind, call_ in calls[:2]:
            timer.start()
            call_()
            tq[kind] = timer.stop()
        # select over the quad select lines
        timer.start()
        wt.buildQuadSelectLines()
        qsel_ms = timer.stop()
        timer.start()
        calls[2][1]()
        tq["select"] = timer.stop()
        # access and rank over the hex lines (level counts that are multiples of 4)
        hex_ms, th = None, {}
        if int(wt.getNumLevels()) % 4 == 0:
            timer.start()
            wt.buildHexLines()
            hex_ms = timer.stop()
            for kind, call_ in calls[:2]:
                timer.start()
                call_()
                th[kind] = timer.stop()
        build_ms = float(np.median(builds))
        r = {"dtype": np.dtype(dt).name, "n": n, "sigma": sigma,
             "distribution": "uniform" if thr is None else "zipf(1.1)",
             "levels": int(wt.getNumLevels()), "build_ms": build_ms,
             "build_gbps": n * np.dtype(dt).itemsize / (build_ms / 1e3) / 1e9,
             "access_qps": Q / (t["access"] / 1e3), "rank_qps": Q / (t["rank"] / 1e3),
             "select_qps": Q / (t["select"] / 1e3),
             "query_and_select_lines_build_ms": lines_ms,
             "access_qps_lines": Q / (tl["access"] / 1e3), "rank_qps_lines": Q / (tl["rank"] / 1e3),
             "select_qps_lines": Q / (tl["select"] / 1e3),
             "quad_lines_build_ms": quad_ms,
             "access_qps_quad": Q / (tq["access"] / 1e3), "rank_qps_quad": Q / (tq["rank"] / 1e3),
             "quad_select_lines_build_ms": qsel_ms,
             "select_qps_quad": Q / (tq["select"] / 1e3),
             "hex_lines_build_ms": hex_ms,
             "access_qps_hex": Q / (th["access"] / 1e3) if th else None,
             "rank_qps_hex": Q / (th["rank"] / 1e3) if th else None}
        out.append(r)
        log(f"sweep {r['dtype']} sigma={sigma} n={n}: build {build_ms:.1f} ms, access "
            f"{r['access_qps']:.3g}, rank {r['rank_qps']:.3g}, select {r['select_qps']:.3g} q/s; "
            f"with lines ({lines_ms:.1f} ms): access {r['access_qps_lines']:.3g}, rank "
            f"{r['rank_qps_lines']:.3g}, select {r['select_qps_lines']:.3g} q/s; "
            f"quad lines ({quad_ms:.1f} ms): access {r['access_qps_quad']:.3g}, rank "
            f"{r['rank_qps_quad']:.3g} q/s; quad select lines ({qsel_ms:.1f} ms): select "
            f"{r['select_qps_quad']:.3g} q/s"
            + (f"; hex lines ({hex_ms:.1f} ms): access {r['access_qps_hex']:.3g}, rank "
               f"{r['rank_qps_hex']:.3g} q/s" if th else ""))
        del wt, d_text
    return out


SERVE_RATES = (1e3, 2e3, 5e3, 1e4, 2e4, 5e4, 1e5, 2e5)
SERVE_P99_TARGETS_MS = (2.0, 5.0, 10.0, 20.0, 50.0, 100.0, 300.0)


def _serve_sweep(srv, mode, coalesce):
    """Open-loop Poisson load at every rate of SERVE_RATES (at least about
    4000 requests each), stopping once p99 passes one second (the form is
    saturated); returns one row per rate."""
    srv.setMode(coalesce)
    rows = []
    for rate in SERVE_RATES:
        secs = max(0.2, 4000.0 / rate)
        st = srv.load(rate, secs, seed=int(rate), min_size=1, max_size=1_000_000, pool=1 << 20)
        row = {"rate_req_s": rate, "seconds": secs,
               **{k: st[k] for k in ("requests", "queries", "batches", "p50_ms", "p99_ms",
                                     "queries_per_s")}}
        rows.append(row)
        log(f"serving {mode} @ {rate:g} req/s: {st['queries_per_s']:.3g} q/s, "
            f"p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, {st['batches']} batches")
        if st["p99_ms"] > 1000.0:
            break
    return rows


def _best_under(rows, p99_ms):
    """Highest delivered throughput among the rates whose p99 stays within
    p99_ms (0 when none does)."""
    ok = [r["queries_per_s"] for r in rows if r["p99_ms"] <= p99_ms]
    return max(ok) if ok else 0.0


def bench_serving(args):
    """Serving configuration: args.sv_trees resident uint8 trees, lengths
    log-uniform in [2^20, 2^sv_max_log2], sigma in {4, 256}; Poisson request
    streams of mixed batches of log-uniform size in [1, 10^6] against a
    random tree, swept over offered rates, served with coalescing and in the
    plain per-call form (one request per batch, one batch at a time).  The
    two forms are compared at matched p99: for each p99 target, the highest
    throughput each form delivers at a rate whose p99 stays within it.  The
    sweep runs over the trees' reference-shaped directory, then over a
    server whose trees also carry query lines, select hints and select
    lines, then over one whose trees carry the quad layouts (quad lines,
    quad select lines and, at sigma = 256, hex lines: the layouts of the
    headline numbers).  A tree the HBM budget cannot admit is left out and
    counted."""
    import ecl_wavelet_amd as E
    rng = np.random.default_rng(10)
    T = args.sv_trees
    lens = np.floor(2.0 ** rng.uniform(20, args.sv_max_log2, T)).astype(np.uint64)
    sig = rng.choice([4, 256], T)
    out = {"config": (f"{T} resident uint8 trees, lengths log-uniform in [2^20, "
                      f"2^{args.sv_max_log2:g}] (total {int(lens.sum())}), sigma in {{4, 256}}; "
                      f"Poisson requests of log-uniform size in [1, 10^6] mixed queries"),
           "p99_targets_ms": list(SERVE_P99_TARGETS_MS)}
    timer = E.Timer()
    from ecl_wavelet_amd._lib import EclError, ECL_ERROR_OUT_OF_MEMORY_TREE
    for layout, lines in (("directory", 0), ("lines", 2), ("quad", 3)):
        srv = E.Server(max_batch=1 << 22, max_inflight=3, query_lines=lines)
        timer.start()
        admitted, symbols = 0, 0
        for i in range(T):
            d = E.generate_text(int(lens[i]), int(sig[i]), np.uint8, seed=200_000 + i)
            try:
                srv.addTree(d, int(sig[i]), n=int(lens[i]))
                admitted += 1
                symbols += int(lens[i])
            except EclError as e:
                if e.status != ECL_ERROR_OUT_OF_MEMORY_TREE:
                    raise
            d.free()
        admit_ms = timer.stop()
        st0 = srv.stats()
        log(f"serving ({layout}): {admitted} of {T} trees, {symbols} symbols, admitted in "
            f"{admit_ms:.0f} ms, {st0['hbm_used'] / 1e9:.1f} GB resident")
        served = _serve_sweep(srv, "served", True)
        plain = _serve_sweep(srv, "plain", False)
        srv.close()
        matched = []
        for tgt in SERVE_P99_TARGETS_MS:
            a, b = _best_under(served, tgt), _best_under(plain, tgt)
            matched.append({"p99_ms": tgt, "served_qps": a, "plain_qps": b,
                            "ratio": (a / b) if b > 0 else None})
        out[layout] = {"admit_ms": admit_ms, "hbm_used_gb": st0["hbm_used"] / 1e9,
                       "trees_admitted": admitted, "symbols_admitted": symbols,
                       "served": served, "plain": plain, "matched_p99": matched}
    return out


def pmc_workload(args):
    """Kernels for the FETCH_SIZE / WRITE_SIZE passes: the HIP runtime's
    copy kernel over 1 GiB (known traffic: 1 GiB read + 1 GiB written) and
    two builds of the benchmark tree."""
    import ctypes as C
    import ecl_wavelet_amd as E
    hip = C.CDLL("libamdhip64.so")
    nbytes = 1 << 30
    a, b = E.DeviceBuffer(nbytes), E.DeviceBuffer(nbytes)
    rc = hip.hipMemcpy(b.ptr, a.ptr, C.c_size_t(nbytes), 3)
    if rc != 0:
        raise SystemExit(f"hipMemcpy failed: {rc}")
    E.synchronize()
    a.free()
    b.free()
    log(f"copy calibration: {nbytes} bytes device to device")
    d_text = E.generate_text(args.n, args.sigma, np.uint8, seed=4)
    alphabet = np.arange(args.sigma, dtype=np.uint8)
    for _ in range(2):
        wt = E.WaveletTree(d_text, alphabet, dtype=np.uint8, n=args.n)
        E.synchronize()
        del wt
    log(f"built the tree twice (n={args.n}, sigma={args.sigma})")


def pmc_queries(args):
    """Kernels for the query counter passes: the copy calibration, then one
    build and 10^7 device-generated queries per kind over the benchmark
    tree (each batch launched once), then the query lines and the same three
    batches over them, then the select lines and the select batch again,
    then the quad lines and quad select lines and the three batches over
    them, then the hex lines and the access and rank batches over them."""
    import ctypes as C
    import ecl_wavelet_amd as E
    hip = C.CDLL("libamdhip64.so")
    nbytes = 1 << 30
    a, b = E.DeviceBuffer(nbytes), E.DeviceBuffer(nbytes)
    rc = hip.hipMemcpy(b.ptr, a.ptr, C.c_size_t(nbytes), 3)
    if rc != 0:
        raise SystemExit(f"hipMemcpy failed: {rc}")
    E.synchronize()
    a.free()
    b.free()
    d_text = E.generate_text(args.n, args.sigma, np.uint8, seed=4)
    wt = E.WaveletTree(d_text, np.arange(args.sigma, dtype=np.uint8), dtype=np.uint8, n=args.n)
    Q = args.queries
    d_idx, d_rq, d_sq = E.DeviceBuffer(Q * 8), E.DeviceBuffer(Q * 16), E.DeviceBuffer(Q * 16)
    d_out = E.DeviceBuffer(Q * 8)
    wt.generateQueries(0, 100, Q, d_idx)
    wt.generateQueries(1, 200, Q, d_rq)
    wt.generateQueries(2, 300, Q, d_sq)
    E.synchronize()
    wt.accessDevice(d_idx, Q, d_out, flags=2)
    wt.rankDevice(d_rq, Q, d_out, flags=1)
    wt.selectDevice(d_sq, Q, d_out, flags=1)
    E.synchronize()
    # the same batches again over the query lines, then select over select lines
    wt.buildQueryLines()
    wt.accessDevice(d_idx, Q, d_out, flags=2)
    wt.rankDevice(d_rq, Q, d_out, flags=1)
    wt.selectDevice(d_sq, Q, d_out, flags=1)
    E.synchronize()
    wt.buildSelectLines()
    wt.selectDevice(d_sq, Q, d_out, flags=1)
    E.synchronize()
    # access and rank over the quad lines, select over the quad select lines
    wt.buildQuadSelectLines()
    wt.accessDevice(d_idx, Q, d_out, flags=2)
    wt.rankDevice(d_rq, Q, d_out, flags=1)
    wt.selectDevice(d_sq, Q, d_out, flags=1)
    E.synchronize()
    # access and rank over the hex lines
    wt.buildHexLines()
    wt.accessDevice(d_idx, Q, d_out, flags=2)
    wt.rankDevice(d_rq, Q, d_out, flags=1)
    E.synchronize()
    log(f"queries: {Q} access, rank and select over n={args.n}, sigma={args.sigma}, "
        f"without and with query lines; select with select lines; the three over the quad "
        f"lines and quad select lines; access and rank over the hex lines")


def run_queries(wt, timer, Q, d_idx, d_rq, d_sq, d_mq, d_out):
    """Milliseconds of the three per-kind calls and the mixed batch."""
    t = {}
    timer.start()
    wt.accessDevice(d_idx, Q, d_out, flags=2)
    t["access"] = timer.stop()
    timer.start()
    wt.rankDevice(d_rq, Q, d_out, flags=1)
    t["rank"] = timer.stop()
    timer.start()
    wt.selectDevice(d_sq, Q, d_out, flags=1)
    t["select"] = timer.stop()
    timer.start()
    wt.queryBatchDevice(d_mq, 3 * Q, d_out)
    t["mixed"] = timer.stop()
    return t


def main():
    args = parse()
    if args.pmc_workload:
        pmc_workload(args)
        return
    if args.pmc_queries:
        pmc_queries(args)
        return
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

    build_ms, lines_ms, sel_ms, q_ms, mixed_ms = [], [], [], [], []
    quad_ms, plain_total_ms = [], []
    kind_ms = {"access": [], "rank": [], "select": []}
    quad_kind_ms = {"access": [], "rank": [], "select": []}
    quad_build_ms, quad_sel_build_ms, quad_mixed_ms = [], [], []
    # quad lines: power-of-two alphabets with an even number of levels
    use_quad = sigma >= 4 and (sigma & (sigma - 1)) == 0 and (sigma.bit_length() - 1) % 2 == 0
    # hex lines: ... whose level count is a multiple of 4
    use_hex = use_quad and (sigma.bit_length() - 1) % 4 == 0
    hex_ms, hex_build_ms, hex_mixed_ms = [], [], []
    hex_kind_ms = {"access": [], "rank": []}
    plain_ms = {"access": [], "rank": [], "select": [], "mixed": []}
    lines_bytes = sel_bytes = qsel_bytes = 0
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
        plain = run_queries(wt, timer, Q, d_idx, d_rq, d_sq, d_mq, d_out)
        free0 = E.mem_info()[0]
        timer.start()
        wt.buildQueryLines()
        lb = timer.stop()
        free1 = E.mem_info()[0]
        lines_bytes = free0 - free1
        timer.start()
        wt.buildSelectLines()
        sb = timer.stop()
        sel_bytes = free1 - E.mem_info()[0]
        times = run_queries(wt, timer, Q, d_idx, d_rq, d_sq, d_mq, d_out)
        mx = times.pop("mixed")
        qt = {}
        if use_quad:
            # quad lines, then access and rank over them (one line per two levels)
            timer.start()
            wt.buildQuadLines()
            qb = timer.stop()
            timer.start()
            wt.accessDevice(d_idx, Q, d_out, flags=2)
            qt["access"] = timer.stop()
            timer.start()
            wt.rankDevice(d_rq, Q, d_out, flags=1)
            qt["rank"] = timer.stop()
            # quad select lines, then select over them
            free2 = E.mem_info()[0]
            timer.start()
            wt.buildQuadSelectLines()
            qsb = timer.stop()
            qsel_bytes = free2 - E.mem_info()[0]
            timer.start()
            wt.selectDevice(d_sq, Q, d_out, flags=1)
            qt["select"] = timer.stop()
            # the mixed batch over them (a pair of lanes per query)
            timer.start()
            wt.queryBatchDevice(d_mq, 3 * Q, d_out)
            qmx = timer.stop()
        ht = {}
        if use_hex:
            # hex lines, then access and rank over them (one line per four levels)
            timer.start()
            wt.buildHexLines()
            hb = timer.stop()
            timer.start()
            wt.accessDevice(d_idx, Q, d_out, flags=2)
            ht["access"] = timer.stop()
            timer.start()
            wt.rankDevice(d_rq, Q, d_out, flags=1)
            ht["rank"] = timer.stop()
            # the mixed batch: access and rank over the hex lines, select over
            # the quad select lines
            timer.start()
            wt.queryBatchDevice(d_mq, 3 * Q, d_out)
            hmx = timer.stop()
        log(f"step {step}{'' if timed else ' (warmup)'}: build {b:.1f} ms, "
            f"plain access {plain['access']:.2f} ms, rank {plain['rank']:.2f} ms, "
            f"select {plain['select']:.2f} ms, mixed {plain['mixed']:.2f} ms; lines {lb:.2f} ms, "
            f"select lines {sb:.2f} ms; "
            f"access {times['access']:.2f} ms, rank {times['rank']:.2f} ms, "
            f"select {times['select']:.2f} ms, mixed {mx:.2f} ms"
            + (f"; quad lines {qb:.2f} ms, access {qt['access']:.2f} ms, "
               f"rank {qt['rank']:.2f} ms; quad select lines {qsb:.2f} ms, "
               f"select {qt['select']:.2f} ms, mixed {qmx:.2f} ms" if use_quad else "")
            + (f"; hex lines {hb:.2f} ms, access {ht['access']:.2f} ms, "
               f"rank {ht['rank']:.2f} ms, mixed {hmx:.2f} ms" if use_hex else ""))
        if timed:
            build_ms.append(b)
            lines_ms.append(lb)
            sel_ms.append(sb)
            q_ms.append(sum(times.values()))
            mixed_ms.append(mx)
            plain_total_ms.append(sum(plain.values()))
            for k, v in times.items():
                kind_ms[k].append(v)
            for k, v in plain.items():
                plain_ms[k].append(v)
            if use_quad:
                quad_ms.append(qb + qsb + qt["access"] + qt["rank"] + qt["select"] + qmx)
                quad_mixed_ms.append(qmx)
                quad_build_ms.append(qb)
                quad_sel_build_ms.append(qsb)
                for k, v in qt.items():
                    quad_kind_ms[k].append(v)
            if use_hex:
                hex_ms.append(hb + ht["access"] + ht["rank"] + hmx)
                hex_build_ms.append(hb)
                hex_mixed_ms.append(hmx)
                for k, v in ht.items():
                    hex_kind_ms[k].append(v)

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

    # the first queries of each kind (the last step's device-generated
    # arrays) and the GPU's answers to them: the CPU baseline runs on these
    # and checks the answers
    ka, kr = min(Q, 2_000_000), min(Q, args.cpu_check_queries)
    sample = {"access_idx": d_idx.download(np.uint64, ka),
              "rank_q": d_rq.download(E.QUERY_DTYPE, kr),
              "select_q": d_sq.download(E.QUERY_DTYPE, kr)}
    wt.accessDevice(d_idx, ka, d_out, flags=2)
    sample["access_gpu"] = d_out.download(np.uint8, ka)
    wt.rankDevice(d_rq, kr, d_out, flags=1)
    sample["rank_gpu"] = d_out.download(np.uint64, kr)
    wt.selectDevice(d_sq, kr, d_out, flags=1)
    sample["select_gpu"] = d_out.download(np.uint64, kr)

    L = wt.getNumLevels()
    del wt
    # the queries' roofline: random dependent 128-byte lines per second on
    # this device (probe over a buffer the size of the tree's query lines,
    # L lines per probe query, two lanes per line like the query kernels)
    line_peak = E.probe_random_lines(int(lines_bytes), Q, steps=L, lanes=2)
    line_peak_half = (E.probe_random_lines(int(lines_bytes), Q, steps=L // 2, lanes=2)
                      if use_quad else None)
    line_peak_quarter = (E.probe_random_lines(int(lines_bytes), Q, steps=L // 4, lanes=2)
                         if use_hex else None)
    build = float(np.median(build_ms))
    qtime = float(np.mean(q_ms)) / 1e3
    lines_value = 3 * Q / qtime
    # value: each kind over the fastest layout the tree has (access and rank
    # over the hex lines, else the quad lines, else the query lines; select
    # over the quad select lines, else the select lines)
    kind_best = {k: float(np.mean(v)) for k, v in kind_ms.items()}
    if use_quad:
        for k, v in quad_kind_ms.items():
            kind_best[k] = float(np.mean(v))
    if use_hex:
        for k, v in hex_kind_ms.items():
            kind_best[k] = float(np.mean(v))
    ar_layout, ar_lines, ar_peak = (("hex lines", L // 4, line_peak_quarter) if use_hex else
                                    ("quad lines", L // 2, line_peak_half) if use_quad else
                                    ("query lines", L, line_peak))
    value = 3 * Q / (sum(kind_best.values()) / 1e3)
    pb = pass_bytes(n, sigma)
    pab = pass_algorithmic_bytes(n, sigma)
    pass_ms = float(np.median(pass_launch_ms))
    achieved = pab / (pass_ms / 1e3) / 1e9
    compulsory = pb / (pass_ms / 1e3) / 1e9
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
        "ms_per_step": float(np.mean(np.array(build_ms) + np.array(plain_total_ms)
                                     + np.array(lines_ms) + np.array(sel_ms)
                                     + np.array(q_ms) + np.array(mixed_ms)
                                     + (np.array(quad_ms) if use_quad else 0.0)
                                     + (np.array(hex_ms) if use_hex else 0.0))),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": (f"wavelet tree over a uint8 text of n={n} uniform over "
                         f"sigma={sigma} (device-resident); per step: 1 build, "
                         f"{Q} each of access/rank/select + one mixed batch of {3 * Q} "
                         f"over the plain directory, the query lines and select "
                         f"lines, the same queries over them, then the quad lines and "
                         f"quad select lines and the three kinds and the mixed batch over "
                         f"them, then the hex lines and access, rank and the mixed batch "
                         f"over them"),
            "n": n, "sigma": sigma, "queries_per_kind": Q, "levels": L,
        },
        "build_ms": build,
        "query_lines_build_ms": float(np.median(lines_ms)),
        "query_lines_bytes": int(lines_bytes),
        "select_lines_build_ms": float(np.median(sel_ms)),
        "select_lines_bytes": int(sel_bytes),
        "build_gbps": n / (build / 1e3) / 1e9,
        "build_algorithmic_gbps": build_bytes / (build / 1e3) / 1e9,
        "build_kernels_ms": kern,
        "access_qps": Q / (kind_best["access"] / 1e3),
        "rank_qps": Q / (kind_best["rank"] / 1e3),
        "select_qps": Q / (kind_best["select"] / 1e3),
        "layouts": {
            "access_rank": ar_layout,
            "select": "quad select lines" if use_quad else "select lines",
        },
        "hex_lines_build_ms": float(np.median(hex_build_ms)) if use_hex else None,
        "hex_lines_bytes": (L // 4) * (n // 128 + 1) * 128 if use_hex else None,
        "quad_lines_layout": ({
            "access_qps": Q / (np.mean(quad_kind_ms["access"]) / 1e3),
            "rank_qps": Q / (np.mean(quad_kind_ms["rank"]) / 1e3),
            "mixed_batch_qps": 3 * Q / (np.mean(quad_mixed_ms) / 1e3),
        } if use_quad else None),
        "quad_lines_build_ms": float(np.median(quad_build_ms)) if use_quad else None,
        "quad_lines_bytes": (L // 2) * (n // 448 + 1) * 128 if use_quad else None,
        "quad_select_lines_build_ms": float(np.median(quad_sel_build_ms)) if use_quad else None,
        "quad_select_lines_bytes": int(qsel_bytes) if use_quad else None,
        "query_lines_layout": {
            "access_qps": Q / (np.mean(kind_ms["access"]) / 1e3),
            "rank_qps": Q / (np.mean(kind_ms["rank"]) / 1e3),
            "select_qps": Q / (np.mean(kind_ms["select"]) / 1e3),
            "per_kind_qps": lines_value,
            "mixed_batch_qps": 3 * Q / (np.mean(mixed_ms) / 1e3),
        },
        "mixed_batch_qps": 3 * Q / (np.mean(hex_mixed_ms if use_hex else
                                            quad_mixed_ms if use_quad else mixed_ms) / 1e3),
        "small_call_rank_p50_ms": float(np.percentile(lat, 50)),
        "small_call_rank_p99_ms": float(np.percentile(lat, 99)),
        "rank_kernel_ms": rank_k[1] / max(rank_k[0], 1),
        "query_roofline": {
            "bound": "random dependent 128-byte line reads",
            "unit": "lines/s",
            "peak": line_peak,
            "peak_source": (f"ecl_probe_random_lines over {int(lines_bytes)} bytes, {Q} probe "
                            f"queries of {L} dependent lines, two lanes per line"),
            "peak_half_chains": line_peak_half,
            "peak_half_chains_source": (f"the same probe with {L // 2} dependent lines per "
                                        f"query (the quad lines' chains)" if use_quad else None),
            "peak_quarter_chains": line_peak_quarter,
            "peak_quarter_chains_source": (f"the same probe with {L // 4} dependent lines per "
                                           f"query (the hex lines' chains)" if use_hex else None),
            **{kind: {"layout": lay, "lines_per_query": lpq,
                      "achieved": Q / (kind_best[kind] / 1e3) * lpq,
                      "peak": pk,
                      "frac": Q / (kind_best[kind] / 1e3) * lpq / pk}
               for kind, lay, lpq, pk in (
                   ("access", ar_layout, ar_lines, ar_peak),
                   ("rank", ar_layout, ar_lines, ar_peak),
                   ("select", "quad select lines" if use_quad else "select lines",
                    L // 2 if use_quad else L, line_peak_half if use_quad else line_peak))},
        },
        "plain_directory": {
            "access_qps": Q / (np.mean(plain_ms["access"]) / 1e3),
            "rank_qps": Q / (np.mean(plain_ms["rank"]) / 1e3),
            "select_qps": Q / (np.mean(plain_ms["select"]) / 1e3),
            "mixed_batch_qps": 3 * Q / (np.mean(plain_ms["mixed"]) / 1e3),
            "per_kind_qps": 3 * Q / (np.mean(np.array(plain_ms["access"]) + np.array(plain_ms["rank"])
                                             + np.array(plain_ms["select"])) / 1e3),
        },
        "roofline": {
            "kernel": "wt_pass (all levels of the build in one pass)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": pass_traffic(n, sigma),
            "traffic_source": "profiles/wt_pass_counters_u8_n6e9_sigma256.txt",
            "bytes_per_launch": pab,
            "bytes_basis": ("algorithmic: the level-at-a-time build's level reads, "
                            "partition writes and bit writes, which this one pass "
                            "replaces"),
            "compulsory_bytes_per_launch": pb,
            "compulsory_gbps": compulsory,
            "compulsory_frac": compulsory / HBM_PEAK_GBPS,
            "limiter": "VALU issue (0.94 of peak in the committed counter profile)",
            "launch_ms": pass_ms,
        },
    }
    del d_idx, d_rq, d_sq, d_mq, d_out
    wt = None
    if args.host_queries > 0:
        res["host_to_host"] = bench_host_path(args, d_text, n, sigma)
    d_text.free()
    if args.bt_texts > 0:
        res["batched"] = bench_batched(args)
    if args.sv_trees > 0:
        res["serving"] = bench_serving(args)
    if args.bin_log2 > 0:
        res["binary_rank_select"] = bench_binary(args)
    if args.full_sweep:
        res["full_size_sweep"] = bench_full_sweep(args)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(n, sigma, seed, args, sample)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
