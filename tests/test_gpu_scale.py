"""The batched and served configurations at their full size on the GPU,
checked against the oracle.

Batched: 10^4 uint8 texts, lengths log-uniform in [2^10, 2^24] (about
1.7e10 symbols, so arena and text offsets pass 2^32), sigma from
{4, 20, 64, 256}, uniform or Zipf(1.1) -- the texts the benchmark draws --
built as one forest; the benchmark's own 10^8 mixed queries (per-text count
proportional to length) run as the autotuned launch replayed from a HIP
graph.  The first three queries of every text (one of each kind) and 3e4
more drawn over the batch are checked against direct scans of the texts
copied back from the device; the graph replayed over the forest's query
lines gives the same 10^8 answers.

Served: 64 uint8 trees, lengths log-uniform in [2^20, 2^32] (about 3.3e10
symbols), sigma in {4, 256}, admitted into one server that also builds
their query lines, select hints and select lines.  Mixed requests against
every tree, among them one that spans several device batches, are
submitted together and answered in the mix; access is checked against the
text, rank against the oracle's one-pass sweep, select by construction
(select(c, rank(c, p) + 1) = p for c = text[p], test_benchmark_utils.cuh:
417-528), and the same requests served one per batch (the plain form) give
the same answers.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[scale]", *a, flush=True)


def test_forest_full_batched_configuration(gpu):
    import ecl_wavelet_amd as E
    rng = np.random.default_rng(9)
    N = 10_000
    lens = np.floor(2.0 ** rng.uniform(10, 24, N)).astype(np.uint64)
    sig = rng.choice([4, 20, 64, 256], N).astype(np.uint64)
    zipf = rng.random(N) < 0.5
    starts = np.zeros(N, np.uint64)
    pos = 0
    for i in range(N):
        starts[i] = pos
        pos += int((int(lens[i]) + 15) // 16 * 16)
    total = int(lens.sum())
    assert pos > 2 ** 32
    _say(f"forest: {N} texts, {total} symbols: generate")
    buf = E.DeviceBuffer(pos)
    thr = {int(x): E.zipf_thresholds(int(x), 1.1) for x in (4, 20, 64, 256)}
    for i in range(N):
        E.generate_text_into(buf, int(starts[i]), int(lens[i]), int(sig[i]), np.uint8,
                             seed=100_000 + i, thresholds=thr[int(sig[i])] if zipf[i] else None)
    E.synchronize()
    _say("forest: build")
    forest = E.Forest(buf, starts, lens, sig)
    host = buf.download(np.uint8, pos)
    per = np.maximum(1, np.round(1e8 * lens.astype(np.float64) / total)).astype(np.uint64)
    qoff = np.concatenate([[0], np.cumsum(per)]).astype(np.uint64)
    m = int(qoff[-1])
    d_q = E.DeviceBuffer(m * E.FOREST_QUERY_DTYPE.itemsize)
    d_out = E.DeviceBuffer(m * 8)
    forest.generateQueries(qoff, 77, d_q)
    E.synchronize()
    _say(f"forest: autotune + graph over {m} queries")
    forest.autotune(d_q, m, d_out)
    g = forest.graph(d_q, m, d_out)
    g.launch()
    E.synchronize()
    out = d_out.download(np.uint64, m)
    q = d_q.download(E.FOREST_QUERY_DTYPE, m)
    # every text's first three queries (kinds q % 3 cover all three) and 3e4
    # drawn over the whole batch
    first = (qoff[:-1, None] + np.arange(3, dtype=np.uint64)[None, :]).ravel()
    first = first[first < np.repeat(qoff[1:], 3)]
    pick = np.unique(np.concatenate([first, rng.integers(0, m, 30_000).astype(np.uint64)]))
    _say(f"forest: check {len(pick)} answers")
    qs, got = q[pick], out[pick]
    exp = np.empty(len(pick), np.uint64)
    order = np.argsort(qs["tree"], kind="stable")
    bounds = np.searchsorted(qs["tree"][order], np.arange(N + 1))
    checked = np.zeros(N, bool)
    for t in range(N):
        sel = order[bounds[t]:bounds[t + 1]]
        if len(sel) == 0:
            continue
        checked[t] = True
        text = host[int(starts[t]):int(starts[t]) + int(lens[t])]
        qq = qs[sel]
        kind = qq["kind"].astype(np.int64)
        arg = qq["arg"].astype(np.uint64)
        sym = qq["symbol"].astype(np.uint64)
        e = np.empty(len(sel), np.uint64)
        a, r, s = kind == 0, kind == 1, kind == 2
        e[a] = text[arg[a]]
        if r.any():
            e[r] = O.scan_rank_batch(text, arg[r], sym[r])
        if s.any():
            e[s] = O.scan_select_batch(text, arg[s], sym[s])
        exp[sel] = e
    assert checked.all()
    bad = np.flatnonzero(got != exp)
    assert len(bad) == 0, (len(bad), pick[bad[:5]].tolist(), got[bad[:5]].tolist(),
                           exp[bad[:5]].tolist())
    # the same graph over the forest's query lines and select hints
    _say("forest: query lines")
    forest.buildQueryLines()
    g.launch()
    E.synchronize()
    assert np.array_equal(d_out.download(np.uint64, m), out)
    # with the select lines, then over the quad layouts and hex lines (a
    # graph captured afterwards runs the pair-lane kernel): all 1e8 answers
    # equal the directory's, which the sample above checked
    _say("forest: select lines")
    forest.buildSelectLines()
    g.launch()
    E.synchronize()
    assert np.array_equal(d_out.download(np.uint64, m), out)
    _say("forest: quad layouts")
    forest.buildQuadLayouts()
    forest.autotune(d_q, m, d_out)
    gq = forest.graph(d_q, m, d_out)
    gq.launch()
    E.synchronize()
    assert np.array_equal(d_out.download(np.uint64, m), out)
    del gq


def _requests_for(text, sigma, sizes, rng):
    """Mixed requests (MIXED_DTYPE) with their answers: access at uniform
    positions, rank at uniform positions with symbols from the text or
    uniform, select by construction (and k = count + 1, which gives n)."""
    from ecl_wavelet_amd import MIXED_DTYPE
    n = len(text)
    tot = sum(sizes)
    kinds = rng.integers(0, 3, tot)
    na, nr, ns = (kinds == 0).sum(), (kinds == 1).sum(), (kinds == 2).sum()
    ai = rng.integers(0, n, na).astype(np.uint64)
    ri = rng.integers(0, n + 1, nr).astype(np.uint64)
    rc = rng.integers(0, sigma, nr).astype(np.uint64)
    half = nr // 2
    inside = ri[:half] < n
    rc[:half][inside] = text[ri[:half][inside]]
    sp = rng.integers(0, n, ns).astype(np.uint64)
    sc = text[sp].astype(np.uint64)
    both = O.sweep_rank(text, np.concatenate([ri, sp]), np.concatenate([rc, sc]), sigma=256)
    sk = both[nr:] + 1
    sexp = sp.copy()
    cnt = O.histogram(text, 256)
    past = rng.random(ns) < 0.01
    sk[past] = cnt[sc[past]] + 1
    sexp[past] = n
    q = np.zeros(tot, dtype=MIXED_DTYPE)
    e = np.zeros(tot, np.uint64)
    q["kind"] = kinds
    q["arg"][kinds == 0], e[kinds == 0] = ai, text[ai]
    q["arg"][kinds == 1], q["symbol"][kinds == 1], e[kinds == 1] = ri, rc, both[:nr]
    q["arg"][kinds == 2], q["symbol"][kinds == 2], e[kinds == 2] = sk, sc, sexp
    out, at = [], 0
    for m in sizes:
        out.append((q[at:at + m], e[at:at + m]))
        at += m
    return out


def test_server_full_serving_configuration(gpu):
    import ecl_wavelet_amd as E
    rng = np.random.default_rng(10)
    T = 64
    lens = np.floor(2.0 ** rng.uniform(20, 32, T)).astype(np.uint64)
    sig = rng.choice([4, 256], T)
    srv = E.Server(max_batch=1 << 22, max_inflight=3, query_lines=2)
    big = int(np.argmax(lens))
    reqs = []  # (tree, queries, expected)
    for i in range(T):
        n, s = int(lens[i]), int(sig[i])
        _say(f"server: tree {i}: n={n} sigma={s}")
        d = E.generate_text(n, s, np.uint8, seed=200_000 + i)
        tree = srv.addTree(d, s, n=n)
        text = d.download(np.uint8, n)
        d.free()
        sizes = [int(2 ** rng.uniform(0, 16)) for _ in range(3)]
        if i == big:
            sizes.append(9_000_000)  # spans three batches of 2^22
        for q, e in _requests_for(text, s, sizes, rng):
            reqs.append((tree, q, e))
        del text
    st = srv.stats()
    _say(f"server: {st['trees']} trees, {st['hbm_used'] / 1e9:.1f} GB resident; "
         f"{len(reqs)} requests")
    order = rng.permutation(len(reqs))
    ids = [(j, srv.submit(reqs[j][0], reqs[j][1])) for j in order]
    got = {}
    for j, rid in ids:
        got[j] = srv.wait(rid)[0]
    for j, (tree, q, e) in enumerate(reqs):
        assert np.array_equal(got[j], e), (j, tree, len(q), int(np.sum(got[j] != e)))
    # the plain per-call form gives the same answers
    srv.setMode(False)
    for j in order[:24]:
        out, _ = srv.wait(srv.submit(reqs[j][0], reqs[j][1]))
        assert np.array_equal(out, reqs[j][2])
    srv.close()
