"""Forests (many trees built and queried together) on the GPU vs the oracle:
every answer of a mixed batch over many trees equals the direct scan of its
text, equals the answer of the same tree built alone, and is unchanged by
the construction tile size, the autotuned launch shape and graph replays."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SIGMAS = [4, 20, 64, 256]


def _texts(seed, count, lo_log2=6, hi_log2=15):
    """count texts, log-uniform lengths, sigma from SIGMAS, uniform or Zipf
    symbols; returns (texts, sigmas)."""
    rng = np.random.default_rng(seed)
    texts, sigmas = [], []
    for i in range(count):
        n = int(2 ** rng.uniform(lo_log2, hi_log2))
        sigma = int(rng.choice(SIGMAS))
        if rng.random() < 0.5:
            t = O.gen_uniform(seed * 1000 + i, n, sigma, np.uint8)
        else:
            t = O.gen_cdf(seed * 1000 + i, n, O.zipf_thresholds(sigma, 1.1), np.uint8)
        texts.append(t.astype(np.uint8))
        sigmas.append(sigma)
    return texts, sigmas


def _forest(texts, sigmas, tile_rows=0):
    """A forest over texts of any length, each starting on a 16-byte
    boundary of one device buffer (the gaps hold other bytes)."""
    from ecl_wavelet_amd import DeviceBuffer, Forest
    starts, pos = [], 0
    for t in texts:
        starts.append(pos)
        pos += (len(t) + 15) // 16 * 16 + 16
    buf = np.full(pos, 0xEE, dtype=np.uint8)  # gap bytes outside every alphabet
    for t, s0 in zip(texts, starts):
        buf[s0:s0 + len(t)] = t
    d = DeviceBuffer.from_array(buf)
    f = Forest(d, np.array(starts, dtype=np.uint64),
               np.array([len(t) for t in texts], dtype=np.uint64),
               np.array(sigmas, dtype=np.uint64), tile_rows=tile_rows)
    return f, d, starts


def _padded_texts(seed, count, **kw):
    return _texts(seed, count, **kw)


def _queries(texts, sigmas, per_text, seed):
    """Mixed queries (kind, tree, arg, symbol) with their direct-scan answers."""
    from ecl_wavelet_amd import FOREST_QUERY_DTYPE
    rng = np.random.default_rng(seed)
    qs, exp = [], []
    for ti, (t, s) in enumerate(zip(texts, sigmas)):
        n = len(t)
        kinds = rng.integers(0, 3, per_text)
        q = np.zeros(per_text, dtype=FOREST_QUERY_DTYPE)
        q["kind"] = kinds
        q["tree"] = ti
        e = np.zeros(per_text, dtype=np.uint64)
        a = kinds == 0
        idx = rng.integers(0, n, a.sum()).astype(np.uint64)
        q["arg"][a] = idx
        e[a] = t[idx]
        r = kinds == 1
        ri = rng.integers(0, n + 1, r.sum()).astype(np.uint64)
        rc = rng.integers(0, s, r.sum()).astype(np.uint64)
        q["arg"][r] = ri
        q["symbol"][r] = rc
        e[r] = O.scan_rank_batch(t, ri, rc)
        sl = kinds == 2
        sc = rng.integers(0, s, sl.sum()).astype(np.uint64)
        cnt = np.bincount(t, minlength=256)
        k = np.array([int(rng.integers(1, cnt[c] + 2)) for c in sc], dtype=np.uint64)
        q["arg"][sl] = k
        q["symbol"][sl] = sc
        e[sl] = O.scan_select_batch(t, k, sc)
        qs.append(q)
        exp.append(e)
    q = np.concatenate(qs)
    e = np.concatenate(exp)
    perm = rng.permutation(len(q))  # trees interleaved in the batch
    return q[perm], e[perm]


def _run(f, q):
    from ecl_wavelet_amd import DeviceBuffer, synchronize
    dq = DeviceBuffer.from_array(q)
    dout = DeviceBuffer(len(q) * 8)
    f.queryDevice(dq, len(q), dout)
    synchronize()
    return dout.download(np.uint64, len(q))


def test_forest_queries_vs_oracle(gpu):
    texts, sigmas = _padded_texts(11, 60)
    f, _, _ = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 300, seed=5)
    got = _run(f, q)
    assert np.array_equal(got, e)
    for i, (t, s) in enumerate(zip(texts, sigmas)):
        n, sigma, L = f.treeInfo(i)
        assert (n, sigma, L) == (len(t), s, O.ceil_log2(s))


@pytest.mark.parametrize("rows", [1, 2, 4])
def test_forest_tile_sizes_agree(gpu, rows):
    texts, sigmas = _padded_texts(12, 40, lo_log2=8, hi_log2=17)
    q, e = _queries(texts, sigmas, 200, seed=6)
    f, _, _ = _forest(texts, sigmas, tile_rows=rows)
    assert np.array_equal(_run(f, q), e)


def test_forest_equals_single_trees(gpu):
    """Every level of every forest tree answers like the tree built alone;
    checked through queries on both."""
    from ecl_wavelet_amd import WaveletTree
    texts, sigmas = _padded_texts(13, 12, lo_log2=10, hi_log2=16)
    f, _, _ = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 400, seed=7)
    got = _run(f, q)
    assert np.array_equal(got, e)
    for i, (t, s) in enumerate(zip(texts, sigmas)):
        wt = WaveletTree(t, np.arange(s, dtype=np.uint8))
        sel = q["tree"] == i
        qi, gi = q[sel], got[sel]
        a = qi["kind"] == 0
        assert np.array_equal(wt.access(qi["arg"][a]).astype(np.uint64), gi[a])
        r = qi["kind"] == 1
        rq = wt.make_queries(qi["arg"][r], qi["symbol"][r].astype(np.uint8))
        assert np.array_equal(wt.rank(rq), gi[r])
        sl = qi["kind"] == 2
        sq = wt.make_queries(qi["arg"][sl], qi["symbol"][sl].astype(np.uint8))
        assert np.array_equal(wt.select(sq), gi[sl])


def test_forest_autotune_and_graph_replays(gpu):
    from ecl_wavelet_amd import DeviceBuffer, synchronize
    texts, sigmas = _padded_texts(14, 30)
    f, _, _ = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 500, seed=8)
    dq = DeviceBuffer.from_array(q)
    dout = DeviceBuffer(len(q) * 8)
    block, grid = f.autotune(dq, len(q), dout)
    assert block in (64, 128, 256, 512) and grid > 0
    f.queryDevice(dq, len(q), dout)
    synchronize()
    assert np.array_equal(dout.download(np.uint64, len(q)), e)
    g = f.graph(dq, len(q), dout)
    zeros = np.zeros(len(q), dtype=np.uint64)
    for _ in range(3):
        dout.upload(zeros)  # eager work between the replays
        synchronize()
        g.launch()
        synchronize()
        assert np.array_equal(dout.download(np.uint64, len(q)), e)


def test_forest_generated_queries(gpu):
    """Queries drawn on the device (per-tree counts) are well formed and
    answered exactly."""
    from ecl_wavelet_amd import DeviceBuffer, FOREST_QUERY_DTYPE, synchronize
    texts, sigmas = _padded_texts(15, 25)
    f, _, _ = _forest(texts, sigmas)
    per = np.array([max(1, len(t) // 64) for t in texts], dtype=np.uint64)
    qoff = np.concatenate([[0], np.cumsum(per)]).astype(np.uint64)
    m = int(qoff[-1])
    dq = DeviceBuffer(m * FOREST_QUERY_DTYPE.itemsize)
    f.generateQueries(qoff, 77, dq)
    synchronize()
    q = dq.download(FOREST_QUERY_DTYPE, m)
    assert np.array_equal(np.repeat(np.arange(len(texts)), per.astype(np.int64)), q["tree"])
    assert np.array_equal(q["kind"], np.arange(m) % 3)
    got = _run(f, q)
    for i, t in enumerate(texts):
        sel = q["tree"] == i
        qi, gi = q[sel], got[sel]
        for kind in range(3):
            s = qi["kind"] == kind
            if kind == 0:
                exp = t[qi["arg"][s]]
            elif kind == 1:
                exp = O.scan_rank_batch(t, qi["arg"][s], qi["symbol"][s].astype(np.uint64))
            else:
                exp = O.scan_select_batch(t, qi["arg"][s], qi["symbol"][s].astype(np.uint64))
            assert np.array_equal(gi[s], exp.astype(np.uint64))


def test_forest_rejects_bad_input(gpu):
    from ecl_wavelet_amd import DeviceBuffer, Forest
    from ecl_wavelet_amd._lib import EclError
    d = DeviceBuffer.from_array(np.zeros(64, dtype=np.uint8))
    u = np.uint64
    with pytest.raises(EclError):  # unaligned second text
        Forest(d, np.array([0, 8], dtype=u), np.array([8, 8], dtype=u), np.array([4, 4], dtype=u))
    with pytest.raises(EclError):  # empty text
        Forest(d, np.array([0, 16], dtype=u), np.array([16, 0], dtype=u), np.array([4, 4], dtype=u))
    with pytest.raises(EclError):  # alphabet too large
        Forest(d, np.array([0], dtype=u), np.array([64], dtype=u), np.array([300], dtype=u))


@pytest.mark.parametrize("span", [2, 5])
def test_forest_build_spans(gpu, span):
    """Workgroups spanning several tiles (and so several trees) build every
    tree exactly."""
    from ecl_wavelet_amd import set_build_span
    texts, sigmas = _padded_texts(16 + span, 50, lo_log2=6, hi_log2=17)
    set_build_span(span)
    try:
        f, _, _ = _forest(texts, sigmas)
    finally:
        set_build_span(0)
    q, e = _queries(texts, sigmas, 300, seed=9)
    assert np.array_equal(_run(f, q), e)


def test_forest_query_lines(gpu):
    """With query lines and select hints built for every tree (three
    launches over all levels), a mixed batch answers like the oracle, and a
    graph captured before the lines existed replays with the same answers."""
    from ecl_wavelet_amd import DeviceBuffer, synchronize
    texts, sigmas = _padded_texts(21, 40, lo_log2=4, hi_log2=18)
    f, _, _ = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 400, seed=9)
    dq = DeviceBuffer.from_array(q)
    dout = DeviceBuffer(len(q) * 8)
    g = f.graph(dq, len(q), dout)
    assert np.array_equal(_run(f, q), e)
    f.buildQueryLines()
    f.buildQueryLines()  # a second call keeps the lines
    assert np.array_equal(_run(f, q), e)
    zeros = np.zeros(len(q), dtype=np.uint64)
    for _ in range(2):
        dout.upload(zeros)
        synchronize()
        g.launch()
        synchronize()
        assert np.array_equal(dout.download(np.uint64, len(q)), e)


def test_forest_select_lines(gpu):
    """Select lines for every tree in one launch (alone, then with the query
    lines and hints): mixed batches answer like the oracle, including
    texts of one symbol (a level with no ones or no zeros) and texts shorter
    than a 512-bit block, and a graph captured before the lines existed
    replays with the same answers.  Every select of every tree's symbols
    at its first, last and one-past-last occurrence is checked too."""
    from ecl_wavelet_amd import DeviceBuffer, FOREST_QUERY_DTYPE, synchronize
    texts, sigmas = _padded_texts(41, 50, lo_log2=4, hi_log2=19)
    texts.append(np.full(3000, 3, dtype=np.uint8))  # one symbol of 4
    sigmas.append(4)
    texts.append(np.array([1, 0, 1], dtype=np.uint8))
    sigmas.append(4)
    f, _, _ = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 400, seed=41)
    # the ends of every symbol's occurrences
    extra, eexp = [], []
    for ti, (t, s) in enumerate(zip(texts, sigmas)):
        cnt = np.bincount(t, minlength=s)
        for c in range(s):
            for k in {1, int(cnt[c]), int(cnt[c]) + 1}:
                if k == 0:
                    continue
                extra.append((2, ti, k, c))
                eexp.append(O.scan_select_batch(t, np.array([k], dtype=np.uint64),
                                                np.array([c], dtype=np.uint64))[0])
    xq = np.zeros(len(extra), dtype=FOREST_QUERY_DTYPE)
    xa = np.array(extra, dtype=np.uint64)
    xq["kind"], xq["tree"], xq["arg"], xq["symbol"] = xa[:, 0], xa[:, 1], xa[:, 2], xa[:, 3]
    q = np.concatenate([q, xq])
    e = np.concatenate([e, np.array(eexp, dtype=np.uint64)])
    dq = DeviceBuffer.from_array(q)
    dout = DeviceBuffer(len(q) * 8)
    g = f.graph(dq, len(q), dout)
    f.buildSelectLines()
    f.buildSelectLines()  # a second call keeps the lines
    assert np.array_equal(_run(f, q), e)
    f.buildQueryLines()
    assert np.array_equal(_run(f, q), e)
    zeros = np.zeros(len(q), dtype=np.uint64)
    for _ in range(2):
        dout.upload(zeros)
        synchronize()
        g.launch()
        synchronize()
        assert np.array_equal(dout.download(np.uint64, len(q)), e)


def _select_ends(texts, sigmas):
    """Selects of every symbol of every tree at its first, last and
    one-past-last occurrence, with their direct-scan answers."""
    from ecl_wavelet_amd import FOREST_QUERY_DTYPE
    extra, exp = [], []
    for ti, (t, s) in enumerate(zip(texts, sigmas)):
        cnt = np.bincount(t, minlength=s)
        for c in range(s):
            for k in sorted({1, int(cnt[c]), int(cnt[c]) + 1} - {0}):
                extra.append((2, ti, k, c))
    xa = np.array(extra, dtype=np.uint64)
    q = np.zeros(len(extra), dtype=FOREST_QUERY_DTYPE)
    q["kind"], q["tree"], q["arg"], q["symbol"] = xa[:, 0], xa[:, 1], xa[:, 2], xa[:, 3]
    for ti, t in enumerate(texts):
        sel = xa[:, 1] == ti
        exp.append((np.nonzero(sel)[0], O.scan_select_batch(t, xa[sel, 2], xa[sel, 3])))
    e = np.zeros(len(extra), dtype=np.uint64)
    for idx, v in exp:
        e[idx] = v
    return q, e


def test_forest_quad_layouts(gpu):
    """Quad lines and quad select lines for the trees over 4, 16, 64 and 256
    symbols (three launches over all their levels), hex lines for those over
    16 and 256 (one more launch); the trees over 20
    symbols keep their binary levels in the same pair-lane kernel.  Mixed
    batches and every symbol's first, last and one-past-last select answer
    like the oracle: texts of one symbol, of 3 symbols, of exactly one and
    one more than one quad line (448 and 449), skewed ones.  A graph captured
    before the call keeps the binary kernel, one captured after runs the
    pair-lane kernel; both replay with the oracle's answers.  A build-graph
    replay over new texts drops the layouts of the rebuilt trees (answers
    stay right), and a second call builds them again."""
    from ecl_wavelet_amd import DeviceBuffer, synchronize
    texts, sigmas = _padded_texts(51, 40, lo_log2=4, hi_log2=18)
    rng = np.random.default_rng(52)
    for n, s in ((3000, 4), (3, 4), (448, 16), (449, 64), (20000, 16), (70000, 256)):
        if n == 3000:
            t = np.full(n, 3, dtype=np.uint8)
        elif n == 3:
            t = np.array([1, 0, 1], dtype=np.uint8)
        elif s == 256:
            t = O.gen_cdf(53, n, O.zipf_thresholds(s, 1.5), np.uint8).astype(np.uint8)
        else:
            t = rng.integers(0, s, n).astype(np.uint8)
        texts.append(t)
        sigmas.append(s)
    f, d, starts = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 300, seed=51)
    xq, xe = _select_ends(texts, sigmas)
    q, e = np.concatenate([q, xq]), np.concatenate([e, xe])
    dq = DeviceBuffer.from_array(q)
    dout = DeviceBuffer(len(q) * 8)
    g_before = f.graph(dq, len(q), dout)
    f.buildQuadLayouts()
    f.buildQuadLayouts()  # a second call keeps them
    assert np.array_equal(_run(f, q), e)
    block, grid = f.autotune(dq, len(q), dout)  # tunes the pair-lane kernel's grid
    assert block == 256 and grid > 0
    assert np.array_equal(_run(f, q), e)
    f.buildQueryLines()
    f.buildSelectLines()
    assert np.array_equal(_run(f, q), e)
    g_after = f.graph(dq, len(q), dout)
    zeros = np.zeros(len(q), dtype=np.uint64)
    for g in (g_before, g_after, g_after):
        dout.upload(zeros)
        synchronize()
        g.launch()
        synchronize()
        assert np.array_equal(dout.download(np.uint64, len(q)), e)
    # new texts, rebuilt by a build-graph replay: the layouts are dropped
    gb = f.buildGraph()
    new = [rng.integers(0, s, len(t)).astype(np.uint8) for t, s in zip(texts, sigmas)]
    _rewrite(d, starts, new)
    synchronize()
    gb.launch()
    synchronize()
    q2, e2 = _queries(new, sigmas, 300, seed=54)
    xq2, xe2 = _select_ends(new, sigmas)
    q2, e2 = np.concatenate([q2, xq2]), np.concatenate([e2, xe2])
    assert np.array_equal(_run(f, q2), e2)
    f.buildQuadLayouts()
    assert np.array_equal(_run(f, q2), e2)
    del g_before, g_after, gb


def _rewrite(d, starts, texts):
    """The texts written in place into the forest's device buffer."""
    for t, s0 in zip(texts, starts):
        dv = d.view(s0, len(t))
        dv.upload(t)


def test_forest_build_graph_rebuilds_from_new_texts(gpu):
    """The build captured as one graph: every replay rebuilds all trees from
    the texts the buffer holds at that moment (same lengths and alphabets,
    new symbols, so the non-power-of-two level sizes change), and the trees
    then answer like the oracle over those texts.  Eager work between the
    replays; query lines built before a replay are dropped by it and built
    again after."""
    from ecl_wavelet_amd import synchronize
    texts, sigmas = _padded_texts(31, 45, lo_log2=5, hi_log2=17)
    f, d, starts = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 200, seed=31)
    assert np.array_equal(_run(f, q), e)
    g = f.buildGraph()
    rng = np.random.default_rng(32)
    for rep in range(3):
        new = []
        for t, s in zip(texts, sigmas):
            if rep == 1:  # skewed: some symbols absent, level sizes move
                x = O.gen_cdf(100 + rep, len(t), O.zipf_thresholds(s, 2.0), np.uint8)
            else:
                x = rng.integers(0, s, len(t)).astype(np.uint8)
            new.append(x)
        _rewrite(d, starts, new)
        synchronize()
        if rep == 2:
            f.buildQueryLines()
        g.launch()
        synchronize()
        q, e = _queries(new, sigmas, 200, seed=40 + rep)
        assert np.array_equal(_run(f, q), e), rep
        f.buildQueryLines()
        assert np.array_equal(_run(f, q), e), rep
    del g


def test_forest_autotune_build(gpu):
    """The build autotuner picks a tile size, leaves the forest built (same
    answers), and later build graphs use the chosen tile."""
    from ecl_wavelet_amd import synchronize
    texts, sigmas = _padded_texts(33, 30, lo_log2=8, hi_log2=18)
    f, d, starts = _forest(texts, sigmas)
    q, e = _queries(texts, sigmas, 200, seed=33)
    rows, ms = f.autotuneBuild()
    assert rows in (1, 2, 4) and ms > 0
    assert np.array_equal(_run(f, q), e)
    g = f.buildGraph()
    g.launch()
    synchronize()
    assert np.array_equal(_run(f, q), e)
