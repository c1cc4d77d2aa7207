"""The query server (continuous batching) on the GPU: every request's results
are the same when it runs alone and when it runs in a mix of concurrent
requests of different sizes and trees, and equal the direct scans of the
texts.  Admission by HBM budget, removal, the plain per-call mode and the
synthetic load driver are covered too."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _text(seed, n, sigma):
    return O.gen_uniform(seed, n, sigma, np.uint8)


def _requests(text, sigma, sizes, seed):
    """Mixed-kind requests (MIXED_DTYPE) with their exact answers."""
    from ecl_wavelet_amd import MIXED_DTYPE
    rng = np.random.default_rng(seed)
    n = len(text)
    cnt = np.bincount(text, minlength=256)
    out = []
    for m in sizes:
        q = np.zeros(m, dtype=MIXED_DTYPE)
        kinds = rng.integers(0, 3, m)
        q["kind"] = kinds
        e = np.zeros(m, dtype=np.uint64)
        a = kinds == 0
        idx = rng.integers(0, n, a.sum()).astype(np.uint64)
        q["arg"][a] = idx
        e[a] = text[idx]
        r = kinds == 1
        ri = rng.integers(0, n + 1, r.sum()).astype(np.uint64)
        rc = rng.integers(0, sigma, r.sum()).astype(np.uint64)
        q["arg"][r], q["symbol"][r] = ri, rc
        e[r] = O.scan_rank_batch(text, ri, rc)
        s = kinds == 2
        sc = rng.integers(0, sigma, s.sum()).astype(np.uint64)
        k = np.array([int(rng.integers(1, cnt[c] + 2)) for c in sc], dtype=np.uint64)
        q["arg"][s], q["symbol"][s] = k, sc
        e[s] = O.scan_select_batch(text, k, sc)
        out.append((q, e))
    return out


@pytest.mark.parametrize("max_batch", [0, 1000])
def test_server_alone_equals_mix(gpu, max_batch):
    """max_batch 1000 forces requests to be split across batches."""
    from ecl_wavelet_amd import Server
    srv = Server(max_batch=max_batch, max_inflight=3)
    texts = [(_text(1, 50_000, 4), 4), (_text(2, 120_000, 256), 256), (_text(3, 70_000, 64), 64)]
    trees = [srv.addTree(t, s) for t, s in texts]
    rng = np.random.default_rng(9)
    work = []
    for j in range(40):
        ti = int(rng.integers(0, len(trees)))
        size = int(2 ** rng.uniform(0, 12))
        (q, e), = _requests(texts[ti][0], texts[ti][1], [size], seed=100 + j)
        work.append((trees[ti], q, e))
    # alone: each request waits for the previous one
    alone = []
    for tree, q, e in work:
        out, lat = srv.wait(srv.submit(tree, q))
        assert np.array_equal(out, e)
        assert lat >= 0
        alone.append(out)
    # in the mix: all submitted before any is waited for
    ids = [srv.submit(tree, q) for tree, q, _ in work]
    for rid, a, (_, _, e) in zip(ids, alone, work):
        out, _ = srv.wait(rid)
        assert np.array_equal(out, a)
        assert np.array_equal(out, e)
    st = srv.stats()
    assert st["requests"] == 80
    assert st["queries"] == 2 * sum(len(q) for _, q, _ in work)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert st["trees"] == 3
    srv.close()


def test_server_plain_mode_same_results(gpu):
    from ecl_wavelet_amd import Server
    srv = Server(coalesce=False)
    t = _text(4, 200_000, 20)
    tree = srv.addTree(t, 20)
    reqs = _requests(t, 20, [1, 17, 300, 5000], seed=5)
    ids = [srv.submit(tree, q) for q, _ in reqs]
    for rid, (_, e) in zip(ids, reqs):
        out, _ = srv.wait(rid)
        assert np.array_equal(out, e)
    assert srv.stats()["batches"] == 4  # one batch per request
    srv.close()


def test_server_budget_admission_and_removal(gpu):
    from ecl_wavelet_amd import Server
    from ecl_wavelet_amd._lib import EclError, ECL_ERROR_OUT_OF_MEMORY_TREE
    srv = Server(hbm_budget=3 << 20)  # room for about one tree of 2^21 symbols
    t = _text(6, 1 << 21, 256)
    tree = srv.addTree(t, 256)
    with pytest.raises(EclError) as ex:
        srv.addTree(t, 256)
    assert ex.value.status == ECL_ERROR_OUT_OF_MEMORY_TREE
    assert str(ex.value) == "Not enough memory available for the wavelet tree."
    (q, e), = _requests(t, 256, [1000], seed=3)
    rid = srv.submit(tree, q)
    srv.removeTree(tree)  # memory is returned once the request is done
    out, _ = srv.wait(rid)
    assert np.array_equal(out, e)
    with pytest.raises(EclError):
        srv.submit(tree, q)
    tree2 = srv.addTree(t, 256)  # fits again
    out, _ = srv.wait(srv.submit(tree2, q))
    assert np.array_equal(out, e)
    srv.close()


def test_server_device_text_and_load(gpu):
    from ecl_wavelet_amd import DeviceBuffer, Server
    srv = Server()
    t = _text(7, 1 << 20, 256)
    d = DeviceBuffer.from_array(t)
    tree = srv.addTree(d, 256, n=t.size)
    (q, e), = _requests(t, 256, [4096], seed=8)
    out, _ = srv.wait(srv.submit(tree, q))
    assert np.array_equal(out, e)
    st = srv.load(rate=2000, seconds=0.5, seed=3, min_size=1, max_size=20000, pool=1 << 16)
    assert st["requests"] > 0 and st["queries"] > 0
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_server_warm_start_from_files(gpu, tmp_path):
    """Trees saved to files are admitted without rebuilding and answer as
    built from their texts; files the server cannot hold are rejected and
    leave the budget untouched."""
    from ecl_wavelet_amd import Server, WaveletTree
    from ecl_wavelet_amd._lib import EclError
    texts = [(_text(21, 300_000, 256), 256), (_text(22, 90_000, 4), 4)]
    paths = []
    for j, (t, sg) in enumerate(texts):
        p = tmp_path / f"t{j}.ecl"
        WaveletTree(t, np.arange(sg, dtype=np.uint8)).save(p)
        paths.append(p)
    bad_alpha = tmp_path / "alpha.ecl"
    WaveletTree(np.array([1, 3, 5, 3] * 100, dtype=np.uint8),
                np.array([1, 3, 5], dtype=np.uint8)).save(bad_alpha)
    bad_type = tmp_path / "u16.ecl"
    WaveletTree(np.arange(1000, dtype=np.uint16) % 300, np.arange(300, dtype=np.uint16)).save(bad_type)
    srv = Server()
    trees = [srv.addTreeFile(p) for p in paths]
    used = srv.stats()["hbm_used"]
    for p in (bad_alpha, bad_type, tmp_path / "missing.ecl"):
        with pytest.raises(EclError):
            srv.addTreeFile(p)
    assert srv.stats()["hbm_used"] == used
    for tree, (t, sg) in zip(trees, texts):
        (q, e), = _requests(t, sg, [3000], seed=31)
        out, _ = srv.wait(srv.submit(tree, q))
        assert np.array_equal(out, e)
    srv.close()


def test_server_query_lines(gpu, tmp_path):
    """A server configured for query lines builds them for every admitted
    tree (from a text and from a saved file) and answers every request like
    the direct scans."""
    from ecl_wavelet_amd import Server, WaveletTree
    t0, t1 = _text(31, 300_001, 256), _text(32, 150_000, 4)
    WaveletTree(t1, np.arange(4, dtype=np.uint8)).save(tmp_path / "t1.ewt")
    srv = Server(query_lines=True)
    trees = [srv.addTree(t0, 256), srv.addTreeFile(tmp_path / "t1.ewt")]
    work = [(trees[0], *r) for r in _requests(t0, 256, [1, 999, 20000], seed=41)]
    work += [(trees[1], *r) for r in _requests(t1, 4, [7, 3000], seed=42)]
    ids = [srv.submit(tree, q) for tree, q, _ in work]
    for rid, (_, _, e) in zip(ids, work):
        out, _ = srv.wait(rid)
        assert np.array_equal(out, e)
    srv.close()


@pytest.mark.parametrize("max_batch", [0, 1000])
def test_server_quad_layouts(gpu, tmp_path, max_batch):
    """A server configured for the quad layouts (query_lines 3) gives every
    tree over a power-of-two alphabet with an even number of levels its quad
    lines and quad select lines, and hex lines where the level count is a
    multiple of 4 (sigma 16 and 256 here); the other trees (sigma 20 and 8)
    get query and select lines.  Batches then mix both kinds of tree, with a
    pair of lanes per query; every request equals the direct scans alone and
    in the mix, from texts and from a saved file, on a skewed text too."""
    from ecl_wavelet_amd import Server, WaveletTree
    rng = np.random.default_rng(77)
    skew = np.minimum(rng.zipf(1.3, 250_003) - 1, 255).astype(np.uint8)
    texts = [(_text(51, 300_001, 256), 256), (_text(52, 150_000, 4), 4),
             (_text(53, 90_001, 16), 16), (_text(54, 80_000, 20), 20),
             (_text(55, 60_000, 64), 64), (skew, 256), (_text(56, 40_000, 8), 8)]
    WaveletTree(texts[1][0], np.arange(4, dtype=np.uint8)).save(tmp_path / "t1.ewt")
    srv = Server(max_batch=max_batch, query_lines=3)
    trees = [srv.addTree(t, sg) if i != 1 else srv.addTreeFile(tmp_path / "t1.ewt")
             for i, (t, sg) in enumerate(texts)]
    work = []
    for j in range(36):
        ti = int(rng.integers(0, len(trees)))
        size = int(2 ** rng.uniform(0, 13))
        (q, e), = _requests(texts[ti][0], texts[ti][1], [size], seed=300 + j)
        work.append((trees[ti], q, e))
    alone = []
    for tree, q, e in work:
        out, _ = srv.wait(srv.submit(tree, q))
        assert np.array_equal(out, e)
        alone.append(out)
    ids = [srv.submit(tree, q) for tree, q, _ in work]
    for rid, a, (_, _, e) in zip(ids, alone, work):
        out, _ = srv.wait(rid)
        assert np.array_equal(out, a)
    srv.close()


def test_server_tree_from_raw_text_file(gpu, tmp_path):
    """A raw uint8 text file over a sparse alphabet is admitted after its
    conversion to the minimal alphabet (convertDataToMinAlphabet,
    test_benchmark_utils.cuh:628-679); requests over the minimal symbols
    answer like the direct scans of the converted text."""
    from ecl_wavelet_amd import Server
    rng = np.random.default_rng(21)
    symbols = np.array([3, 50, 51, 200, 255], np.uint8)
    raw = symbols[rng.integers(0, 5, 300_000)]
    path = tmp_path / "raw.u8"
    raw.tofile(path)
    srv = Server(max_inflight=2)
    tree, alpha = srv.addTreeTextFile(path, 250_000)
    assert np.array_equal(alpha, np.unique(raw[:250_000]))
    text = np.searchsorted(alpha, raw[:250_000]).astype(np.uint8)
    for q, e in _requests(text, len(alpha), [1, 700, 5000], seed=3):
        out, _ = srv.wait(srv.submit(tree, q))
        assert np.array_equal(out, e)
    srv.close()
