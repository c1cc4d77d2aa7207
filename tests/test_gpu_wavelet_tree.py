"""WaveletTree<T> on the GPU vs the oracle.  Mirrors the reference's
test_wavelet_tree.cu: structure KATs, codes, node ranks, histogram, and
randomized access/rank/select against direct scans of the text."""
import gc

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

DT = [np.uint8, np.uint16, np.uint32, np.uint64]


def _bits(rs, level):
    return rs.bit_array_.bits(level)


def _check_structure_vs_oracle(wt, sym, sigma):
    """Every level bit array, level size, count and node rank equals the
    oracle's tree over the same minimal-alphabet text."""
    ot = O.OracleTree(sym, sigma)
    rs = wt.rank_select_
    sizes = ot.level_sizes()
    assert wt.getNumLevels() == ot.num_levels
    for l in range(ot.num_levels):
        assert rs.bit_array_.size(l) == sizes[l]
        if sizes[l]:
            assert np.array_equal(rs.bit_array_.words(l), ot.level_words(l)), l
    assert np.array_equal(wt.getCounts(), ot.counts())
    r, s, lv = wt.nodeRanks()
    os_, olv = ot.node_starts()
    assert np.array_equal(s, os_) and np.array_equal(lv, olv)
    assert np.array_equal(r, ot.ranks())
    assert wt.getCodesStart() == ot.codes_start
    if ot.num_levels > 1:
        assert np.array_equal(wt.numNodesAtLevel(), ot.nodes_at_level()) or \
            O.is_pow2(sigma)
    return ot


def _check_queries(wt, text, sym_of, nq=3000, seed=0):
    """access/rank/select through the host API vs direct scans (exact)."""
    rng = np.random.default_rng(seed)
    n = len(text)
    alphabet = wt.getAlphabet()
    idx = rng.integers(0, n, nq).astype(np.uint64)
    got = wt.access(idx)
    assert np.array_equal(got, text[idx])
    # rank: positions in [0, n], symbols drawn from the alphabet
    ri = rng.integers(0, n + 1, nq).astype(np.uint64)
    rc = alphabet[rng.integers(0, len(alphabet), nq)]
    exp = O.scan_rank_batch(text, np.minimum(ri, n), rc.astype(np.uint64))
    got = wt.rank(wt.make_queries(ri, rc))
    assert np.array_equal(got, exp)
    # select: k in [1, count(c) + 1] (one past the last occurrence gives n)
    sc = alphabet[rng.integers(0, len(alphabet), nq)]
    hist = {int(a): int(c) for a, c in zip(*np.unique(text, return_counts=True))}
    k = np.array([int(rng.integers(1, hist.get(int(c), 0) + 2)) for c in sc],
                 dtype=np.uint64)
    exp = O.scan_select_batch(text, k, sc.astype(np.uint64))
    got = wt.select(wt.make_queries(k, sc))
    assert np.array_equal(got, exp)


# --- construction ------------------------------------------------------------
@pytest.mark.parametrize("dt", DT)
def test_constructor_small(gpu, dt):
    from ecl_wavelet_amd import WaveletTree
    WaveletTree(np.array([1, 2, 3, 4, 5, 6, 7, 8, 9], dtype=dt),
                np.arange(1, 10, dtype=dt))
    WaveletTree(np.array([0, 1, 2, 3], dtype=dt), np.arange(4, dtype=dt))
    WaveletTree(np.array([0, 1, 2], dtype=dt), np.arange(3, dtype=dt))


@pytest.mark.parametrize("dt", DT)
def test_structure_sigma10(gpu, dt):
    from ecl_wavelet_amd import WaveletTree
    data = (np.arange(100) % 10).astype(dt)
    wt = WaveletTree(data, np.arange(10, dtype=dt))
    rs = wt.rank_select_
    b = _bits(rs, 0)
    assert list(b) == [1 if i % 10 >= 8 else 0 for i in range(100)]
    b = _bits(rs, 1)
    assert list(b[:80]) == [1 if i % 8 >= 4 else 0 for i in range(80)]
    assert list(b[80:100]) == [i % 2 for i in range(80, 100)]
    assert list(_bits(rs, 2)[:80]) == [1 if i % 4 >= 2 else 0 for i in range(80)]
    assert list(_bits(rs, 3)[:80]) == [i % 2 for i in range(80)]


@pytest.mark.parametrize("dt", DT)
def test_structure_sigma5(gpu, dt):
    from ecl_wavelet_amd import WaveletTree
    data = np.array([4, 1, 1, 2, 1, 1, 3, 4, 4, 0, 0, 2, 3, 2, 4, 2, 4, 1, 3, 2], dtype=dt)
    wt = WaveletTree(data, np.arange(5, dtype=dt))
    rs = wt.rank_select_
    assert "".join(map(str, _bits(rs, 0))) == "10000001100000101000"
    assert "".join(map(str, _bits(rs, 1))) == "001001001111011"
    assert "".join(map(str, _bits(rs, 2))) == "111100101010010"


@pytest.mark.parametrize("sigma", [72, 63, 75, 20, 5, 10, 256, 1000])
def test_codes_vs_oracle(gpu, sigma):
    from ecl_wavelet_amd import WaveletTree
    dt = np.uint8 if sigma <= 256 else np.uint16
    wt = WaveletTree((np.arange(sigma * 2) % sigma).astype(dt), np.arange(sigma, dtype=dt))
    code, ln, total = O.minimal_codes(sigma)
    codes, lens = wt.codes()
    assert wt.getCodesStart() == sigma - total
    assert np.array_equal(codes, code[sigma - total:])
    assert np.array_equal(lens, ln[sigma - total:])


@pytest.mark.parametrize("sigma", [3, 4, 5, 10, 17, 64, 75, 100, 255, 256])
def test_structure_random_vs_oracle(gpu, sigma):
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(sigma)
    n = int(rng.integers(1000, 300000))
    text = rng.integers(0, sigma, n).astype(np.uint8)
    wt = WaveletTree(text, np.arange(sigma, dtype=np.uint8))
    _check_structure_vs_oracle(wt, text, sigma)


def test_histogram_kats(gpu):
    """TestGlobalHistogram (test_wavelet_tree.cu:261-379): counts of four
    fixed texts over the alphabet 0..9."""
    from ecl_wavelet_amd import WaveletTree
    alphabet = np.arange(10, dtype=np.uint8)
    texts = [np.arange(1000) % 10]
    t = np.zeros(1000, dtype=np.int64)
    t[:10] = np.arange(10)
    t[10:] = 9
    texts.append(t)
    t = np.arange(1000) % 8
    t[:5] = 9
    t[5:15] = 8
    texts.append(t)
    for text in texts:
        text = text.astype(np.uint8)
        wt = WaveletTree(text, alphabet)
        hist = np.bincount(text, minlength=10)
        exp = np.concatenate([[0], np.cumsum(hist)[:-1]]).astype(np.uint64)
        assert np.array_equal(wt.getCounts(), exp)
        for c in range(10):
            assert wt.getTotalAppearances(c) == hist[c]


# --- queries -----------------------------------------------------------------
@pytest.mark.parametrize("dt", DT)
def test_queries_kat_mod10(gpu, dt):
    from ecl_wavelet_amd import WaveletTree
    data = (np.arange(100) % 10).astype(dt)
    wt = WaveletTree(data, np.arange(10, dtype=dt))
    for i in range(1, 100):  # every prefix length as a batch (reference loop)
        idx = np.arange(i, dtype=np.uint64)
        assert np.array_equal(wt.access(idx), data[:i])
    j = np.arange(100)
    q = wt.make_queries(j, data)
    assert list(wt.rank(q)) == [x // 10 for x in j]
    q = wt.make_queries(j // 10 + 1, data)
    assert list(wt.select(q)) == list(j)
    assert list(wt.select(wt.make_queries([11], [0]))) == [100]


def test_queries_kat_mod256(gpu):
    from ecl_wavelet_amd import WaveletTree
    data = (np.arange(1280) % 256).astype(np.uint8)
    wt = WaveletTree(data, np.arange(256, dtype=np.uint8))
    j = np.arange(1280)
    assert np.array_equal(wt.access(j), data)
    assert list(wt.rank(wt.make_queries(j, data))) == list(j // 256)
    assert list(wt.select(wt.make_queries(j // 256 + 1, data))) == list(j)
    assert list(wt.select(wt.make_queries([11], [0]))) == [1280]


@pytest.mark.parametrize("sigma", [2, 3, 4, 5, 7, 16, 20, 32, 64, 100, 128, 200, 256])
def test_queries_random_u8(gpu, sigma):
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(1000 + sigma)
    n = int(rng.integers(10**5, 3 * 10**6))
    text = rng.integers(0, sigma, n).astype(np.uint8)
    wt = WaveletTree(text, np.arange(sigma, dtype=np.uint8))
    _check_queries(wt, text, None, seed=sigma)


@pytest.mark.parametrize("sigma", [300, 4097, 65536])
def test_queries_random_u16(gpu, sigma):
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(2000 + sigma)
    n = 2 * 10**6
    text = rng.integers(0, sigma, n).astype(np.uint16)
    wt = WaveletTree(text, np.arange(sigma, dtype=np.uint16))
    _check_queries(wt, text, None, nq=2000, seed=sigma)


def test_queries_zipf_u16_full_alphabet(gpu):
    """sigma = 2^16 with a Zipf(1.1) text: many tiny and empty nodes."""
    from ecl_wavelet_amd import WaveletTree
    thr = O.zipf_thresholds(65536, 1.1)
    text = O.gen_cdf(8, 3 * 10**6, thr, np.uint16)
    wt = WaveletTree(text, np.arange(65536, dtype=np.uint16))
    _check_queries(wt, text, None, nq=2000, seed=3)


@pytest.mark.parametrize("dt", DT)
def test_queries_random_alphabet(gpu, dt):
    """A random sorted, non-minimal alphabet (generateRandomAlphabet): the
    tree remaps symbols, access returns original symbols."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(np.dtype(dt).itemsize)
    info = np.iinfo(dt)
    for sigma in (3, 50, 200):
        alphabet = np.unique(rng.integers(0, int(info.max), 3 * sigma, dtype=np.uint64)
                             .astype(dt))[:sigma]
        text = alphabet[rng.integers(0, len(alphabet), 200000)]
        wt = WaveletTree(text, alphabet)
        assert not wt.isMinAlphabet()
        assert np.array_equal(wt.getAlphabet(), alphabet)
        _check_queries(wt, text, None, nq=1500, seed=sigma)


@pytest.mark.parametrize("dt", DT)
def test_alphabet_discovery(gpu, dt):
    """genAlphabetRandom: an empty alphabet is discovered from the text."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(5)
    sigma = 1000 if np.dtype(dt).itemsize > 1 else 200
    alphabet = np.arange(sigma, dtype=dt)
    text = np.concatenate([alphabet[rng.integers(0, sigma, 100000)], alphabet])
    wt = WaveletTree(text, None)
    assert np.array_equal(wt.getAlphabet(), alphabet)
    _check_queries(wt, text, None, nq=500, seed=1)


def test_device_resident_build_and_queries(gpu):
    """Text generated on the device (the bench's path) vs the oracle's copy of
    the same text; device-resident query kernels and the mixed batch."""
    from ecl_wavelet_amd import DeviceBuffer, WaveletTree, generate_text
    n, sigma = 5 * 10**6 + 17, 256
    d_text = generate_text(n, sigma, np.uint8, seed=4)
    text = O.gen_uniform(4, n, sigma, np.uint8)
    assert np.array_equal(d_text.download(np.uint8, n), text)
    wt = WaveletTree(d_text, np.arange(sigma, dtype=np.uint8), dtype=np.uint8, n=n)
    _check_structure_vs_oracle(wt, text, sigma) if n < 10**6 else None
    m = 30000
    d_q = DeviceBuffer(m * 16)
    d_out = DeviceBuffer(m * 8)
    # access
    wt.generateQueries(0, 9, m, d_q)
    idx = d_q.download(np.uint64, m)
    wt.accessDevice(d_q, m, d_out)
    assert np.array_equal(d_out.download(np.uint8, m), text[idx])
    # rank
    wt.generateQueries(1, 10, m, d_q)
    q = d_q.download(np.dtype([("index", "<u8"), ("symbol", "<u8")]), m)
    wt.rankDevice(d_q, m, d_out)
    exp = O.scan_rank_batch(text, q["index"], q["symbol"] & 0xFF)
    assert np.array_equal(d_out.download(np.uint64, m), exp)
    # select
    wt.generateQueries(2, 11, m, d_q)
    q = d_q.download(np.dtype([("index", "<u8"), ("symbol", "<u8")]), m)
    wt.selectDevice(d_q, m, d_out)
    exp = O.scan_select_batch(text, q["index"], q["symbol"] & 0xFF)
    assert np.array_equal(d_out.download(np.uint64, m), exp)
    # mixed 1:1:1 in one launch == the three single-kind answers
    from ecl_wavelet_amd import MIXED_DTYPE
    d_mq = DeviceBuffer(m * 16)
    wt.generateQueries(3, 12, m, d_mq)
    mq = d_mq.download(MIXED_DTYPE, m)
    wt.queryBatchDevice(d_mq, m, d_out)
    got = d_out.download(np.uint64, m)
    kinds = mq["kind"]
    a = mq["arg"]
    s = mq["symbol"].astype(np.uint64)
    assert np.array_equal(got[kinds == 0], text[a[kinds == 0]].astype(np.uint64))
    assert np.array_equal(got[kinds == 1],
                          O.scan_rank_batch(text, a[kinds == 1], s[kinds == 1]))
    assert np.array_equal(got[kinds == 2],
                          O.scan_select_batch(text, a[kinds == 2], s[kinds == 2]))


def test_symbol_with_no_occurrence(gpu):
    from ecl_wavelet_amd import WaveletTree
    text = np.array([0, 1, 3, 3, 1, 0, 4, 4] * 100, dtype=np.uint8)
    wt = WaveletTree(text, np.arange(5, dtype=np.uint8))  # symbol 2 absent
    assert list(wt.rank(wt.make_queries([0, 400, 800], [2, 2, 2]))) == [0, 0, 0]
    assert list(wt.select(wt.make_queries([1], [2]))) == [800]


def test_oom_message(gpu):
    """Asking for more memory than the device has raises the reference's
    exact message (wavelet_tree.cuh:652-653)."""
    from ecl_wavelet_amd import DeviceBuffer, WaveletTree
    from ecl_wavelet_amd._lib import EclError, ECL_ERROR_OUT_OF_MEMORY_TREE
    d = DeviceBuffer(1024)
    with pytest.raises(EclError) as e:
        WaveletTree(d, np.arange(256, dtype=np.uint8), dtype=np.uint8, n=1 << 42)
    assert e.value.status == ECL_ERROR_OUT_OF_MEMORY_TREE
    assert str(e.value) == "Not enough memory available for the wavelet tree."


@pytest.mark.parametrize("case", ["u8_256", "u8_200", "u16_1000", "u16_65536", "dev_200"])
@pytest.mark.parametrize("where", ["host", "split"])
def test_build_stages_on_host_when_device_is_short(gpu, case, where):
    """With the device's free memory capped just above the resident part of a
    build (wavelet_tree.cuh:655-689: the text, its sorted copy and the sort
    scratch go to unified memory when they do not fit), the text copy and the
    code buffers go to pinned host memory -- all of them ("host"), or all but
    the text copy ("split") -- and the tree equals the oracle's; below the
    resident part the build fails with the reference's message."""
    from ecl_wavelet_amd import (DeviceBuffer, WaveletTree, release_cached_memory,
                                 set_device_memory_cap, tree_build_memory)
    from ecl_wavelet_amd._lib import EclError, ECL_ERROR_OUT_OF_MEMORY_TREE
    kind, sigma = case.split("_")
    sigma = int(sigma)
    dt = np.uint16 if kind == "u16" else np.uint8
    rng = np.random.default_rng(sigma + (7 if where == "split" else 0))
    n = 3 * 10**6 + int(rng.integers(0, 1000))
    text = rng.integers(0, sigma, n).astype(dt)
    on_device = kind == "dev"
    resident, staging = tree_build_memory(dt, n, sigma, host_text=not on_device)
    text_bytes = 0 if on_device else n * np.dtype(dt).itemsize
    d = DeviceBuffer.from_array(text) if on_device else None
    alphabet = np.arange(sigma, dtype=dt)

    def build():
        if on_device:
            return WaveletTree(d, alphabet, dtype=dt, n=n)
        return WaveletTree(text, alphabet)

    try:
        release_cached_memory()
        set_device_memory_cap(0, resident - 1)
        with pytest.raises(EclError) as e:
            build()
        assert e.value.status == ECL_ERROR_OUT_OF_MEMORY_TREE
        assert str(e.value) == "Not enough memory available for the wavelet tree."
        del e  # its traceback holds this frame (and the trees) in a cycle
        release_cached_memory()
        extra = (text_bytes if where == "split" else 0) + (1 << 20)
        set_device_memory_cap(0, resident + extra)
        wt = build()
    finally:
        set_device_memory_cap(0, 0)
    hb = wt.buildHostBytes()
    if case == "u8_256":
        # its own code: the text copy is the only staging buffer
        assert hb == (0 if where == "split" else n)
    else:
        assert 0 < hb <= staging - (text_bytes if where == "split" else 0)
    _check_structure_vs_oracle(wt, text.astype(np.uint64), sigma)
    _check_queries(wt, text, None, nq=2000, seed=sigma)
    # without the cap the same text builds on the device alone
    assert build().buildHostBytes() == 0


def test_query_oom_message(gpu):
    """Host queries whose device staging does not fit raise the reference's
    exact message for their kind (wavelet_tree.cuh:972-973, 1268-1269,
    1587-1588); with the memory back, the same calls answer."""
    from ecl_wavelet_amd import DeviceBuffer, WaveletTree, mem_info, release_cached_memory
    from ecl_wavelet_amd._lib import EclError, ECL_ERROR_OUT_OF_MEMORY_QUERIES
    rng = np.random.default_rng(3)
    text = rng.integers(0, 4, 100_000).astype(np.uint8)
    wt = WaveletTree(text, np.arange(4, dtype=np.uint8))
    m = 1 << 22  # one full staging chunk: 2 x 2^22 x (16 + 8) bytes on the device
    q = wt.make_queries(rng.integers(0, len(text), m), rng.integers(0, 4, m))
    idx = q["index"].copy()
    # objects of earlier tests that only the cycle collector frees would
    # otherwise give their device memory back in the middle of this test
    gc.collect()
    release_cached_memory()
    fillers = []
    try:
        # leave less than the 75 MB a chunk pair of access queries needs
        # (rank and select need 201 MB)
        take = 64 << 30
        while mem_info()[0] > (64 << 20) and take >= (8 << 20):
            want = min(take, mem_info()[0] - (48 << 20))
            try:
                fillers.append(DeviceBuffer(want))
            except EclError:
                take //= 2
        assert mem_info()[0] < (72 << 20)
        for kind, call in (("access", lambda: wt.access(idx)), ("rank", lambda: wt.rank(q)),
                           ("select", lambda: wt.select(q))):
            with pytest.raises(EclError) as e:
                call()
            assert e.value.status == ECL_ERROR_OUT_OF_MEMORY_QUERIES
            assert str(e.value) == f"Not enough GPU memory available for the {kind} queries."
    finally:
        for f in fillers:
            f.free()
    _check_queries(wt, text, None, nq=500, seed=4)


def test_scratch_cache_reuse_release_and_threads(gpu):
    """Builds reuse the per-device construction scratch (growing, shrinking,
    released in between); two threads building at once (one of them on a
    private scratch) both give exact trees."""
    import threading
    from ecl_wavelet_amd import WaveletTree, release_cached_memory
    rng = np.random.default_rng(77)
    cases = [(3 * 10**6, 256), (10**5, 4), (2 * 10**6, 20), (5 * 10**6, 64)]
    for j, (n, sigma) in enumerate(cases):
        text = rng.integers(0, sigma, n).astype(np.uint8)
        wt = WaveletTree(text, np.arange(sigma, dtype=np.uint8))
        _check_queries(wt, text, None, nq=1000, seed=j)
        if j == 1:
            release_cached_memory()
    texts = [rng.integers(0, s, n).astype(np.uint8) for n, s in cases]
    out, errs = {}, []

    def work(i):
        try:
            n, sigma = cases[i]
            out[i] = WaveletTree(texts[i], np.arange(sigma, dtype=np.uint8))
        except Exception as e:  # noqa: BLE001
            errs.append(e)
    ths = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errs
    for i, t in enumerate(texts):
        _check_queries(out[i], t, None, nq=1000, seed=10 + i)


@pytest.mark.parametrize("span", [2, 3, 8])
def test_build_spans_vs_oracle(gpu, span):
    """Construction workgroups that take several tiles in turn carry the
    runs' partial words from tile to tile: every level stays bit-exact
    (single-pass power-of-2 and not, Zipf texts with tiny nodes, multi-pass
    16-bit trees with node-aligned tiles)."""
    from ecl_wavelet_amd import WaveletTree, set_build_span
    rng = np.random.default_rng(300 + span)
    set_build_span(span)
    try:
        for sigma, n in [(256, 1_500_000), (20, 700_000), (5, 400_000), (2, 300_000)]:
            text = rng.integers(0, sigma, n).astype(np.uint8)
            wt = WaveletTree(text, np.arange(sigma, dtype=np.uint8))
            _check_structure_vs_oracle(wt, text, sigma)
        thr = O.zipf_thresholds(200, 1.3)
        text = O.gen_cdf(span, 900_000, thr, np.uint8)
        wt = WaveletTree(text, np.arange(200, dtype=np.uint8))
        _check_structure_vs_oracle(wt, text, 200)
        for sigma, n in [(1000, 600_000), (4097, 500_000)]:
            text = rng.integers(0, sigma, n).astype(np.uint16)
            wt = WaveletTree(text, np.arange(sigma, dtype=np.uint16))
            _check_structure_vs_oracle(wt, text, sigma)
            _check_queries(wt, text, None, nq=1000, seed=span)
    finally:
        set_build_span(0)


def test_host_text_staged_upload_and_chunked_host_queries(gpu):
    """A host text above the staging size (uploaded through the pinned
    double buffer with parallel host copies) builds the same levels as the
    same text already on the device; host query arrays spanning several
    pipeline chunks answer as the device-resident calls."""
    from ecl_wavelet_amd import DeviceBuffer, WaveletTree
    rng = np.random.default_rng(42)
    n = 300_000_000 + 12345
    text = rng.integers(0, 256, n, dtype=np.uint8)
    alpha = np.arange(256, dtype=np.uint8)
    wt_h = WaveletTree(text, alpha)
    d = DeviceBuffer.from_array(text)
    wt_d = WaveletTree(d, alpha, dtype=np.uint8, n=n)
    ba_h, ba_d = wt_h.rank_select_.bit_array_, wt_d.rank_select_.bit_array_
    for l in range(8):
        assert np.array_equal(ba_h.words(l), ba_d.words(l)), l
    m = 9_000_000
    d_q = DeviceBuffer(m * 16)
    d_out = DeviceBuffer(m * 8)
    for kind in (1, 2):
        wt_d.generateQueries(kind, 70 + kind, m, d_q)
        q = d_q.download(np.dtype([("index", "<u8"), ("symbol", "<u8")]), m)
        if kind == 1:
            wt_d.rankDevice(d_q, m, d_out, flags=1)
            host = wt_h.rank(q, flags=1)
        else:
            wt_d.selectDevice(d_q, m, d_out, flags=1)
            host = wt_h.select(q, flags=1)
        assert np.array_equal(host, d_out.download(np.uint64, m))
    idx = rng.integers(0, n, m).astype(np.uint64)
    assert np.array_equal(wt_h.access(idx, flags=2), text[idx])


@pytest.mark.parametrize("case", ["u8_256", "u8_20", "u16_1000", "u32_alpha", "u8_zipf"])
def test_save_load_round_trip(gpu, tmp_path, case):
    """A tree written to a file and read back has the same levels,
    directory, counts, node ranks and codes, and answers every query as the
    original (multi-pass and non-minimal alphabets included)."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(hash(case) % 1000)
    if case == "u8_256":
        text = rng.integers(0, 256, 2_000_000).astype(np.uint8)
        alpha = np.arange(256, dtype=np.uint8)
    elif case == "u8_20":
        text = rng.integers(0, 20, 700_000).astype(np.uint8)
        alpha = np.arange(20, dtype=np.uint8)
    elif case == "u16_1000":
        text = rng.integers(0, 1000, 600_000).astype(np.uint16)
        alpha = np.arange(1000, dtype=np.uint16)
    elif case == "u32_alpha":
        alpha = np.sort(rng.choice(1 << 31, 37, replace=False)).astype(np.uint32)
        text = alpha[rng.integers(0, 37, 300_000)]
    else:
        text = O.gen_cdf(5, 800_000, O.zipf_thresholds(200, 1.2), np.uint8)
        alpha = np.arange(200, dtype=np.uint8)
    wt = WaveletTree(text, alpha)
    path = tmp_path / "tree.ecl"
    wt.save(path)
    lt = WaveletTree.load(path)
    assert lt.dtype == wt.dtype
    assert lt.getNumLevels() == wt.getNumLevels()
    assert np.array_equal(lt.getAlphabet(), wt.getAlphabet())
    assert np.array_equal(lt.getCounts(), wt.getCounts())
    r0, s0, l0 = wt.nodeRanks()
    r1, s1, l1 = lt.nodeRanks()
    assert np.array_equal(r0, r1) and np.array_equal(s0, s1) and np.array_equal(l0, l1)
    rs0, rs1 = wt.rank_select_, lt.rank_select_
    for l in range(wt.getNumLevels()):
        assert np.array_equal(rs0.bit_array_.words(l), rs1.bit_array_.words(l))
        assert np.array_equal(rs0.l1Entries(l), rs1.l1Entries(l))
        assert np.array_equal(rs0.l2Entries(l), rs1.l2Entries(l))
        for v in (0, 1):
            assert np.array_equal(rs0.selectSamples(l, v), rs1.selectSamples(l, v))
    _check_queries(lt, text, None, nq=2000, seed=9)


def test_load_rejects_damaged_files(gpu, tmp_path):
    from ecl_wavelet_amd import WaveletTree
    from ecl_wavelet_amd._lib import EclError
    text = (np.arange(100_003) % 7).astype(np.uint8)
    wt = WaveletTree(text, np.arange(7, dtype=np.uint8))
    path = tmp_path / "t.ecl"
    wt.save(path)
    data = path.read_bytes()
    # layout: 40-byte header, 7 alphabet bytes, 7 u64 counts, then per level
    # a u64 size and its words.  Level 1 holds node 0..3 at [0, C(4)) and
    # node 4..6 at [C(4), n).
    n = len(text)
    lvl0 = 40 + 7 + 56 + 8
    lvl1 = lvl0 + (n + 31) // 32 * 4 + 8
    bits1 = np.unpackbits(np.frombuffer(data[lvl1:lvl1 + (n + 31) // 32 * 4], np.uint8),
                          bitorder="little")[:n]
    c4 = int(np.count_nonzero(text < 4))
    one_left = int(np.flatnonzero(bits1[:c4] == 1)[0])
    zero_right = c4 + int(np.flatnonzero(bits1[c4:] == 0)[0])

    def flip(blob, base, bit):
        b = bytearray(blob)
        b[base + bit // 8] ^= 1 << (bit % 8)
        return bytes(b)

    cases = {
        "truncated": data[:-5],
        "trailing": data + b"x",
        "magic": b"X" + data[1:],
        "counts": data[:40 + 7] + bytes([data[47] ^ 1]) + data[48:],  # first symbol count
        # one level-0 bit flipped: the level's ones no longer match the counts
        "bit": flip(data, lvl0, 12345),
        # a one moved from node 0..3 into node 4..6 of level 1: the level's
        # total is unchanged, but both nodes' ones are wrong, and a descent
        # into the right node could index past the last level's array
        "moved_bit": flip(flip(data, lvl1, one_left), lvl1, zero_right),
    }
    for name, blob in cases.items():
        p = tmp_path / f"{name}.ecl"
        p.write_bytes(blob)
        with pytest.raises(EclError, match="level bits do not match" if "bit" in name else None):
            WaveletTree.load(p)
    with pytest.raises(EclError):
        WaveletTree.load(tmp_path / "missing.ecl")
    # set bits past n in the last word of level 0 are not part of the array:
    # the file loads and answers as the intact one
    last = lvl0 + ((n + 31) // 32 - 1) * 4
    p = tmp_path / "pad.ecl"
    assert n % 32
    p.write_bytes(flip(data, last, 31))
    _check_queries(WaveletTree.load(p), text, None, nq=500, seed=2)
    lt = WaveletTree.load(path)  # the intact file still loads
    _check_queries(lt, text, None, nq=500, seed=1)


@pytest.mark.parametrize("case", ["u8_256", "u8_20", "u16_1000", "u32_alpha"])
def test_query_lines_same_answers(gpu, case):
    """With query lines (one 128-byte line per level and rank) every access,
    rank and select answer is the one without them, including positions at
    line boundaries (multiples of 960 and their neighbours)."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(len(case))
    if case == "u8_256":
        text = rng.integers(0, 256, 1_500_017).astype(np.uint8)
        alpha = np.arange(256, dtype=np.uint8)
    elif case == "u8_20":
        text = rng.integers(0, 20, 700_001).astype(np.uint8)
        alpha = np.arange(20, dtype=np.uint8)
    elif case == "u16_1000":
        text = rng.integers(0, 1000, 400_003).astype(np.uint16)
        alpha = np.arange(1000, dtype=np.uint16)
    else:
        alpha = np.sort(rng.choice(1 << 31, 29, replace=False)).astype(np.uint32)
        text = alpha[rng.integers(0, 29, 250_000)]
    wt = WaveletTree(text, alpha)
    n = len(text)
    idx = np.concatenate([rng.integers(0, n, 5000),
                          np.arange(0, n, 960), np.arange(959, n, 960),
                          np.arange(961, n, 960), [n - 1]]).astype(np.uint64)
    ri = np.concatenate([idx, [0, n]]).astype(np.uint64)
    rc = alpha[rng.integers(0, len(alpha), len(ri))]
    q = wt.make_queries(ri, rc)
    # select: every occurrence (and one past) of two symbols, plus random k
    syms, cnts = np.unique(text, return_counts=True)
    sk, sc = [], []
    for c, cnt in list(zip(syms, cnts))[:2]:
        sk.append(np.arange(1, cnt + 2, dtype=np.uint64))
        sc.append(np.full(cnt + 1, c, dtype=text.dtype))
    c3 = syms[rng.integers(0, len(syms), 4000)]
    sk.append(np.array([int(rng.integers(1, cnts[np.searchsorted(syms, c)] + 2)) for c in c3],
                       dtype=np.uint64))
    sc.append(c3.astype(text.dtype))
    sq = wt.make_queries(np.concatenate(sk), np.concatenate(sc))
    before = (wt.access(idx), wt.rank(q.copy()), wt.select(sq.copy()))
    wt.buildQueryLines()
    wt.buildQueryLines()  # a second call keeps the lines
    assert np.array_equal(wt.access(idx), before[0])
    assert np.array_equal(wt.rank(q.copy()), before[1])
    got = wt.select(sq.copy())
    assert np.array_equal(got, before[2])
    assert np.array_equal(got, O.scan_select_batch(text, np.concatenate(sk),
                                                   np.concatenate(sc).astype(np.uint64)))
    wt.buildSelectLines()
    wt.buildSelectLines()  # a second call keeps the lines
    assert np.array_equal(wt.select(sq.copy()), before[2])
    assert np.array_equal(wt.access(idx), before[0])
    assert np.array_equal(wt.access(idx), text[idx])
    _check_queries(wt, text, None, nq=3000, seed=4)


@pytest.mark.parametrize("lanes", [0, 1, 2])
@pytest.mark.parametrize("case", ["u8_256", "u16_1000", "u32_sparse", "tiny"])
def test_query_lanes(gpu, lanes, case):
    """access and rank with one or two lanes per query (the pair fetches half
    of each level's line per lane; 0 chooses two lanes over query lines) and
    select (one lane) equal the oracle's scans over the plain directory, over
    query lines with select hints, and with select lines too, including
    positions at block (512) and line (960) boundaries and every occurrence
    of one symbol."""
    from ecl_wavelet_amd import WaveletTree, set_query_lanes
    rng = np.random.default_rng(lanes * 7 + len(case))
    if case == "u8_256":
        text = rng.integers(0, 256, 900_001).astype(np.uint8)
    elif case == "u16_1000":
        text = rng.integers(0, 1000, 300_007).astype(np.uint16)
    elif case == "u32_sparse":
        alpha = np.sort(rng.choice(1 << 31, 37, replace=False)).astype(np.uint32)
        text = alpha[rng.integers(0, 37, 200_003)]
    else:
        text = np.array([3, 1, 4, 1, 5, 9, 2, 6, 5], dtype=np.uint8)
    n = len(text)
    set_query_lanes(lanes)
    try:
        wt = WaveletTree(text, np.unique(text))
        idx = np.concatenate([rng.integers(0, n, 4000), np.arange(0, n, 512),
                              np.arange(511, n, 512), np.arange(0, n, 960),
                              np.arange(959, n, 960), [0, n - 1]]).astype(np.uint64)
        ri = np.concatenate([idx, [0, n]]).astype(np.uint64)
        syms = np.unique(text)
        rc = syms[rng.integers(0, len(syms), len(ri))]
        exp_r = O.scan_rank_batch(text, np.minimum(ri, n), rc.astype(np.uint64))
        # select: every occurrence (and one past) of one symbol, plus random k
        cnt = {int(a): int(c) for a, c in zip(*np.unique(text, return_counts=True))}
        s0 = syms[0]
        sc = np.concatenate([np.full(cnt[int(s0)] + 1, s0), syms[rng.integers(0, len(syms), 3000)]])
        sk = np.concatenate([np.arange(1, cnt[int(s0)] + 2),
                             [int(rng.integers(1, cnt[int(c)] + 2)) for c in sc[cnt[int(s0)] + 1:]]])
        sc = sc.astype(text.dtype)
        sk = sk.astype(np.uint64)
        exp_s = O.scan_select_batch(text, sk, sc.astype(np.uint64))
        for stage in ("plain", "query_lines", "select_lines"):
            if stage == "query_lines":
                wt.buildQueryLines()
            elif stage == "select_lines":
                wt.buildSelectLines()
            assert np.array_equal(wt.access(idx), text[idx]), stage
            assert np.array_equal(wt.rank(wt.make_queries(ri, rc)), exp_r), stage
            assert np.array_equal(wt.select(wt.make_queries(sk, sc)), exp_s), stage
    finally:
        set_query_lanes(0)


def _skewed_text(rng, n):
    """Halves whose level-0 ones density is ~0.02 and ~0.98 (98% of the
    symbols below 128, then 98% at or above): the level's density is ~0.5,
    so its select lines (one every ~400 Values) overflow their 960 bits
    throughout, and the selects fall back to the hints and the directory."""
    half = n // 2
    lo = np.where(rng.random(half) < 0.98, rng.integers(0, 128, half), rng.integers(128, 256, half))
    hi = np.where(rng.random(n - half) < 0.98, rng.integers(128, 256, n - half),
                  rng.integers(0, 128, n - half))
    return np.concatenate([lo, hi]).astype(np.uint8)


@pytest.mark.parametrize("text_kind", ["uniform", "skewed"])
@pytest.mark.parametrize("m", [1, 3, 777, 70001])
def test_mixed_batch_query_lines(gpu, m, text_kind):
    """A mixed batch (kinds grouped per chunk; with query lines, access and
    rank on two lanes each) answers like the oracle over the directory, then
    with the query lines and hints, then with the select lines too, then
    with the quad lines and quad select lines (every query on a pair of
    lanes), and with hex lines beside them, for odd batch sizes and for balanced, one-sided and kind > 2
    (read as select) mixes; on a skewed text the select lines overflow and
    the mixed kernel's select goes from the line to the hint and to the
    directory search, and rare digits overflow their quad select lines."""
    from ecl_wavelet_amd import MIXED_DTYPE, DeviceBuffer, WaveletTree
    rng = np.random.default_rng(m)
    text = (rng.integers(0, 256, 300_007).astype(np.uint8) if text_kind == "uniform"
            else _skewed_text(rng, 300_007))
    n = len(text)
    wt = WaveletTree(text, np.arange(256, dtype=np.uint8))
    cnt = np.bincount(text, minlength=256)
    for mix in ("balanced", "access_rank", "select_high"):
        q = np.zeros(m, dtype=MIXED_DTYPE)
        if mix == "balanced":
            kinds = rng.integers(0, 3, m)
        elif mix == "access_rank":
            kinds = rng.integers(0, 2, m)
        else:
            kinds = rng.choice([0, 2, 7], m)
        sym = rng.integers(0, 256, m)
        arg = np.where(kinds == 0, rng.integers(0, n, m),
                       np.where(kinds == 1, rng.integers(0, n + 1, m),
                                [int(rng.integers(1, cnt[c] + 2)) for c in sym]))
        q["kind"], q["symbol"], q["arg"] = kinds, sym, arg
        exp = np.empty(m, dtype=np.uint64)
        a, s = q["arg"], q["symbol"].astype(np.uint64)
        k0, k1, k2 = kinds == 0, kinds == 1, kinds >= 2
        exp[k0] = text[a[k0]]
        exp[k1] = O.scan_rank_batch(text, a[k1], s[k1])
        exp[k2] = O.scan_select_batch(text, a[k2], s[k2])
        d_q, d_out = DeviceBuffer(m * 16), DeviceBuffer(m * 8)
        d_q.upload(q)
        for stage in ("directory", "query lines", "select lines", "quad lines", "hex lines"):
            if stage == "query lines":
                wt.buildQueryLines()
            elif stage == "select lines":
                wt.buildSelectLines()
            elif stage == "quad lines":
                wt.buildQuadSelectLines()  # with the quad lines: the pair-lane mixed kernel
            elif stage == "hex lines":
                wt.buildHexLines()  # mixed batches keep the quad layouts
            wt.queryBatchDevice(d_q, m, d_out)
            assert np.array_equal(d_out.download(np.uint64, m), exp), (mix, stage)


QUAD_CASES = ["u8_4", "u8_16", "u8_64", "u8_256", "u8_256_skewed", "u16_16_alpha",
              "u16_65536", "u8_256_n448k"]


def _quad_text(case, rng):
    """(text, alphabet) for a quad-lines case: power-of-two alphabets with an
    even number of levels, uniform or skewed (with absent symbols), minimal
    or not, n a multiple of 448 or not."""
    if case == "u8_256_skewed":
        text = np.minimum(rng.zipf(1.3, 1_000_003) - 1, 255).astype(np.uint8)
        text[text == 7] = 8  # symbol 7 absent
        return text, np.arange(256, dtype=np.uint8)
    if case == "u16_16_alpha":
        alpha = np.sort(rng.choice(60000, 16, replace=False)).astype(np.uint16)
        return alpha[rng.integers(0, 16, 777_777)], alpha
    if case == "u16_65536":
        return rng.integers(0, 65536, 2_000_011).astype(np.uint16), np.arange(65536, dtype=np.uint16)
    if case == "u8_256_n448k":
        return rng.integers(0, 256, 448 * 1000).astype(np.uint8), np.arange(256, dtype=np.uint8)
    if case == "u8_256_n128k":
        return rng.integers(0, 256, 128 * 7001).astype(np.uint8), np.arange(256, dtype=np.uint8)
    if case == "u16_4096":
        return rng.integers(0, 4096, 1_500_001).astype(np.uint16), np.arange(4096, dtype=np.uint16)
    if case == "u16_65536_skewed":
        text = np.minimum(rng.zipf(1.1, 1_000_001) - 1, 65535).astype(np.uint16)
        return text, np.arange(65536, dtype=np.uint16)
    sigma = int(case.split("_")[1])
    return rng.integers(0, sigma, 1_234_567).astype(np.uint8), np.arange(sigma, dtype=np.uint8)


@pytest.mark.parametrize("case", QUAD_CASES)
def test_quad_lines_same_answers(gpu, case):
    """Access and rank over the quad lines (2-bit digits of two levels per
    128-byte line of 448) equal the direct scans of the text: random
    queries, every line boundary (448 j - 1, 448 j, 448 j + 1), every node
    boundary of every quad level (the C array entries), positions 0, n - 1
    and n, and every symbol including absent ones; select keeps its own
    path and stays exact."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(sum(map(ord, case)))
    text, alpha = _quad_text(case, rng)
    n = len(text)
    wt = WaveletTree(text, alpha)
    wt.buildQuadLines()
    wt.buildQuadLines()  # a second call keeps the lines
    _check_queries(wt, text, None, nq=4000, seed=len(case))
    # access at line boundaries and node starts
    lb = np.arange(0, n + 448, 448, dtype=np.int64)
    pos = np.concatenate([lb - 1, lb, lb + 1, [0, n - 1], wt.getCounts().astype(np.int64)])
    pos = np.unique(pos[(pos >= 0) & (pos < n)]).astype(np.uint64)
    assert np.array_equal(wt.access(pos), text[pos])
    # rank at the same positions and n, for a spread of symbols (absent ones too)
    ri = np.unique(np.concatenate([pos, [n]])).astype(np.uint64)
    syms = alpha[rng.integers(0, len(alpha), 8)]
    for c in np.unique(np.concatenate([syms, alpha[:1], alpha[-1:]])):
        cs = np.full(len(ri), c, dtype=alpha.dtype)
        exp = O.scan_rank_batch(text, ri, cs.astype(np.uint64))
        assert np.array_equal(wt.rank(wt.make_queries(ri, cs)), exp), int(c)


@pytest.mark.parametrize("case", QUAD_CASES)
def test_quad_select_lines_same_answers(gpu, case):
    """Select over the quad select lines (a line every r occurrences of each
    digit of each quad level; an occurrence past a line's 448 digits is found
    by a walk over the quad lines) equals the text's occurrence positions:
    random queries, and every occurrence of the rarest and the most frequent
    symbol and of a random one (each line boundary, both ends), plus
    count + 1 (n).  The skewed text has digits rare enough that their lines
    overflow."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(sum(map(ord, case)) + 1)
    text, alpha = _quad_text(case, rng)
    n = len(text)
    wt = WaveletTree(text, alpha)
    wt.buildQuadSelectLines()  # builds the quad lines too
    wt.buildQuadSelectLines()
    _check_queries(wt, text, None, nq=4000, seed=len(case) + 1)
    vals, cnts = np.unique(text, return_counts=True)
    picks = {vals[np.argmin(cnts)], vals[np.argmax(cnts)], vals[rng.integers(0, len(vals))]}
    for c in picks:
        occ = np.flatnonzero(text == c).astype(np.uint64)
        if len(occ) > 300_000:
            occ = occ[:300_000]
        k = np.arange(1, len(occ) + 2, dtype=np.uint64)
        exp = np.concatenate([occ, [n]]).astype(np.uint64)
        if len(occ) == 300_000:
            k, exp = k[:-1], exp[:-1]
        cs = np.full(len(k), c, dtype=alpha.dtype)
        assert np.array_equal(wt.select(wt.make_queries(k, cs)), exp), int(c)


HEX_CASES = ["u8_16", "u8_256", "u8_256_skewed", "u16_16_alpha", "u16_4096", "u16_65536",
             "u16_65536_skewed", "u8_256_n128k"]


@pytest.mark.parametrize("case", HEX_CASES)
def test_hex_lines_same_answers(gpu, case):
    """Access and rank over the hex lines (4-bit digits of four levels per
    128-byte line of 128, built from the quad lines) equal the direct scans
    of the text: random queries, every line boundary (128 j - 1, 128 j,
    128 j + 1), every node boundary (the C array entries), positions 0,
    n - 1 and n, and a spread of symbols including absent ones.  Select,
    which keeps its own layouts, stays exact beside them.  The
    alphabets of 2^16 symbols put many node starts inside the 32-position
    words that the build handles together."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(sum(map(ord, case)) + 7)
    text, alpha = _quad_text(case, rng)
    n = len(text)
    wt = WaveletTree(text, alpha)
    wt.buildHexLines()  # builds the quad lines too
    wt.buildHexLines()  # a second call keeps the lines
    _check_queries(wt, text, None, nq=4000, seed=len(case) + 2)
    lb = np.arange(0, n + 128, 128, dtype=np.int64)
    pos = np.concatenate([lb - 1, lb, lb + 1, [0, n - 1], wt.getCounts().astype(np.int64)])
    pos = np.unique(pos[(pos >= 0) & (pos < n)]).astype(np.uint64)
    assert np.array_equal(wt.access(pos), text[pos])
    ri = np.unique(np.concatenate([pos, [n]])).astype(np.uint64)
    if len(ri) > 60_000:
        ri = np.unique(np.concatenate([ri[rng.integers(0, len(ri), 60_000)], [0, n]])).astype(np.uint64)
    syms = alpha[rng.integers(0, len(alpha), 6)]
    for c in np.unique(np.concatenate([syms, alpha[:1], alpha[-1:]])):
        cs = np.full(len(ri), c, dtype=alpha.dtype)
        exp = O.scan_rank_batch(text, ri, cs.astype(np.uint64))
        assert np.array_equal(wt.rank(wt.make_queries(ri, cs)), exp), int(c)
    # with the quad select lines as well: select over them, access and rank
    # still over the hex lines
    wt.buildQuadSelectLines()
    _check_queries(wt, text, None, nq=2000, seed=len(case) + 3)


@pytest.mark.parametrize("sigma", [4, 8, 64, 20])
def test_hex_lines_reject_other_trees(gpu, sigma):
    """Alphabets that are not a power of two, or whose level count is not a
    multiple of 4, get no hex lines, and their queries keep working."""
    from ecl_wavelet_amd import WaveletTree
    from ecl_wavelet_amd._lib import EclError, ECL_ERROR_INVALID_ARGUMENT
    rng = np.random.default_rng(sigma + 100)
    text = rng.integers(0, sigma, 100_003).astype(np.uint8)
    wt = WaveletTree(text, np.arange(sigma, dtype=np.uint8))
    with pytest.raises(EclError) as e:
        wt.buildHexLines()
    assert e.value.status == ECL_ERROR_INVALID_ARGUMENT
    del e
    _check_queries(wt, text, None, nq=500, seed=sigma)


@pytest.mark.parametrize("sigma", [2, 8, 20, 128])
def test_quad_lines_reject_other_trees(gpu, sigma):
    """Alphabets that are not a power of two, or whose level count is odd,
    get no quad lines, and their queries keep working."""
    from ecl_wavelet_amd import WaveletTree
    from ecl_wavelet_amd._lib import EclError, ECL_ERROR_INVALID_ARGUMENT
    rng = np.random.default_rng(sigma)
    text = rng.integers(0, sigma, 100_003).astype(np.uint8)
    wt = WaveletTree(text, np.arange(sigma, dtype=np.uint8))
    for build in (wt.buildQuadLines, wt.buildQuadSelectLines):
        with pytest.raises(EclError) as e:
            build()
        assert e.value.status == ECL_ERROR_INVALID_ARGUMENT
        del e
    _check_queries(wt, text, None, nq=500, seed=sigma)


def test_tree_from_raw_text_file(gpu, tmp_path):
    """A tree over a prefix of a raw text file (readDataFromFile,
    test_benchmark_utils.cuh:588-619) with its alphabet discovered equals the
    tree over the same text given in memory, and answers every query like
    the direct scans; a file that is too short raises the reference's
    message."""
    from ecl_wavelet_amd import WaveletTree
    from ecl_wavelet_amd._lib import EclError
    rng = np.random.default_rng(17)
    alpha = np.sort(rng.choice(60000, 300, replace=False)).astype(np.uint16)
    full = alpha[rng.integers(0, 300, 400_000)]
    path = tmp_path / "text.u16"
    full.tofile(path)
    n = 399_990
    wt = WaveletTree.fromFile(path, n, np.uint16)
    ref = WaveletTree(full[:n], np.unique(full[:n]))
    assert np.array_equal(wt.getAlphabet(), ref.getAlphabet())
    for l in range(wt.getNumLevels()):
        assert np.array_equal(wt.rank_select_.bit_array_.words(l),
                              ref.rank_select_.bit_array_.words(l))
    _check_queries(wt, full[:n], None, nq=1000, seed=5)
    with pytest.raises(EclError, match="Data size is larger than file size"):
        WaveletTree.fromFile(path, 400_001, np.uint16)


def test_tree_autotune_lanes(gpu):
    """The tree's query autotuner picks one or two lanes per access/rank
    query (with and without query lines) and keeps it for the tree; every
    setting answers like the direct scans."""
    from ecl_wavelet_amd import WaveletTree
    rng = np.random.default_rng(61)
    text = rng.integers(0, 256, 3_000_000).astype(np.uint8)
    wt = WaveletTree(text, np.arange(256, dtype=np.uint8))
    for lines in (False, True):
        if lines:
            wt.buildQueryLines()
        lanes, ms = wt.autotune(m=1 << 20, seed=3)
        assert lanes in (1, 2) and ms[0] > 0 and ms[1] > 0
        _check_queries(wt, text, None, nq=2000, seed=62 + lines)
        for forced in (1, 2, 0):
            wt.setQueryLanes(forced)
            _check_queries(wt, text, None, nq=1000, seed=70 + forced)
