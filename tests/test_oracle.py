"""Pins for the CPU oracle: known-answer vectors from the reference's own
tests, closed forms, and brute force.  No GPU needed."""
import numpy as np
import pytest

import oracle as O


# --- generator --------------------------------------------------------------
def test_hash_is_splitmix64():
    # splitmix64 seeded with 0: first output 0xE220A8397B1DCDAF, second
    # 0x6E789E6AA1B965F4 (published test vector of the generator).
    assert O.hash64(0, 0) == 0xE220A8397B1DCDAF
    assert O.hash64(0, 1) == 0x6E789E6AA1B965F4


def test_bounded_is_mulhi():
    rng = np.random.default_rng(0)
    for h in rng.integers(0, 2**63, 50, dtype=np.uint64):
        for m in (1, 4, 256, 65536, 6_000_000_000):
            assert O.lib().orc_bounded(int(h), m) == (int(h) * m) >> 64


def test_gen_uniform_matches_hash():
    t = O.gen_uniform(7, 1000, 20, np.uint8)
    for i in range(0, 1000, 37):
        assert t[i] == (O.hash64(7, i) * 20) >> 64


def test_gen_cdf_zipf():
    thr = O.zipf_thresholds(100, 1.1)
    assert thr[-1] == np.uint64(2**64 - 1)
    assert np.all(np.diff(thr.astype(object)) >= 0)
    t = O.gen_cdf(3, 20000, thr, np.uint16)
    h = np.bincount(t, minlength=100)
    assert h[0] > h[1] > h[10] > 0  # Zipf: decreasing frequencies
    for i in range(0, 20000, 997):
        x = O.hash64(3, i)
        assert t[i] == int(np.searchsorted(thr.astype(object), x, side="right"))


# --- utils (test_utils.cu:30-57) ---------------------------------------------
def test_prev_pow2_closed_form():
    expected = 0
    for i in range(1024):
        assert O.prev_pow2(i) == expected
        if i & (i - 1) == 0:
            expected = i


def test_ceil_log2_closed_form():
    expected = 0
    for i in range(1, 1024):
        assert O.ceil_log2(i) == expected
        if i & (i - 1) == 0:
            expected += 1


# --- minimal codes (test_wavelet_tree.cu:174-217) ----------------------------
def test_codes_pow2_empty():
    for s in (4, 8, 16, 32, 256, 65536):
        assert O.minimal_codes(s)[2] == 0


def test_codes_kat_72():
    code, ln, total = O.minimal_codes(72)
    assert total == 8
    for i in range(64, 72):
        assert code[i] == 64 + 8 * (i - 64) and ln[i] == 4


def test_codes_kat_63():
    code, ln, total = O.minimal_codes(63)
    assert code[62] == 62 and ln[62] == 5
    assert total == 1  # only symbol 62 keeps a code


def test_codes_kat_75():
    code, ln, total = O.minimal_codes(75)
    assert total == 11
    for i in range(64, 72):
        assert code[i] == 64 + 4 * (i - 64) and ln[i] == 5
    assert (code[72], ln[72]) == (96, 4)
    assert (code[73], ln[73]) == (104, 4)
    assert (code[74], ln[74]) == (112, 3)


def test_codes_sigma20():
    code, ln, total = O.minimal_codes(20)
    assert total == 4
    assert [(int(code[i]), int(ln[i])) for i in range(16, 20)] == [
        (16, 3), (20, 3), (24, 3), (28, 3)]


@pytest.mark.parametrize("sigma", [3, 5, 6, 7, 9, 10, 11, 20, 33, 63, 72, 75,
                                   100, 129, 255, 257, 1000, 40000])
def test_codes_are_a_prefix_code(sigma):
    """Codes are increasing, prefix-free, and lengths never increase with
    the symbol: the structure every level formula relies on."""
    code, ln, _ = O.minimal_codes(sigma)
    L = O.ceil_log2(sigma)
    prev = -1
    for c in range(sigma):
        assert int(code[c]) > prev
        prev = int(code[c])
        assert int(code[c]) < (1 << L)
        # low bits below the code length are zero
        assert int(code[c]) & ((1 << (L - int(ln[c]))) - 1) == 0
    assert np.all(np.diff(ln.astype(int)) <= 0)
    words = {(int(code[c]) >> (L - int(ln[c])), int(ln[c])) for c in range(sigma)}
    for w, l in words:
        for k in range(1, l):
            assert (w >> (l - k), k) not in words


# --- level structure (test_wavelet_tree.cu:446-565) ------------------------
def _bits(words, m):
    return [(int(words[i >> 5]) >> (i & 31)) & 1 for i in range(m)]


def test_structure_sigma10():
    data = np.arange(100) % 10
    wt = O.OracleTree(data, 10)
    sizes = wt.level_sizes()
    # symbols 8 and 9 get 2-bit codes (1000, 1100): levels 0-1 hold all 100
    # symbols, levels 2-3 only the 80 symbols 0..7
    assert list(sizes) == [100, 100, 80, 80]
    b0 = _bits(wt.level_words(0), 100)
    assert b0 == [1 if i % 10 >= 8 else 0 for i in range(100)]
    b1 = _bits(wt.level_words(1), 100)
    assert b1[:80] == [1 if i % 8 >= 4 else 0 for i in range(80)]
    assert b1[80:] == [i % 2 for i in range(80, 100)]
    b2 = _bits(wt.level_words(2), 80)
    assert b2 == [1 if i % 4 >= 2 else 0 for i in range(80)]
    b3 = _bits(wt.level_words(3), 80)
    assert b3 == [i % 2 for i in range(80)]


def test_structure_sigma5():
    data = [4, 1, 1, 2, 1, 1, 3, 4, 4, 0, 0, 2, 3, 2, 4, 2, 4, 1, 3, 2]
    wt = O.OracleTree(np.array(data), 5)
    assert list(wt.level_sizes()) == [20, 15, 15]
    assert "".join(map(str, _bits(wt.level_words(0), 20))) == "10000001100000101000"
    assert "".join(map(str, _bits(wt.level_words(1), 15))) == "001001001111011"
    assert "".join(map(str, _bits(wt.level_words(2), 15))) == "111100101010010"


@pytest.mark.parametrize("sigma", [3, 5, 10, 20, 75, 256])
def test_level_order_short_codes_sort_last(sigma):
    """The reference sorts the whole text by the top l bits and keeps the
    first n_l elements (wavelet_tree.cuh:806-818); the oracle keeps only the
    present elements.  Both must agree."""
    rng = np.random.default_rng(sigma)
    sym = rng.integers(0, sigma, 3000).astype(np.uint64)
    code, ln, _ = O.minimal_codes(sigma)
    L = O.ceil_log2(sigma)
    for l in range(L):
        keys = (code[sym] >> np.uint64(L - l)) if l else np.zeros_like(sym)
        full = np.argsort(keys, kind="stable")
        present = O.level_order(sym, code, ln, L, l)
        assert np.array_equal(full[: len(present)].astype(np.uint64), present)


# --- node positions (test_wavelet_tree.cu:219-259) ---------------------------
@pytest.mark.parametrize("sigma", [4, 5, 10, 20, 75, 100, 255, 256, 1000, 4099])
def test_node_pos_sequential(sigma):
    st, lv = O.node_infos(sigma)
    L = O.ceil_log2(sigma)
    counter = 0
    for j in range(len(st)):
        if j > 0 and lv[j] != lv[j - 1]:
            counter = 0
        assert O.node_pos_at_level(sigma, L, int(st[j]), int(lv[j])) == counter
        counter += 1


def test_node_count_pow2():
    for s in (4, 16, 256):
        st, _ = O.node_infos(s)
        assert len(st) == s - O.ceil_log2(s) - 1


@pytest.mark.parametrize("sigma", [5, 10, 20, 75, 100, 1000])
def test_nodes_at_level_loop_matches_per_level_count(sigma):
    wt = O.OracleTree(np.arange(sigma * 3) % sigma, sigma)
    _, lv = wt.node_starts()
    per_level = [int(np.sum(lv == l)) for l in range(wt.num_levels - 1)]
    assert list(wt.nodes_at_level()) == per_level


# --- queries: tree traversal == direct scan ---------------------------------
def _check_tree(data, sigma, nq=400, seed=0):
    data = np.asarray(data)
    wt = O.OracleTree(data, sigma)
    rng = np.random.default_rng(seed)
    n = len(data)
    hist = np.bincount(data, minlength=sigma)
    for i in rng.integers(0, n, nq):
        assert wt.access(int(i)) == data[i]
    for i, c in zip(rng.integers(0, n, nq), rng.integers(0, sigma, nq)):
        assert wt.rank(int(c), int(i)) == int(np.sum(data[:i] == c))
    for c in rng.integers(0, sigma, nq):
        cnt = int(hist[c])
        k = int(rng.integers(1, cnt + 2))  # includes one past the count
        pos = np.flatnonzero(data == c)
        exp = int(pos[k - 1]) if k <= cnt else n
        assert wt.select(int(c), k) == exp


def test_queries_kat_mod10():
    data = np.arange(100) % 10
    wt = O.OracleTree(data, 10)
    for j in range(100):
        assert wt.access(j) == j % 10
        assert wt.select(j % 10, j // 10 + 1) == j
        assert wt.rank(j % 10, j) == j // 10
    assert wt.select(0, 11) == 100


def test_queries_kat_mod256():
    data = np.arange(1280) % 256
    wt = O.OracleTree(data, 256)
    for j in range(0, 1280, 7):
        assert wt.access(j) == j % 256
        assert wt.select(j % 256, j // 256 + 1) == j
    assert wt.select(0, 11) == 1280


@pytest.mark.parametrize("sigma", [2, 3, 4, 5, 7, 10, 20, 33, 64, 75, 100,
                                   255, 256, 257, 1000])
def test_tree_vs_scan_random(sigma):
    rng = np.random.default_rng(sigma)
    n = int(rng.integers(sigma, 20000))
    _check_tree(rng.integers(0, sigma, n), sigma, seed=sigma)


def test_tree_vs_scan_skewed_and_missing_symbols():
    rng = np.random.default_rng(5)
    sigma = 75
    data = np.minimum(rng.zipf(1.3, 15000) - 1, sigma - 1)
    data[data == 70] = 3  # one symbol absent
    _check_tree(data, sigma, seed=1)


def test_by_construction_queries():
    """generateRandomDataAndRSQueries (test_benchmark_utils.cuh:417-528):
    rank(p, text[p]) = occurrences before p; select(occurrences up to and
    including p, text[p]) = p."""
    rng = np.random.default_rng(11)
    text = rng.integers(0, 37, 5000).astype(np.uint8)
    for p in rng.integers(0, 5000, 200):
        c = int(text[p])
        r = O.scan_rank(text, c, int(p))
        assert O.scan_select(text, c, r + 1) == p


# --- rank/select over bits (test_rank_select.cu) -----------------------------
def test_rs_index_sizes_kat():
    sizes = [1, 511, 513, 1023, 65535, 65537, 131071]
    l1 = [O.rs_sizes(m)[0] for m in sizes]
    l2 = [O.rs_sizes(m)[1] for m in sizes]
    last = [O.rs_sizes(m)[2] for m in sizes]
    assert l1 == [1, 1, 1, 1, 1, 2, 2]
    assert l2 == [1, 1, 2, 2, 128, 129, 256]
    assert last == [1, 1, 2, 2, 128, 1, 128]


@pytest.mark.parametrize("m", [1, 513, 1023, 65535, 65537, 131071, 131073,
                               196608])
@pytest.mark.parametrize("val", [0, 1])
def test_rs_all_ones_zeros(m, val):
    w = np.full((m + 31) // 32, 0xFFFFFFFF if val else 0, dtype=np.uint32)
    l1, l2, s1, s0, ones = O.rs_index(w, m)
    assert ones == (m if val else 0)
    if val:
        assert list(l1) == [65536 * k for k in range(len(l1))]
        assert list(l2) == [(512 * j) % 65536 for j in range(len(l2))]
        assert list(s1) == [4096 * (j + 1) - 1 for j in range(m // 4096)]
    rng = np.random.default_rng(m)
    for pos in rng.integers(1, m, 20) if m > 1 else []:
        pos = int(pos)
        assert O.bits_rank(w, m, val, pos) == pos
        assert O.bits_select(w, m, val, pos) == pos - 1
    assert O.bits_rank(w, m, val, m) == m
    assert O.bits_select(w, m, val, 1) == 0
    assert O.bits_select(w, m, val, m) == m - 1
    assert O.bits_rank(w, m, 1 - val, m) == 0
    assert O.bits_select(w, m, 1 - val, 1) == m


@pytest.mark.parametrize("m", [513, 1023, 65535, 65537, 131071, 131073])
def test_rs_pattern_ffff0000(m):
    """Closed forms of test_rank_select.cu:665-761."""
    w = np.full((m + 31) // 32, 0xFFFF0000, dtype=np.uint32)
    rng = np.random.default_rng(m)
    for pos in rng.integers(1, m, 50):
        pos = int(pos)
        last = pos % 32 - 16
        nwords = pos // 32
        rank1 = 32 * nwords // 2 + (last if last > 0 else 0)
        assert O.bits_rank(w, m, 1, pos) == rank1
        if rank1:
            sel1 = 32 * (rank1 // 16) - 1
            if rank1 % 16:
                sel1 += 16 + rank1 % 16
            assert O.bits_select(w, m, 1, rank1) == sel1
        rank0 = 32 * nwords // 2 + min(16, pos % 32)
        assert O.bits_rank(w, m, 0, pos) == rank0
        if rank0:
            sel0 = 32 * (rank0 // 16) - 1
            if rank0 % 16:
                sel0 += min(16, rank0 % 16)
            else:
                sel0 -= 16
            assert O.bits_select(w, m, 0, rank0) == sel0


@pytest.mark.parametrize("m", [2000, 65536, 100001, 300000])
def test_rs_index_random_vs_prefix(m):
    rng = np.random.default_rng(m)
    w = rng.integers(0, 2**32, (m + 31) // 32, dtype=np.uint64).astype(np.uint32)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")[:m].astype(np.int64)
    pref = np.concatenate([[0], np.cumsum(bits)])
    l1, l2, s1, s0, ones = O.rs_index(w, m)
    assert ones == pref[-1]
    assert np.array_equal(l1, pref[np.arange(len(l1)) * 65536])
    j = np.arange(len(l2))
    assert np.array_equal(l2.astype(np.int64),
                          pref[j * 512] - pref[(j // 128) * 65536])
    one_pos = np.flatnonzero(bits == 1)
    zero_pos = np.flatnonzero(bits == 0)
    assert np.array_equal(s1, one_pos[4095::4096][: len(s1)])
    assert len(s1) == ones // 4096
    assert np.array_equal(s0, zero_pos[4095::4096][: len(s0)])
    for k in rng.integers(1, max(ones, 1) + 1, 30):
        assert O.bits_select(w, m, 1, int(k)) == (one_pos[k - 1] if k <= ones else m)
    for i in rng.integers(0, m + 1, 30):
        assert O.bits_rank(w, m, 1, int(i)) == pref[i]


def test_rs_samples_kat():
    """All-ones / all-zeros samples are j*4096-1 (test_rank_select.cu:370-403)."""
    for m in [1, 4095, 4096, 4097, 8191, 8192, 8193, 6 * 4096, 6 * 512 + 102]:
        for val in (0, 1):
            w = np.full((m + 31) // 32, 0xFFFFFFFF if val else 0, dtype=np.uint32)
            _, _, s1, s0, ones = O.rs_index(w, m)
            s = s1 if val else s0
            assert list(s) == [j * 4096 - 1 for j in range(1, m // 4096 + 1)]


# --- sweep rank and the blocked oracle (variant B) ---------------------------
def _numpy_rank(text, pos, sym):
    """count of sym in text[:pos] from per-symbol prefix sums (numpy)."""
    out = np.empty(len(pos), dtype=np.uint64)
    for c in np.unique(sym):
        pref = np.concatenate([[0], np.cumsum(text == c, dtype=np.int64)])
        sel = sym == c
        out[sel] = pref[pos[sel].astype(np.int64)]
    return out


def _numpy_select(text, k, sym):
    n = len(text)
    out = np.empty(len(k), dtype=np.uint64)
    for c in np.unique(sym):
        occ = np.flatnonzero(text == c)
        sel = sym == c
        kk = k[sel].astype(np.int64)
        ok = (kk >= 1) & (kk <= len(occ))
        r = np.full(len(kk), n, dtype=np.uint64)
        r[ok] = occ[kk[ok] - 1]
        out[sel] = r
    return out


def _edge_positions(n, rng, extra):
    b = np.arange(0, n + 1, 65536, dtype=np.int64)
    e = np.concatenate([[0, 1, n - 1, n], b, b - 1, b + 1,
                        rng.integers(0, n + 1, extra)])
    return np.unique(np.clip(e, 0, n)).astype(np.uint64)


@pytest.mark.parametrize("case", [("u8", 4, 300_001), ("u8", 256, 200_000), ("u8", 20, 131_072),
                                  ("u16", 1000, 150_000), ("u8", 7, 5000), ("u8", 3, 1)])
def test_sweep_rank_vs_prefix_sums(case):
    """orc_sweep_rank equals numpy prefix sums for every (symbol, position),
    including position 0, n, block and range boundaries, repeated positions
    and symbols that do not occur."""
    dt, sigma, n = case
    rng = np.random.default_rng(sigma * 7 + n)
    text = rng.integers(0, sigma, n).astype(np.uint8 if dt == "u8" else np.uint16)
    pos = _edge_positions(n, rng, 3000)
    pos = np.concatenate([pos, pos[:50]])  # repeated positions
    sym = rng.integers(0, sigma + 2, len(pos)).astype(np.uint64)  # includes absent symbols
    sym = np.minimum(sym, sigma - 1) if sigma > 250 else sym
    got = O.sweep_rank(text, pos, sym, sigma=sigma + 2)
    assert np.array_equal(got, _numpy_rank(text, pos, sym))


@pytest.mark.parametrize("case", [("u8", 4, 300_001), ("u8", 256, 400_000), ("u8", 20, 131_072),
                                  ("u16", 1000, 200_000), ("u8", 5, 65_536), ("u8", 2, 3)])
def test_blocked_oracle_vs_numpy(case):
    """Variant B (per-symbol block counts + one-block scan) answers rank and
    select exactly as numpy prefix sums / occurrence lists, at block
    boundaries, position n, k = count + 1 (gives n) and absent symbols."""
    dt, sigma, n = case
    rng = np.random.default_rng(sigma * 11 + n)
    text = rng.integers(0, sigma, n).astype(np.uint8 if dt == "u8" else np.uint16)
    B = O.BlockedOracle(text, sigma + 1)  # symbol sigma never occurs
    pos = _edge_positions(n, rng, 2000)
    sym = rng.integers(0, sigma + 1, len(pos)).astype(np.uint64)
    assert np.array_equal(B.rank_batch(pos, sym), _numpy_rank(text, pos, sym))
    hist = np.bincount(text, minlength=sigma + 1)
    sym = rng.integers(0, sigma + 1, 3000).astype(np.uint64)
    k = np.array([rng.integers(0, hist[c] + 2) for c in sym], dtype=np.uint64)
    # every block's first and last occurrence of symbol 0
    occ0 = np.flatnonzero(text == 0)
    if len(occ0):
        ks = np.unique(np.searchsorted(occ0, np.arange(0, n + 65536, 65536)))
        ks = ks[(ks >= 0) & (ks < len(occ0))] + 1
        k = np.concatenate([k, ks.astype(np.uint64), ks.astype(np.uint64) + 1])
        sym = np.concatenate([sym, np.zeros(2 * len(ks), dtype=np.uint64)])
    assert np.array_equal(B.select_batch(k, sym), _numpy_select(text, k, sym))
    # the direct scans (variant A) agree on a sample
    assert np.array_equal(B.select_batch(k[:200], sym[:200]),
                          O.scan_select_batch(text, k[:200], sym[:200]))


def test_by_construction_round_trip():
    """The by-construction pairs of test_benchmark_utils.cuh:417-528: for a
    position p with c = text[p], rank(c, p) is the number of earlier c, and
    select(c, rank(c, p) + 1) is p."""
    rng = np.random.default_rng(5)
    n = 250_000
    text = O.gen_cdf(3, n, O.zipf_thresholds(64, 1.1), np.uint8)
    p = rng.integers(0, n, 4000).astype(np.uint64)
    c = text[p].astype(np.uint64)
    r = O.sweep_rank(text, p, c, sigma=64)
    B = O.BlockedOracle(text, 64)
    assert np.array_equal(B.select_batch(r + 1, c), p)
    assert np.array_equal(r, B.rank_batch(p, c))


@pytest.mark.parametrize("sigma", [2, 4, 256])
def test_direct_scans_at_count_block_edges(sigma):
    """The u8 direct scans count whole 16320-byte blocks at a time: rank at
    and around block edges, and select of occurrences that are the first or
    last byte of a block, equal numpy prefix sums / occurrence lists."""
    rng = np.random.default_rng(sigma)
    blk = 64 * 255
    n = 5 * blk + 77
    text = rng.integers(0, sigma, n).astype(np.uint8)
    edges = np.array([j * blk + d for j in range(6) for d in (-1, 0, 1)] + [n], np.int64)
    edges = np.unique(np.clip(edges, 0, n)).astype(np.uint64)
    for c in range(min(sigma, 4)):
        sym = np.full(len(edges), c, np.uint64)
        assert np.array_equal(O.scan_rank_batch(text, edges, sym), _numpy_rank(text, edges, sym))
        occ = np.flatnonzero(text == c)
        at = [q for q in range(len(occ)) if occ[q] % blk in (0, blk - 1)][:20]
        k = np.array([q + 1 for q in at] + [len(occ), len(occ) + 1, 1], np.uint64)
        sym = np.full(len(k), c, np.uint64)
        assert np.array_equal(O.scan_select_batch(text, k, sym), _numpy_select(text, k, sym))
