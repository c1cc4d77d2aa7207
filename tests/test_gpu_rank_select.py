"""BitArray + RankSelect on the GPU vs the oracle.  Mirrors the reference's
test_bit_array.cu and test_rank_select.cu; every value is compared bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

L1 = 65536
L2 = 512
SR = 4096


def _rand_words(rng, m):
    return rng.integers(0, 2**32, (m + 31) // 32, dtype=np.uint64).astype(np.uint32)


def _build(gpu, sizes, words_list=None, init=False):
    from ecl_wavelet_amd import BitArray, RankSelect
    ba = BitArray(sizes, init)
    if words_list is not None:
        for i, w in enumerate(words_list):
            ba.writeWords(i, 0, w)
    return RankSelect(ba)


# --- BitArray (test_bit_array.cu) -------------------------------------------
def test_bit_array_sizes_offsets_init(gpu):
    from ecl_wavelet_amd import BitArray
    sizes = [2, 4, 8, 64, 128, 1024]
    for init in (True, False):
        ba = BitArray(sizes, init)
        assert ba.numArrays() == len(sizes)
        off = 0
        for i, s in enumerate(sizes):
            assert ba.size(i) == s
            assert ba.getOffset(i) == off
            off += ((s + 31) // 32 + 31) // 32 * 32
            assert np.all(ba.bits(i) == (1 if init else 0))


def test_bit_array_write_read_words(gpu):
    from ecl_wavelet_amd import BitArray
    ba = BitArray([2, 4, 8, 64, 128, 1024], False)
    ba.writeWords(2, 0, [1 << 1])
    b = ba.bits(2)
    assert b[0] == 0 and b[1] == 1
    for i in (0, 1, 3, 4, 5):
        assert not ba.bits(i).any()
    ba = BitArray([66000], False)
    ba.writeWords(0, 0, [0xFFFFFFFF])
    ba.writeWords(0, 1, [0])
    w = ba.words(0, 0, 2)
    assert (int(w[1]) << 32 | int(w[0])) == 0xFFFFFFFF
    ba.writeWords(0, 0, [0, 0xFFFFFFFF])
    w = ba.words(0, 0, 2)
    assert (int(w[1]) << 32 | int(w[0])) == 0xFFFFFFFF00000000


# --- RankSelect index (test_rank_select.cu:178-403) --------------------------
def test_index_sizes_kat(gpu):
    sizes = [1, L2 - 1, L2 + 1, 2 * L2 - 1, L1 - 1, L1 + 1, 2 * L1 - 1]
    rs = _build(gpu, sizes)
    assert [rs.getNumL1Blocks(i) for i in range(7)] == [1, 1, 1, 1, 1, 2, 2]
    assert [rs.getNumL2Blocks(i) for i in range(7)] == [1, 1, 2, 2, 128, 129, 256]
    assert [rs.getNumLastL2Blocks(i) for i in range(7)] == [1, 1, 2, 2, 128, 1, 128]


def test_index_all_zero_content(gpu):
    sizes = [1, L2 - 1, L2 + 1, 2 * L2 - 1, L1 - 1, L1 + 1, 2 * L1 - 1]
    rs = _build(gpu, sizes)
    for i in range(len(sizes)):
        assert not rs.l1Entries(i).any()
        assert not rs.l2Entries(i).any()
        assert rs.getTotalNumVals(i, 1) == 0


def test_index_all_ones_content(gpu):
    sizes = [1, L2 + 1, 2 * L2 - 1, L1 - 1, L1 + 1, 2 * L1 - 1, 2 * L1 + 1]
    rs = _build(gpu, sizes, init=True)
    for i, m in enumerate(sizes):
        assert list(rs.l1Entries(i)) == [L1 * k for k in range(rs.getNumL1Blocks(i))]
        assert list(rs.l2Entries(i)) == [(L2 * j) % L1 for j in range(rs.getNumL2Blocks(i))]
        assert rs.getTotalNumVals(i, 1) == m


@pytest.mark.parametrize("val", [True, False])
def test_samples_kat(gpu, val):
    sizes = [1, SR - 1, SR, SR + 1, 2 * SR - 1, 2 * SR, 2 * SR + 1, 6 * SR,
             6 * L2 + L2 // 5]
    rs = _build(gpu, sizes, init=val)
    for i, m in enumerate(sizes):
        assert rs.getTotalNumVals(i, val) == m
        assert rs.getTotalNumVals(i, not val) == 0
        s = rs.selectSamples(i, val)
        assert list(s) == [j * SR - 1 for j in range(1, m // SR + 1)]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_index_random_vs_oracle(gpu, seed):
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(200, 10**6, 10)]
    words = [_rand_words(rng, m) for m in sizes]
    rs = _build(gpu, sizes, words)
    for i, m in enumerate(sizes):
        l1, l2, s1, s0, ones = O.rs_index(words[i], m)
        assert rs.getTotalNumVals(i, 1) == ones
        assert np.array_equal(rs.l1Entries(i), l1)
        assert np.array_equal(rs.l2Entries(i), l2)
        assert np.array_equal(rs.selectSamples(i, 1), s1)
        assert np.array_equal(rs.selectSamples(i, 0), s0)


# --- rank/select operations (test_rank_select.cu:557-855) ---------------------
@pytest.mark.parametrize("val", [1, 0])
def test_operations_all_same(gpu, val):
    sizes = [L2 + 1, 2 * L2 - 1, L1 - 1, L1 + 1, 2 * L1 - 1, 2 * L1 + 1, 3 * L1]
    rs = _build(gpu, sizes, init=bool(val))
    rng = np.random.default_rng(7)
    for i, m in enumerate(sizes):
        pos = rng.integers(1, m, 100).astype(np.uint64)
        assert np.array_equal(rs.rank(i, pos, val), pos)
        assert np.array_equal(rs.select(i, pos, val), pos - 1)
        edge = np.array([0, m], dtype=np.uint64)
        assert list(rs.rank(i, edge, val)) == [0, m]
        assert list(rs.select(i, np.array([1, m]), val)) == [0, m - 1]
        assert list(rs.rank(i, np.array([m]), 1 - val)) == [0]
        assert list(rs.select(i, np.array([1]), 1 - val)) == [m]


def test_operations_pattern_ffff0000(gpu):
    sizes = [L2 + 1, 2 * L2 - 1, L1 - 1, L1 + 1, 2 * L1 - 1, 2 * L1 + 1]
    words = [np.full((m + 31) // 32, 0xFFFF0000, dtype=np.uint32) for m in sizes]
    rs = _build(gpu, sizes, words)
    rng = np.random.default_rng(9)
    for i, m in enumerate(sizes):
        pos = [int(p) for p in rng.integers(1, m, 100)]
        r1 = [32 * (p // 32) // 2 + max(p % 32 - 16, 0) for p in pos]
        r0 = [32 * (p // 32) // 2 + min(16, p % 32) for p in pos]
        assert list(rs.rank(i, pos, 1)) == r1
        assert list(rs.rank(i, pos, 0)) == r0
        k1 = [r for r in r1 if r]
        s1 = [32 * (r // 16) - 1 + (16 + r % 16 if r % 16 else 0) for r in k1]
        assert list(rs.select(i, k1, 1)) == s1
        k0 = [r for r in r0 if r]
        s0 = [32 * (r // 16) - 1 + (min(16, r % 16) if r % 16 else -16) for r in k0]
        assert list(rs.select(i, k0, 0)) == s0


@pytest.mark.parametrize("seed", [11, 12])
def test_operations_random_vs_oracle(gpu, seed):
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(2000, 10**6, 10)]
    words = [_rand_words(rng, m) for m in sizes]
    rs = _build(gpu, sizes, words)
    for i, m in enumerate(sizes):
        idx = rng.integers(0, m + 1, 500).astype(np.uint64)
        for v in (0, 1):
            exp = O.scan_bits_batch(words[i], m, v, idx)
            assert np.array_equal(rs.rank(i, idx, v), exp)
            total = rs.getTotalNumVals(i, v)
            k = rng.integers(1, total + 2, 500).astype(np.uint64)  # one past
            exp = O.scan_bits_batch(words[i], m, 2 + v, k)
            assert np.array_equal(rs.select(i, k, v), exp)


@pytest.mark.parametrize("p", [0.5, 0.05, 0.001])
def test_large_generated_bits_and_queries(gpu, p):
    """A 2^26-bit array filled on the device (Bernoulli p), directory and
    device-generated queries checked against the oracle on the same bits."""
    from ecl_wavelet_amd import BitArray, DeviceBuffer, RankSelect
    m = (1 << 26) + 12345
    thr = O.bernoulli_threshold(p)
    ba = BitArray([m])
    ba.fillRandom(0, 5, thr)
    rs = RankSelect(ba)
    words = O.gen_bits(5, m, thr)
    assert np.array_equal(rs.bit_array_.words(0), words)
    l1, l2, s1, s0, ones = O.rs_index(words, m)
    assert rs.getTotalNumVals(0, 1) == ones
    assert np.array_equal(rs.l1Entries(0), l1)
    assert np.array_equal(rs.l2Entries(0), l2)
    assert np.array_equal(rs.selectSamples(0, 1), s1)
    assert np.array_equal(rs.selectSamples(0, 0), s0)
    q = 20000
    out = DeviceBuffer(q * 8)
    rs.rankGenerated(0, 77, q, out, 1)
    got = out.download(np.uint64)
    idx = np.array([(O.hash64(77, j) * m) >> 64 for j in range(q)], dtype=np.uint64)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:m]
    pref = np.concatenate([[0], np.cumsum(bits, dtype=np.uint64)])
    assert np.array_equal(got, pref[idx])
    rs.selectGenerated(0, 78, q, out, 1)
    got = out.download(np.uint64)
    k = np.array([1 + ((O.hash64(78, j) * ones) >> 64) for j in range(q)], dtype=np.uint64)
    one_pos = np.flatnonzero(bits).astype(np.uint64)
    assert np.array_equal(got, one_pos[k - 1])


def _rank_from_index(words, l1, l2, idx):
    """rank1(i) from the oracle's directory: L1 + L2 + ones of the words of
    the 512-bit block before i (vectorized; independent of the kernels)."""
    idx = idx.astype(np.uint64)
    blk = (idx >> np.uint64(9)).astype(np.int64)
    base = l1[blk >> 7] + l2[blk].astype(np.uint64)
    w0 = blk * 16
    out = base.copy()
    wi = (idx >> np.uint64(5)).astype(np.int64)
    for j in range(16):
        w = w0 + j
        full = w < wi
        out += np.where(full, np.bitwise_count(words[np.minimum(w, len(words) - 1)]), 0).astype(np.uint64)
    part = (idx & np.uint64(31)).astype(np.uint32)
    last = words[np.minimum(wi, len(words) - 1)] & ((np.uint64(1) << part.astype(np.uint64)) - np.uint64(1)).astype(np.uint32)
    out += np.where(part > 0, np.bitwise_count(last), 0).astype(np.uint64)
    return out


@pytest.mark.parametrize("p", [0.5, 0.05])
def test_binary_2e34_bits(gpu, p):
    """The binary RankSelect configuration at its full size: 2^34 Bernoulli
    bits drawn on the device.  L1, L2 and both sample arrays equal the
    oracle's directory over the same words; device-generated rank1 queries
    equal L1 + L2 + in-block ones; every generated select1(k) answer j has
    bit j set and rank1(j) = k - 1."""
    from ecl_wavelet_amd import BitArray, DeviceBuffer, RankSelect
    m = 1 << 34
    thr = O.bernoulli_threshold(p)
    ba = BitArray([m])
    ba.fillRandom(0, 2 if p == 0.5 else 3, thr)
    rs = RankSelect(ba)
    words = rs.bit_array_.words(0)
    l1, l2, s1, s0, ones = O.rs_index(words, m)
    assert rs.getTotalNumVals(0, 1) == ones
    assert np.array_equal(rs.l1Entries(0), l1)
    assert np.array_equal(rs.l2Entries(0), l2)
    assert np.array_equal(rs.selectSamples(0, 1), s1)
    assert np.array_equal(rs.selectSamples(0, 0), s0)
    q = 50_000
    out = DeviceBuffer(q * 8)
    rs.rankGenerated(0, 91, q, out, 1)
    got = out.download(np.uint64)
    idx = np.array([(O.hash64(91, j) * m) >> 64 for j in range(q)], dtype=np.uint64)
    assert np.array_equal(got, _rank_from_index(words, l1, l2, idx))
    rs.selectGenerated(0, 92, q, out, 1)
    pos = out.download(np.uint64)
    k = np.array([1 + ((O.hash64(92, j) * ones) >> 64) for j in range(q)], dtype=np.uint64)
    assert np.all(pos < m)
    bit = (words[(pos >> np.uint64(5)).astype(np.int64)] >> (pos & np.uint64(31)).astype(np.uint32)) & 1
    assert np.all(bit == 1)
    assert np.array_equal(_rank_from_index(words, l1, l2, pos), k - np.uint64(1))


@pytest.mark.parametrize("order", ["query_lines_first", "select_lines_first"])
@pytest.mark.parametrize("p", [0.3, 0.02, 0.985])
def test_query_lines_binary(gpu, p, order):
    """Binary rank and select of ones and zeros with query lines and select
    hints built equal the oracle's scans and the answers without them, for
    balanced (0.3), sparse (0.02) and dense (0.985) arrays: at 0.3 every
    sample bracket fits one hint (at most 24 blocks); the rare Value of the
    sparse and dense arrays has brackets wider than that and takes the
    directory search.  Select lines (a line every r = floor(800 density)
    ones or zeros) are built before or after the query lines and give the
    same answers.  The select arguments include every sample multiple
    (4096 j) and every multiple of r, their neighbours, the total and one
    past it."""
    from ecl_wavelet_amd import BitArray, RankSelect
    m = (1 << 22) + 4321
    thr = O.bernoulli_threshold(p)
    sizes = [m, 960 * 7, 5, 4096 * 3 + 17]
    ba = BitArray(sizes)
    for i in range(len(sizes)):
        ba.fillRandom(i, 40 + i, thr)
    rs = RankSelect(ba)
    words = [rs.bit_array_.words(i) for i in range(len(sizes))]
    rng = np.random.default_rng(3)
    idxs, ks = [], []
    for i, mm in enumerate(sizes):
        idxs.append(np.concatenate([rng.integers(0, mm + 1, 3000), np.arange(0, mm + 1, 960),
                                    np.arange(959, mm, 960), [0, mm]]).astype(np.uint64))
        kv = []
        for v in (0, 1):
            total = rs.getTotalNumVals(i, v)
            samp = np.arange(4096, total + 1, 4096, dtype=np.int64)
            r = max(1, 800 * total // mm) if total else 1
            lines = np.arange(r, total + 1, r, dtype=np.int64)[:4000]
            k = np.concatenate([rng.integers(1, total + 2, 3000), samp, samp - 1, samp + 1,
                                lines, lines - 1, lines + 1, [0, 1, total, total + 1]])
            kv.append(k[k >= 0].astype(np.uint64))
        ks.append(kv)
    before = [[rs.select(i, ks[i][v], v) for v in (0, 1)] for i in range(len(sizes))]
    steps = [rs.buildQueryLines, rs.buildSelectLines]
    if order == "select_lines_first":
        steps.reverse()
    for step in steps:
        step()
        for i, mm in enumerate(sizes):
            for v in (0, 1):
                assert np.array_equal(rs.rank(i, idxs[i], v),
                                      O.scan_bits_batch(words[i], mm, v, idxs[i]))
                got = rs.select(i, ks[i][v], v)
                assert np.array_equal(got, O.scan_bits_batch(words[i], mm, 2 + v, ks[i][v]))
                assert np.array_equal(got, before[i][v])
