"""Full-size trees on the GPU against the oracle.

The full-size configurations -- uint8 texts of 6e9 symbols uniform over
sigma = 256 and sigma = 4, and a uint16 text of 2e9 symbols drawn from
Zipf(1.1) over sigma = 2^16 -- are built on the device, and 10^7 queries of
each kind run in single launches of the shipped kernels (8192 workgroups,
grid-stride), over the reference-shaped directory, then with the query lines
and select hints, then with the select lines, then with the quad lines (which
access and rank then read: all three alphabets are powers of two with an even
number of levels), then with the quad select lines (which select then
reads), then with the hex lines (which access and rank then read, where the
level count is a multiple of 4: sigma = 256 and 2^16); the mixed-kind kernel
runs the same 3 * 10^7 queries in one launch in every layout.

Every answer is checked, over the whole launch:
  * access(i) against the text copied back from the device;
  * rank(c, i) against the oracle's one-pass sweep over that text
    (orc_sweep_rank), for uniform positions with c = text[i] and with c
    uniform over the alphabet, plus i = 0 and i = n;
  * select(c, k) by construction, as the reference's queriesRandom test
    does (test_benchmark_utils.cuh:417-528, test_wavelet_tree.cu:709-794):
    for a uniform position p with c = text[p], select(c, rank(c, p) + 1)
    is p; k = count(c) + 1 gives n.
These sizes take the 64-bit tile, word and line offsets of construction and
queries past 2^32.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

M = 10_000_000

CASES = {
    "u8_sigma256": dict(n=6_000_000_000, sigma=256, dtype=np.uint8, seed=4, zipf=None),
    "u8_sigma4": dict(n=6_000_000_000, sigma=4, dtype=np.uint8, seed=5, zipf=None),
    "u16_zipf": dict(n=2_000_000_000, sigma=1 << 16, dtype=np.uint16, seed=8, zipf=1.1),
}
LAYOUTS = ["plain", "query_lines", "select_lines", "quad_lines", "quad_select", "hex_lines"]


class Full:
    pass


def _say(*a):
    print("[fullsize]", *a, flush=True)


@pytest.fixture(scope="module", params=list(CASES))
def full(request, gpu):
    from ecl_wavelet_amd import DeviceBuffer, WaveletTree, api
    cfg = CASES[request.param]
    n, sigma, dt = cfg["n"], cfg["sigma"], cfg["dtype"]
    _say(request.param, "generate + build")
    thr = None if cfg["zipf"] is None else api.zipf_thresholds(sigma, cfg["zipf"])
    d_text = api.generate_text(n, sigma, dt, seed=cfg["seed"], thresholds=thr)
    wt = WaveletTree(d_text, np.arange(sigma, dtype=dt), dtype=dt, n=n)
    _say(request.param, "download text")
    text = d_text.download(dt, n)
    d_text.free()
    hist = O.histogram(text, sigma)
    rng = np.random.default_rng(cfg["seed"] + 100)
    F = Full()
    F.name, F.n, F.sigma, F.dtype, F.wt, F.text, F.hist = request.param, n, sigma, dt, wt, text, hist
    # access: uniform positions plus both ends
    acc = rng.integers(0, n, M).astype(np.uint64)
    acc[:2] = [0, n - 1]
    F.acc_idx, F.acc_exp = acc, text[acc]
    # rank: first half c = text[i], second half c uniform; i = 0 and i = n included
    ri = rng.integers(0, n, M).astype(np.uint64)
    ri[:2] = [0, n]
    rc = rng.integers(0, sigma, M).astype(np.uint64)
    half = M // 2
    rc[2:half] = text[ri[2:half]]
    # select by construction: p uniform, c = text[p]
    sp = rng.integers(0, n, M).astype(np.uint64)
    sc = text[sp].astype(np.uint64)
    _say(request.param, "sweep oracle")
    both = O.sweep_rank(text, np.concatenate([ri, sp]), np.concatenate([rc, sc]), sigma=sigma)
    F.rank_q = WaveletTree.make_queries(ri, rc)
    F.rank_exp = both[:M]
    sk = both[M:] + 1
    sel_exp = sp.copy()
    # k = count(c) + 1 for the last 1000 queries: not found, gives n
    tail = slice(M - 1000, M)
    sk[tail] = hist[sc[tail]] + 1
    sel_exp[tail] = n
    F.sel_q = WaveletTree.make_queries(sk, sc)
    F.sel_exp = sel_exp
    # one mixed batch of the same 3M queries in a random order
    from ecl_wavelet_amd import MIXED_DTYPE
    mixed = np.zeros(3 * M, dtype=MIXED_DTYPE)
    mixed["arg"] = np.concatenate([acc, ri, sk])
    mixed["symbol"] = np.concatenate([np.zeros(M, np.uint64), rc, sc]).astype(np.uint32)
    mixed["kind"] = np.repeat(np.arange(3, dtype=np.uint32), M)
    exp = np.concatenate([F.acc_exp.astype(np.uint64), F.rank_exp, sel_exp])
    perm = rng.permutation(3 * M)
    F.mixed, F.mixed_exp = mixed[perm], exp[perm]
    F.d_in = DeviceBuffer(3 * M * 16)
    F.d_out = DeviceBuffer(3 * M * 8)
    _say(request.param, "ready")
    yield F
    F.d_in.free()
    F.d_out.free()
    del F.wt


def _ensure(F, layout):
    if layout in ("query_lines", "select_lines", "quad_lines", "quad_select", "hex_lines"):
        F.wt.buildQueryLines()
    if layout in ("select_lines", "quad_lines", "quad_select", "hex_lines"):
        F.wt.buildSelectLines()
    if layout in ("quad_lines", "quad_select", "hex_lines"):
        F.wt.buildQuadLines()
    if layout in ("quad_select", "hex_lines"):
        F.wt.buildQuadSelectLines()
    if layout == "hex_lines":
        if F.wt.getNumLevels() % 4 != 0:
            pytest.skip("hex lines need a level count that is a multiple of 4")
        F.wt.buildHexLines()


def _first_bad(got, exp):
    bad = np.flatnonzero(got != exp)
    return (len(bad), bad[:5].tolist(), got[bad[:5]].tolist(), exp[bad[:5]].tolist())


@pytest.mark.parametrize("layout", LAYOUTS)
def test_fullsize(full, layout):
    """The three per-kind kernels, then the mixed kernel, in one layout
    (layouts run in the listed order: each adds structures to the tree)."""
    F = full
    _ensure(F, layout)
    F.d_in.upload(F.acc_idx)
    _say(F.name, layout, "access")
    F.wt.accessDevice(F.d_in, M, F.d_out)
    got = F.d_out.download(F.dtype, M)
    assert np.array_equal(got, F.acc_exp), ("access",) + _first_bad(got, F.acc_exp)
    F.d_in.upload(F.rank_q)
    _say(F.name, layout, "rank")
    F.wt.rankDevice(F.d_in, M, F.d_out)
    got = F.d_out.download(np.uint64, M)
    assert np.array_equal(got, F.rank_exp), ("rank",) + _first_bad(got, F.rank_exp)
    F.d_in.upload(F.sel_q)
    _say(F.name, layout, "select")
    F.wt.selectDevice(F.d_in, M, F.d_out)
    got = F.d_out.download(np.uint64, M)
    assert np.array_equal(got, F.sel_exp), ("select",) + _first_bad(got, F.sel_exp)
    F.d_in.upload(F.mixed)
    _say(F.name, layout, "mixed")
    F.wt.queryBatchDevice(F.d_in, 3 * M, F.d_out)
    got = F.d_out.download(np.uint64, 3 * M)
    assert np.array_equal(got, F.mixed_exp), ("mixed",) + _first_bad(got, F.mixed_exp)
