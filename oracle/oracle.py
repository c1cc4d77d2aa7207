"""ctypes view of ``libecl_oracle.so`` (test infrastructure only).

The C sources in this directory restate the reference's semantics; each C
function cites the reference file:line it follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libecl_oracle.so")
_lib = None

u64 = C.c_uint64
u32 = C.c_uint32
vp = C.c_void_p


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_hash": (u64, [u64, u64]),
            "orc_bounded": (u64, [u64, u64]),
            "orc_gen_uniform": (None, [u64, u64, u64, C.c_int, vp]),
            "orc_gen_cdf": (None, [u64, u64, vp, u64, C.c_int, vp]),
            "orc_gen_bits": (None, [u64, u64, u64, vp]),
            "orc_is_pow2": (C.c_int, [u64]),
            "orc_prev_pow2": (u64, [u64]),
            "orc_ceil_log2": (u32, [u64]),
            "orc_minimal_codes": (u64, [u64, vp, vp]),
            "orc_node_infos": (u64, [u64, vp, vp, vp, vp]),
            "orc_node_pos_at_level": (u64, [u64, u32, u64, u32]),
            "orc_num_nodes_at_level": (None, [vp, u64, u32, vp]),
            "orc_bits_rank": (u64, [vp, u64, C.c_int, u64]),
            "orc_bits_select": (u64, [vp, u64, C.c_int, u64]),
            "orc_rs_num_l1": (u64, [u64]),
            "orc_rs_num_last_l2": (u64, [u64]),
            "orc_rs_num_l2": (u64, [u64]),
            "orc_rs_index": (u64, [vp, u64, vp, vp, vp, vp]),
            "orc_wt_build": (C.c_int, [vp, vp, u64, u64]),
            "orc_wt_free": (None, [vp]),
            "orc_wt_access": (u64, [vp, u64]),
            "orc_wt_rank": (u64, [vp, u64, u64]),
            "orc_wt_select": (u64, [vp, u64, u64]),
            "orc_level_order": (None, [vp, u64, vp, vp, u32, u32, vp, vp]),
            "orc_scan_rank": (u64, [vp, C.c_int, u64, u64, u64]),
            "orc_scan_select": (u64, [vp, C.c_int, u64, u64, u64]),
            "orc_scan_access_batch": (C.c_int, [vp, C.c_int, u64, vp, u64, vp]),
            "orc_scan_rank_batch": (C.c_int, [vp, C.c_int, u64, vp, vp, u64, vp]),
            "orc_scan_select_batch": (C.c_int, [vp, C.c_int, u64, vp, vp, u64, vp]),
            "orc_scan_bits_batch": (C.c_int, [vp, u64, C.c_int, vp, u64, vp]),
            "orc_max_threads": (C.c_int, []),
            "orc_histogram": (None, [vp, C.c_int, u64, u64, vp]),
            "orc_sweep_rank": (C.c_int, [vp, C.c_int, u64, u64, vp, vp, u64, vp]),
            "orc_blocked_rows": (u64, [u64]),
            "orc_blocked_build": (None, [vp, C.c_int, u64, u64, vp]),
            "orc_blocked_rank": (u64, [vp, C.c_int, u64, u64, vp, u64, u64]),
            "orc_blocked_select": (u64, [vp, C.c_int, u64, u64, vp, u64, u64]),
            "orc_blocked_rank_batch": (C.c_int, [vp, C.c_int, u64, u64, vp, vp, vp, u64, vp]),
            "orc_blocked_select_batch": (C.c_int, [vp, C.c_int, u64, u64, vp, vp, vp, u64, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(vp)


def hash64(seed, i):
    return lib().orc_hash(seed, i)


def gen_uniform(seed, n, sigma, dtype=np.uint8):
    out = np.empty(n, dtype=dtype)
    lib().orc_gen_uniform(seed, n, sigma, out.itemsize, _p(out))
    return out


def gen_cdf(seed, n, thresholds, dtype=np.uint16):
    thr = np.ascontiguousarray(thresholds, dtype=np.uint64)
    out = np.empty(n, dtype=dtype)
    lib().orc_gen_cdf(seed, n, _p(thr), len(thr), out.itemsize, _p(out))
    return out


def gen_bits(seed, nbits, threshold):
    w = np.empty((nbits + 31) // 32, dtype=np.uint32)
    lib().orc_gen_bits(seed, nbits, threshold, _p(w))
    return w


def zipf_thresholds(sigma, s=1.1):
    """u64 inverse-CDF thresholds for Zipf(s) over sigma symbols (last is
    2^64-1).  Computed once here and handed to both the oracle and the GPU
    generator so both sides draw identical texts."""
    w = 1.0 / np.power(np.arange(1, sigma + 1, dtype=np.float64), s)
    cdf = np.cumsum(w) / w.sum()
    thr = np.minimum(np.floor(cdf * 2.0 ** 64), 2.0 ** 64 - 1)
    out = np.array([int(x) if x < 2.0 ** 64 else 2 ** 64 - 1 for x in thr],
                   dtype=np.uint64)
    out[-1] = np.uint64(2 ** 64 - 1)
    return out


def bernoulli_threshold(p):
    return min(int(p * 2.0 ** 64), 2 ** 64 - 1)


def is_pow2(n):
    return bool(lib().orc_is_pow2(n))


def prev_pow2(n):
    return lib().orc_prev_pow2(n)


def ceil_log2(n):
    return lib().orc_ceil_log2(n)


def minimal_codes(sigma):
    code = np.empty(sigma, dtype=np.uint64)
    ln = np.empty(sigma, dtype=np.uint8)
    total = lib().orc_minimal_codes(sigma, _p(code), _p(ln))
    return code, ln, total


def node_infos(sigma):
    code, ln, _ = minimal_codes(sigma)
    cnt = lib().orc_node_infos(sigma, _p(code), _p(ln), None, None)
    st = np.empty(max(cnt, 1), dtype=np.uint64)
    lv = np.empty(max(cnt, 1), dtype=np.uint8)
    lib().orc_node_infos(sigma, _p(code), _p(ln), _p(st), _p(lv))
    return st[:cnt], lv[:cnt]


def node_pos_at_level(sigma, num_levels, symbol, level):
    return lib().orc_node_pos_at_level(sigma, num_levels, symbol, level)


def bits_rank(words, m, value, i):
    return lib().orc_bits_rank(_p(words), m, value, i)


def bits_select(words, m, value, k):
    return lib().orc_bits_select(_p(words), m, value, k)


def rs_sizes(m):
    L = lib()
    return L.orc_rs_num_l1(m), L.orc_rs_num_l2(m), L.orc_rs_num_last_l2(m)


def rs_index(words, m):
    """(L1, L2, S1, S0, total_ones) of a bit vector of m bits."""
    n1, n2, _ = rs_sizes(m)
    l1 = np.zeros(n1, dtype=np.uint64)
    l2 = np.zeros(n2, dtype=np.uint16)
    s1 = np.zeros(m // 4096 + 1, dtype=np.uint64)
    s0 = np.zeros(m // 4096 + 1, dtype=np.uint64)
    ones = lib().orc_rs_index(_p(words), m, _p(l1), _p(l2), _p(s1), _p(s0))
    return l1, l2, s1[: ones // 4096], s0[: (m - ones) // 4096], ones


class _WT(C.Structure):
    _fields_ = [
        ("n", u64), ("sigma", u64), ("num_levels", u32), ("codes_start", u64),
        ("code", vp), ("len", vp), ("counts", vp), ("level_size", vp),
        ("level_words", vp), ("num_ranks", u64), ("node_start", vp),
        ("node_level", vp), ("ranks", vp), ("nodes_at_level", vp),
        ("nodes_until_last_level", u64),
    ]


class OracleTree:
    """The oracle's wavelet tree over minimal-alphabet symbols."""

    def __init__(self, sym, sigma):
        self._sym = np.ascontiguousarray(sym, dtype=np.uint64)
        self._wt = _WT()
        rc = lib().orc_wt_build(C.byref(self._wt), _p(self._sym),
                                len(self._sym), sigma)
        if rc != 0:
            raise RuntimeError("oracle build failed")

    def __del__(self):
        if getattr(self, "_wt", None) is not None and _lib is not None:
            lib().orc_wt_free(C.byref(self._wt))
            self._wt = None

    def _arr(self, ptr, n, dtype):
        if n == 0:
            return np.zeros(0, dtype=dtype)
        buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype).copy()

    @property
    def num_levels(self):
        return self._wt.num_levels

    @property
    def codes_start(self):
        return self._wt.codes_start

    def counts(self):
        return self._arr(self._wt.counts, self._wt.sigma, np.uint64)

    def level_sizes(self):
        return self._arr(self._wt.level_size, self._wt.num_levels, np.uint64)

    def level_words(self, l):
        m = int(self.level_sizes()[l])
        ptrs = self._arr(self._wt.level_words, self._wt.num_levels, np.uint64)
        return self._arr(int(ptrs[l]), (m + 31) // 32, np.uint32)

    def ranks(self):
        return self._arr(self._wt.ranks, self._wt.num_ranks, np.uint64)

    def node_starts(self):
        return (self._arr(self._wt.node_start, self._wt.num_ranks, np.uint64),
                self._arr(self._wt.node_level, self._wt.num_ranks, np.uint8))

    def nodes_at_level(self):
        L = self._wt.num_levels
        return self._arr(self._wt.nodes_at_level, max(L - 1, 0), np.uint64)

    def access(self, i):
        return lib().orc_wt_access(C.byref(self._wt), i)

    def rank(self, c, i):
        return lib().orc_wt_rank(C.byref(self._wt), c, i)

    def select(self, c, k):
        return lib().orc_wt_select(C.byref(self._wt), c, k)


def level_order(sym, code, ln, num_levels, level):
    sym = np.ascontiguousarray(sym, dtype=np.uint64)
    order = np.empty(max(len(sym), 1), dtype=np.uint64)
    cnt = u64(0)
    lib().orc_level_order(_p(sym), len(sym), _p(code), _p(ln), num_levels,
                          level, _p(order), C.byref(cnt))
    return order[: cnt.value]


def scan_rank(text, c, i):
    return lib().orc_scan_rank(_p(text), text.itemsize, len(text), c, i)


def scan_select(text, c, k):
    return lib().orc_scan_select(_p(text), text.itemsize, len(text), c, k)


def scan_access_batch(text, idx):
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.empty(len(idx), dtype=np.uint64)
    lib().orc_scan_access_batch(_p(text), text.itemsize, len(text), _p(idx),
                                len(idx), _p(out))
    return out


def scan_rank_batch(text, idx, sym):
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    sym = np.ascontiguousarray(sym, dtype=np.uint64)
    out = np.empty(len(idx), dtype=np.uint64)
    lib().orc_scan_rank_batch(_p(text), text.itemsize, len(text), _p(idx),
                              _p(sym), len(idx), _p(out))
    return out


def scan_select_batch(text, k, sym):
    k = np.ascontiguousarray(k, dtype=np.uint64)
    sym = np.ascontiguousarray(sym, dtype=np.uint64)
    out = np.empty(len(k), dtype=np.uint64)
    lib().orc_scan_select_batch(_p(text), text.itemsize, len(text), _p(k),
                                _p(sym), len(k), _p(out))
    return out


def scan_bits_batch(words, m, op, arg):
    arg = np.ascontiguousarray(arg, dtype=np.uint64)
    out = np.empty(len(arg), dtype=np.uint64)
    lib().orc_scan_bits_batch(_p(words), m, op, _p(arg), len(arg), _p(out))
    return out


def histogram(text, sigma):
    h = np.empty(sigma, dtype=np.uint64)
    lib().orc_histogram(_p(text), text.itemsize, len(text), sigma, _p(h))
    return h


def max_threads():
    return lib().orc_max_threads()


def sweep_rank(text, pos, sym, sigma=None):
    """count of sym[j] in text[0:pos[j]] for every j (pos[j] <= len(text)),
    from one pass over the text (orc_sweep_rank); any order of queries."""
    pos = np.ascontiguousarray(pos, dtype=np.uint64)
    sym = np.ascontiguousarray(sym, dtype=np.uint64)
    if sigma is None:
        sigma = 1 << (8 * text.itemsize) if text.itemsize <= 2 else int(text.max()) + 1
    order = np.argsort(pos, kind="stable")
    ps, ss = pos[order], sym[order]
    out_s = np.empty(len(pos), dtype=np.uint64)
    r = lib().orc_sweep_rank(_p(text), text.itemsize, len(text), sigma, _p(ps), _p(ss),
                             len(ps), _p(out_s))
    if r < 0:
        raise MemoryError("orc_sweep_rank")
    out = np.empty_like(out_s)
    out[order] = out_s
    return out


class BlockedOracle:
    """Variant B of the CPU baseline: per-symbol counts before every 2^16
    symbols plus a scan of one block (orc_blocked_*)."""

    def __init__(self, text, sigma):
        self.text = np.ascontiguousarray(text)
        self.sigma = int(sigma)
        rows = lib().orc_blocked_rows(len(self.text))
        self.table = np.empty(rows * self.sigma, dtype=np.uint64)
        lib().orc_blocked_build(_p(self.text), self.text.itemsize, len(self.text),
                                self.sigma, _p(self.table))

    def rank_batch(self, idx, sym):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        sym = np.ascontiguousarray(sym, dtype=np.uint64)
        out = np.empty(len(idx), dtype=np.uint64)
        lib().orc_blocked_rank_batch(_p(self.text), self.text.itemsize, len(self.text),
                                     self.sigma, _p(self.table), _p(idx), _p(sym), len(idx),
                                     _p(out))
        return out

    def select_batch(self, k, sym):
        k = np.ascontiguousarray(k, dtype=np.uint64)
        sym = np.ascontiguousarray(sym, dtype=np.uint64)
        out = np.empty(len(k), dtype=np.uint64)
        lib().orc_blocked_select_batch(_p(self.text), self.text.itemsize, len(self.text),
                                       self.sigma, _p(self.table), _p(k), _p(sym), len(k),
                                       _p(out))
        return out
