"""ecl_rank_select_build_directory_noalloc (the allocation-free directory
build) against ecl_rank_select_create over the same bits, through the C ABI:
L1, L2, both select-sample streams and the total must be identical, on two
calls in a row, the second over a smaller array, each into output arrays
filled with junk first (so every entry the call owns is shown written)."""
import ctypes as C

import numpy as np
import pytest

from ecl_wavelet_amd import _lib

pytestmark = pytest.mark.gpu

u64 = C.c_uint64


def _call(name, *args):
    _lib.call(name, *args)


def _reference_directory(size_bits, seed, threshold):
    """A BitArray of one array filled on the device, its RankSelect built by
    ecl_rank_select_create; returns (rs handle, device words pointer, L1, L2,
    ones samples, zeros samples, total)."""
    sizes = np.array([size_bits], dtype=np.uint64)
    ba = C.c_void_p()
    _call("ecl_bit_array_create", sizes.ctypes.data_as(C.c_void_p), 1, -1, 0, C.byref(ba))
    _call("ecl_bit_array_fill_random", ba, 0, seed, threshold)
    rs = C.c_void_p()
    _call("ecl_rank_select_create", C.byref(ba), 0, C.byref(rs))
    info = _lib.RankSelectInfo()
    _call("ecl_rank_select_info_get", rs, 0, C.byref(info))
    l1 = np.empty(info.num_l1_blocks, np.uint64)
    l2 = np.empty(info.num_l2_blocks, np.uint16)
    s1 = np.empty(info.num_samples_1, np.uint64)
    s0 = np.empty(info.num_samples_0, np.uint64)
    _call("ecl_rank_select_get_l1", rs, 0, 0, len(l1), l1.ctypes.data_as(C.c_void_p))
    _call("ecl_rank_select_get_l2", rs, 0, 0, len(l2), l2.ctypes.data_as(C.c_void_p))
    _call("ecl_rank_select_get_samples", rs, 0, 1, 0, len(s1), s1.ctypes.data_as(C.c_void_p))
    _call("ecl_rank_select_get_samples", rs, 0, 0, 0, len(s0), s0.ctypes.data_as(C.c_void_p))
    bah = C.c_void_p()
    _call("ecl_rank_select_bit_array", rs, C.byref(bah))
    data, offs, szs = C.c_void_p(), C.c_void_p(), C.c_void_p()
    _call("ecl_bit_array_device_pointers", bah, C.byref(data), C.byref(offs), C.byref(szs))
    off = C.c_size_t()
    _call("ecl_bit_array_offset", bah, 0, C.byref(off))
    words = C.c_void_p(data.value + 4 * off.value)
    return rs, words, l1, l2, s1, s0, info.total_ones


def _alloc_junk(nbytes):
    p = C.c_void_p()
    _call("ecl_device_alloc", 0, max(nbytes, 1), C.byref(p))
    junk = np.full(max(nbytes, 1), 0xAB, np.uint8)
    _call("ecl_memcpy_h2d", 0, p, junk.ctypes.data_as(C.c_void_p), junk.nbytes)
    return p


def _download(p, dtype, count):
    out = np.empty(count, dtype)
    if count:
        _call("ecl_memcpy_d2h", 0, out.ctypes.data_as(C.c_void_p), p, out.nbytes)
    return out


def _noalloc_directory(words, size_bits):
    n1, n2, ns = u64(), u64(), u64()
    _call("ecl_rank_select_directory_sizes", size_bits, C.byref(n1), C.byref(n2), C.byref(ns))
    d_l1 = _alloc_junk(8 * n1.value)
    d_l2 = _alloc_junk(2 * n2.value)
    d_s = _alloc_junk(8 * ns.value)
    d_t = _alloc_junk(8)
    print(f"[no_alloc] build directory of {size_bits} bits", flush=True)
    _call("ecl_rank_select_build_directory_noalloc", 0, words, size_bits, d_l1, d_l2, d_s,
          d_t, None)
    _call("ecl_device_synchronize", 0)
    out = (_download(d_l1, np.uint64, n1.value), _download(d_l2, np.uint16, n2.value),
           _download(d_s, np.uint64, ns.value), int(_download(d_t, np.uint64, 1)[0]))
    for p in (d_l1, d_l2, d_s, d_t):
        _call("ecl_device_free", 0, p)
    return out


@pytest.mark.parametrize("calls", [[(3_000_000_123, 0.5, 11), (1_000_077, 0.05, 12)]])
def test_noalloc_directory_equals_rank_select_create(gpu, calls):
    for size_bits, p, seed in calls:
        thr = min(int(p * 2.0 ** 64), 2 ** 64 - 1)
        rs, words, l1, l2, s1, s0, total = _reference_directory(size_bits, seed, thr)
        try:
            g1, g2, gs, gt = _noalloc_directory(words, size_bits)
        finally:
            _call("ecl_rank_select_destroy", rs)
        assert gt == total
        assert np.array_equal(g1, l1)
        assert np.array_equal(g2, l2)
        assert np.array_equal(gs[:len(s1)], s1)
        assert np.array_equal(gs[len(gs) - len(s0):][::-1], s0)
        assert len(s1) == total // 4096 and len(s0) == (size_bits - total) // 4096


def test_noalloc_rejects_oversize(gpu):
    with pytest.raises(_lib.EclError, match="ECL_NOALLOC_MAX_BITS"):
        _call("ecl_rank_select_build_directory_noalloc", 0, None, 2 ** 35 + 1, None, None,
              None, None, None)
