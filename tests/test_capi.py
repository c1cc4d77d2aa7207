"""The C-ABI library loads and exports every symbol the header declares.
No compute call: runs without a GPU."""
import ctypes
import os

import numpy as np
import pytest

from ecl_wavelet_amd import _lib


def test_library_built():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _lib.header_symbols()
    assert len(declared) > 50
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []


def test_binding_covers_header():
    assert sorted(_lib._SIGS) == _lib.header_symbols()


def test_abi_version_and_no_device_error():
    lib = _lib.load()
    assert lib.ecl_abi_version() == 1
    n = ctypes.c_int(-1)
    assert lib.ecl_device_count(ctypes.byref(n)) == 0
    if n.value == 0:
        h = ctypes.c_void_p()
        import numpy as np
        sizes = np.array([100], dtype=np.uint64)
        rc = lib.ecl_bit_array_create(sizes.ctypes.data_as(ctypes.c_void_p), 1, 0, 0,
                                      ctypes.byref(h))
        assert rc == _lib.ECL_ERROR_NO_DEVICE
        assert b"no HIP device" in lib.ecl_last_error()


def test_needed_memory_formulas_monotone():
    lib = _lib.load()
    a = lib.ecl_wavelet_tree_needed_gpu_memory(10**6, 256, 8, 0)
    b = lib.ecl_wavelet_tree_needed_gpu_memory(10**7, 256, 8, 0)
    assert b > a > 0
    assert lib.ecl_rank_select_needed_gpu_memory(2**20, 4) > 0
    assert lib.ecl_bit_array_needed_gpu_memory(2**20, 4) >= 4 * 2**17


@pytest.mark.parametrize("name", ["ecl_wavelet_tree_access", "ecl_rank_select_select"])
def test_null_handle_is_invalid_argument(name):
    lib = _lib.load()
    f = getattr(lib, name)
    rc = f(None, None, 0, None, 0) if name == "ecl_wavelet_tree_access" else f(
        None, 0, 1, None, 0, None)
    assert rc == _lib.ECL_ERROR_INVALID_ARGUMENT


def test_directory_sizes_match_the_oracle():
    """ecl_rank_select_directory_sizes (host only) gives the RankSelect
    layout's entry counts (rank_select.cu:57-117) for edge sizes."""
    import oracle as O
    lib = _lib.load()
    for m in [0, 1, 511, 512, 513, 4095, 4096, 65535, 65536, 65537, 131071, 2**34 + 7]:
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        assert lib.ecl_rank_select_directory_sizes(m, ctypes.byref(a), ctypes.byref(b),
                                                   ctypes.byref(c)) == 0
        n1, n2 = O.rs_sizes(m)[:2] if m else (0, 0)
        assert (a.value, b.value, c.value) == (n1, n2, m // 4096), m


def test_noalloc_rejects_oversize_before_any_device_call():
    lib = _lib.load()
    rc = lib.ecl_rank_select_build_directory_noalloc(0, None, 2**35 + 1, None, None, None,
                                                     None, None)
    assert rc == _lib.ECL_ERROR_INVALID_ARGUMENT
    assert b"ECL_NOALLOC_MAX_BITS" in lib.ecl_last_error()


def test_tree_build_memory_estimate():
    """The build's memory split: the resident part does not depend on where
    the text is; a host text adds its copy to the staging part; multi-pass
    trees add a level-order buffer."""
    from ecl_wavelet_amd import tree_build_memory
    n = 10**7
    r_h, s_h = tree_build_memory(np.uint8, n, 200, host_text=True)
    r_d, s_d = tree_build_memory(np.uint8, n, 200, host_text=False)
    assert r_h == r_d and r_h > n  # bits of 8 levels + directory + headroom
    assert s_h - s_d == n
    assert s_d >= n  # the mapped codes
    r16, s16 = tree_build_memory(np.uint16, n, 65536, host_text=False)
    assert s16 >= 2 * n * 2  # codes and one level-order buffer of 2-byte codes
    assert r16 > r_d
    with pytest.raises(Exception):
        tree_build_memory(np.uint8, n, 1)
