"""ctypes binding of libecl_wavelet_hip.so (the C ABI in include/ecl_wavelet_c.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os
import re

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libecl_wavelet_hip.so")
HEADER = os.path.join(os.path.dirname(_PKG), "include", "ecl_wavelet_c.h")

vp = C.c_void_p
u64 = C.c_uint64
u32 = C.c_uint32
sz = C.c_size_t
i32 = C.c_int
st = C.c_int  # ecl_status

ECL_OK = 0
ECL_ERROR_INVALID_ARGUMENT = 1
ECL_ERROR_OUT_OF_MEMORY_TREE = 2
ECL_ERROR_OUT_OF_MEMORY_QUERIES = 3
ECL_ERROR_HIP = 4
ECL_ERROR_NO_DEVICE = 5
ECL_ERROR_INTERNAL = 6
ECL_SYMBOLS_MINIMAL = 1
ECL_RESULTS_MINIMAL = 2


class EclError(RuntimeError):
    """Error raised for a non-OK ecl_status; str() is the library's message
    (the reference's exact std::runtime_error text for the OOM cases)."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class RankSelectInfo(C.Structure):
    _fields_ = [(n, u64) for n in (
        "size_bits", "offset_words", "num_l1_blocks", "num_l2_blocks",
        "num_last_l2_blocks", "total_ones", "total_zeros", "num_samples_1",
        "num_samples_0")]


class TreeInfo(C.Structure):
    _fields_ = [
        ("data_size", u64), ("alphabet_size", u64), ("num_levels", u32),
        ("is_min_alphabet", u32), ("codes_start", u64), ("num_ranks", u64),
        ("nodes_until_last_level", u64), ("element_size", u32),
        ("is_pow_two", u32)]


class RankSelectDevice(C.Structure):
    _fields_ = [
        ("l1", vp), ("l2", vp), ("l1_offsets", vp), ("l2_offsets", vp),
        ("num_l1_blocks", vp), ("num_last_l2_blocks", vp), ("samples", vp),
        ("sample_offsets", vp), ("sample_caps", vp), ("total_ones", vp),
        ("levels", vp), ("total_num_l2_blocks", u64)]


_SIGS = {
    "ecl_last_error": (C.c_char_p, []),
    "ecl_abi_version": (i32, []),
    "ecl_device_count": (st, [C.POINTER(i32)]),
    "ecl_device_alloc": (st, [i32, sz, C.POINTER(vp)]),
    "ecl_device_free": (st, [i32, vp]),
    "ecl_memcpy_h2d": (st, [i32, vp, vp, sz]),
    "ecl_memcpy_d2h": (st, [i32, vp, vp, sz]),
    "ecl_device_synchronize": (st, [i32]),
    "ecl_device_mem_info": (st, [i32, C.POINTER(sz), C.POINTER(sz)]),
    "ecl_release_cached_memory": (st, [i32]),
    "ecl_set_device_memory_cap": (st, [i32, u64]),
    "ecl_wavelet_tree_build_memory": (st, [i32, sz, sz, i32, C.POINTER(u64), C.POINTER(u64)]),
    "ecl_wavelet_tree_build_host_bytes": (st, [vp, C.POINTER(u64)]),
    "ecl_set_build_span": (st, [C.c_uint32]),
    "ecl_set_query_lanes": (st, [C.c_uint32]),
    "ecl_probe_random_lines": (st, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_double)]),
    "ecl_device_name": (st, [i32, C.c_char_p, sz]),
    "ecl_event_create": (st, [C.POINTER(vp)]),
    "ecl_event_destroy": (st, [vp]),
    "ecl_event_record": (st, [vp, vp]),
    "ecl_event_elapsed": (st, [vp, vp, C.POINTER(C.c_float)]),
    "ecl_profile_enable": (st, [i32]),
    "ecl_profile_reset": (st, []),
    "ecl_profile_get": (st, [C.c_char_p, C.POINTER(u64), C.POINTER(C.c_double)]),
    "ecl_generate_text": (st, [i32, i32, u64, sz, u64, vp, vp]),
    "ecl_bit_array_create": (st, [vp, sz, i32, i32, C.POINTER(vp)]),
    "ecl_bit_array_destroy": (st, [vp]),
    "ecl_bit_array_num_arrays": (st, [vp, C.POINTER(sz)]),
    "ecl_bit_array_size": (st, [vp, sz, C.POINTER(sz)]),
    "ecl_bit_array_offset": (st, [vp, sz, C.POINTER(sz)]),
    "ecl_bit_array_device_pointers": (st, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    "ecl_bit_array_write_words": (st, [vp, sz, sz, vp, sz]),
    "ecl_bit_array_read_words": (st, [vp, sz, sz, vp, sz]),
    "ecl_bit_array_fill_random": (st, [vp, sz, u64, u64]),
    "ecl_bit_array_needed_gpu_memory": (sz, [sz, C.c_uint8]),
    "ecl_rank_select_create": (st, [C.POINTER(vp), i32, C.POINTER(vp)]),
    "ecl_rank_select_destroy": (st, [vp]),
    "ecl_rank_select_bit_array": (st, [vp, C.POINTER(vp)]),
    "ecl_rank_select_info_get": (st, [vp, sz, C.POINTER(RankSelectInfo)]),
    "ecl_rank_select_get_l1": (st, [vp, sz, sz, sz, vp]),
    "ecl_rank_select_get_l2": (st, [vp, sz, sz, sz, vp]),
    "ecl_rank_select_get_samples": (st, [vp, sz, i32, sz, sz, vp]),
    "ecl_rank_select_device_pointers": (st, [vp, C.POINTER(RankSelectDevice)]),
    "ecl_rank_select_rank_device": (st, [vp, sz, i32, vp, sz, vp, vp]),
    "ecl_rank_select_select_device": (st, [vp, sz, i32, vp, sz, vp, vp]),
    "ecl_rank_select_rank": (st, [vp, sz, i32, vp, sz, vp]),
    "ecl_rank_select_select": (st, [vp, sz, i32, vp, sz, vp]),
    "ecl_rank_select_rank_generated": (st, [vp, sz, i32, u64, sz, vp, vp]),
    "ecl_rank_select_select_generated": (st, [vp, sz, i32, u64, sz, vp, vp]),
    "ecl_rank_select_needed_gpu_memory": (sz, [sz, C.c_uint8]),
    "ecl_rank_select_directory_sizes": (st, [u64, C.POINTER(u64), C.POINTER(u64),
                                             C.POINTER(u64)]),
    "ecl_rank_select_build_directory_noalloc": (st, [i32, vp, u64, vp, vp, vp, vp, vp]),
    "ecl_wavelet_tree_create": (st, [i32, vp, sz, vp, sz, i32, C.POINTER(vp)]),
    "ecl_wavelet_tree_create_device": (st, [i32, vp, sz, vp, sz, i32, C.POINTER(vp)]),
    "ecl_read_text_file": (st, [C.c_char_p, i32, sz, vp]),
    "ecl_convert_to_min_alphabet": (st, [i32, vp, sz, vp, sz, vp, C.POINTER(sz)]),
    "ecl_wavelet_tree_create_from_file": (st, [i32, C.c_char_p, sz, vp, sz, i32,
                                               C.POINTER(vp)]),
    "ecl_server_add_tree_text_file": (st, [vp, C.c_char_p, sz, vp, C.POINTER(u64),
                                           C.POINTER(u32)]),
    "ecl_wavelet_tree_destroy": (st, [vp]),
    "ecl_wavelet_tree_build_query_lines": (st, [vp]),
    "ecl_wavelet_tree_build_select_lines": (st, [vp]),
    "ecl_rank_select_build_query_lines": (st, [vp]),
    "ecl_rank_select_build_select_lines": (st, [vp]),
    "ecl_wavelet_tree_save": (st, [vp, C.c_char_p]),
    "ecl_wavelet_tree_load": (st, [C.c_char_p, i32, C.POINTER(vp)]),
    "ecl_wavelet_tree_info_get": (st, [vp, C.POINTER(TreeInfo)]),
    "ecl_wavelet_tree_alphabet": (st, [vp, vp]),
    "ecl_wavelet_tree_counts": (st, [vp, vp]),
    "ecl_wavelet_tree_codes": (st, [vp, vp, vp]),
    "ecl_wavelet_tree_node_ranks": (st, [vp, vp, vp, vp]),
    "ecl_wavelet_tree_nodes_at_level": (st, [vp, vp]),
    "ecl_wavelet_tree_rank_select": (st, [vp, C.POINTER(vp)]),
    "ecl_wavelet_tree_device_view": (st, [vp, vp, sz]),
    "ecl_wavelet_tree_needed_gpu_memory": (sz, [sz, sz, C.c_uint8, sz]),
    "ecl_wavelet_tree_access": (st, [vp, vp, sz, vp, u32]),
    "ecl_wavelet_tree_rank": (st, [vp, vp, sz, vp, u32]),
    "ecl_wavelet_tree_select": (st, [vp, vp, sz, vp, u32]),
    "ecl_wavelet_tree_access_device": (st, [vp, vp, sz, vp, u32, vp]),
    "ecl_wavelet_tree_rank_device": (st, [vp, vp, sz, vp, u32, vp]),
    "ecl_wavelet_tree_select_device": (st, [vp, vp, sz, vp, u32, vp]),
    "ecl_wavelet_tree_query_batch_device": (st, [vp, vp, sz, vp, vp]),
    "ecl_wavelet_tree_generate_queries": (st, [vp, i32, u64, sz, vp, vp]),
    "ecl_wavelet_tree_set_query_lanes": (st, [vp, u32]),
    "ecl_wavelet_tree_build_quad_lines": (st, [vp]),
    "ecl_wavelet_tree_build_quad_select_lines": (st, [vp]),
    "ecl_wavelet_tree_build_hex_lines": (st, [vp]),
    "ecl_wavelet_tree_layout_views": (st, [vp, vp, vp]),
    "ecl_wavelet_tree_autotune": (st, [vp, u64, u64, C.POINTER(u32), vp]),
    "ecl_forest_create": (st, [vp, vp, vp, vp, sz, i32, u32, C.POINTER(vp)]),
    "ecl_forest_destroy": (st, [vp]),
    "ecl_forest_build_graph": (st, [vp, C.POINTER(vp)]),
    "ecl_forest_autotune_build": (st, [vp, C.POINTER(u32), C.POINTER(C.c_float)]),
    "ecl_forest_build_query_lines": (st, [vp]),
    "ecl_forest_build_select_lines": (st, [vp]),
    "ecl_forest_build_quad_layouts": (st, [vp]),
    "ecl_forest_num_trees": (st, [vp, C.POINTER(sz)]),
    "ecl_forest_tree_info": (st, [vp, sz, C.POINTER(u64), C.POINTER(u64), C.POINTER(u32)]),
    "ecl_forest_device_view": (st, [vp, sz, vp, sz]),
    "ecl_forest_query_device": (st, [vp, vp, sz, vp, vp]),
    "ecl_forest_generate_queries": (st, [vp, vp, u64, vp, vp]),
    "ecl_forest_autotune": (st, [vp, vp, sz, vp, C.POINTER(u32), C.POINTER(u32)]),
    "ecl_forest_graph_create": (st, [vp, vp, sz, vp, C.POINTER(vp)]),
    "ecl_graph_launch": (st, [vp, vp]),
    "ecl_graph_destroy": (st, [vp]),
    "ecl_server_create": (st, [i32, vp, C.POINTER(vp)]),
    "ecl_server_destroy": (st, [vp]),
    "ecl_server_add_tree": (st, [vp, vp, i32, sz, u64, C.POINTER(u32)]),
    "ecl_server_add_tree_file": (st, [vp, C.c_char_p, C.POINTER(C.c_uint32)]),
    "ecl_server_remove_tree": (st, [vp, u32]),
    "ecl_server_submit": (st, [vp, u32, vp, sz, vp, C.POINTER(u64)]),
    "ecl_server_wait": (st, [vp, u64, C.POINTER(C.c_double)]),
    "ecl_server_set_mode": (st, [vp, u32]),
    "ecl_server_stats_get": (st, [vp, vp]),
    "ecl_server_stats_reset": (st, [vp]),
    "ecl_server_load": (st, [vp, vp, vp]),
}

_lib = None


def header_symbols():
    """Every ecl_* function declared in include/ecl_wavelet_c.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ecl_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load the library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    L = load()
    rc = getattr(L, name)(*args)
    if rc != ECL_OK:
        msg = L.ecl_last_error().decode(errors="replace")
        raise EclError(rc, msg)
    return rc


def device_count():
    n = i32(0)
    call("ecl_device_count", C.byref(n))
    return n.value
