"""MI355X (gfx950) wavelet-tree / rank-select library.

The product is ``libecl_wavelet_hip.so`` (hand-written HIP kernels behind the
C ABI in ``include/ecl_wavelet_c.h``).  This package builds it
(``build.build_library``), binds it (``_lib``) and mirrors the reference's
class names in Python (``api``).
"""
from .api import (BitArray, DeviceBuffer, Forest, Graph, Profiler, RankSelect,  # noqa: F401
                  Server, Timer, WaveletTree, convert_data_to_min_alphabet, device_name,
                  generate_text, read_data_from_file,
                  generate_text_into, mem_info, release_cached_memory, set_build_span,
                  set_device_memory_cap, tree_build_memory,
                  set_query_lanes, probe_random_lines,
                  zipf_thresholds,
                  synchronize, FOREST_QUERY_DTYPE, MIXED_DTYPE, QUERY_DTYPE)
