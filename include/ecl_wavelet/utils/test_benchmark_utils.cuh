// Forwarding header: the reference include path
// utils/test_benchmark_utils.cuh maps to the HIP drop-in
// utils/test_benchmark_utils.hpp (the raw-text helpers).
#pragma once
#include "ecl_wavelet/utils/test_benchmark_utils.hpp"
