// Forwarding header: the reference include path bitarray/bit_array.cuh maps to the
// HIP drop-in bitarray/bit_array.hpp.
#pragma once
#include "ecl_wavelet/bitarray/bit_array.hpp"
