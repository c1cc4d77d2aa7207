// Forwarding header: the reference include path bitarray/rank_select.cuh maps to the
// HIP drop-in bitarray/rank_select.hpp.
#pragma once
#include "ecl_wavelet/bitarray/rank_select.hpp"
