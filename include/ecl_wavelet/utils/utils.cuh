// Forwarding header: the reference include path utils/utils.cuh maps to the
// HIP drop-in utils/utils.hpp.
#pragma once
#include "ecl_wavelet/utils/utils.hpp"
