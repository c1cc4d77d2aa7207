// Forwarding header: the reference include path tree/wavelet_tree.cuh maps to the
// HIP drop-in tree/wavelet_tree.hpp.
#pragma once
#include "ecl_wavelet/tree/wavelet_tree.hpp"
