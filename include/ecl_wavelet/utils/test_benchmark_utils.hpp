// ecl_wavelet/utils/test_benchmark_utils.hpp -- the raw-text helpers of
// utils/test_benchmark_utils.cuh (mfranzreb/ECL_wavelet, :588-679) for the
// HIP drop-in: readDataFromFile and convertDataToMinAlphabet, with the same
// signatures, results and exception messages, over the library's C ABI.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "ecl_wavelet/utils/utils.hpp"
#include "ecl_wavelet_c.h"

namespace ecl {
namespace utils {

// the first num_symbols elements of a raw little-endian file of T
template <typename T>
std::vector<T> readDataFromFile(std::string const& filename, size_t const num_symbols) {
  static_assert(std::is_unsigned_v<T>, "T must be an unsigned integer type");
  std::vector<T> data(num_symbols);
  throwOnError(ecl_read_text_file(filename.c_str(), static_cast<ecl_type>(sizeof(T)),
                                  num_symbols, data.data()));
  return data;
}

// every element becomes its index among the sorted distinct elements;
// returns the size of that alphabet
template <typename T>
size_t convertDataToMinAlphabet(T* data, size_t const data_size) {
  size_t sigma = 0;
  throwOnError(ecl_convert_to_min_alphabet(static_cast<ecl_type>(sizeof(T)), data, data_size,
                                           nullptr, 0, nullptr, &sigma));
  return sigma;
}

// every element becomes the index of its lower bound in the sorted alphabet
template <typename T>
size_t convertDataToMinAlphabet(T* data, size_t const data_size,
                                std::vector<T> const& alphabet) {
  size_t sigma = 0;
  throwOnError(ecl_convert_to_min_alphabet(static_cast<ecl_type>(sizeof(T)), data, data_size,
                                           alphabet.data(), alphabet.size(), nullptr, &sigma));
  return sigma;
}

}  // namespace utils
}  // namespace ecl
