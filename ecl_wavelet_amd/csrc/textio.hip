// textio.hip -- raw text files and the minimal alphabet (host code).
//
// ecl_read_text_file restates readDataFromFile and
// ecl_convert_to_min_alphabet restates convertDataToMinAlphabet
// (include/ecl_wavelet/utils/test_benchmark_utils.cuh:588-679): the same
// checks in the same order with the same messages, and the same mapping
// (every symbol becomes its index in the sorted set of distinct symbols, or
// in a given sorted alphabet by lower bound).  The conversion runs on the
// host cores: one presence table per thread for 1- and 2-byte symbols,
// sorted distinct runs merged for wider ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "host.hpp"

using namespace ecl::lib;

namespace {

unsigned host_threads(uint64_t n) {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 1;
  if (t > 32) t = 32;
  const uint64_t by_size = n / (1u << 20) + 1;  // at least 1 Mi symbols a thread
  return static_cast<unsigned>(std::min<uint64_t>(t, by_size));
}

template <typename F>
void parallel_ranges(uint64_t n, F f) {
  const unsigned T = host_threads(n);
  std::vector<std::thread> th;
  for (unsigned t = 1; t < T; ++t)
    th.emplace_back([&, t] { f(t, n * t / T, n * (t + 1) / T); });
  f(0, 0, n / T);
  for (auto& x : th) x.join();
}

// distinct symbols of a 1- or 2-byte text, ascending
template <typename T>
std::vector<uint64_t> distinct_small(const T* d, uint64_t n) {
  constexpr size_t R = size_t(1) << (8 * sizeof(T));
  const unsigned T_ = host_threads(n);
  std::vector<std::vector<uint8_t>> seen(T_, std::vector<uint8_t>(R, 0));
  parallel_ranges(n, [&](unsigned t, uint64_t lo, uint64_t hi) {
    uint8_t* s = seen[t].data();
    for (uint64_t i = lo; i < hi; ++i) s[d[i]] = 1;
  });
  std::vector<uint64_t> out;
  for (size_t v = 0; v < R; ++v) {
    bool any = false;
    for (unsigned t = 0; t < T_ && !any; ++t) any = seen[t][v] != 0;
    if (any) out.push_back(v);
  }
  return out;
}

template <typename T>
std::vector<uint64_t> distinct_wide(const T* d, uint64_t n) {
  const unsigned T_ = host_threads(n);
  std::vector<std::vector<T>> part(T_);
  parallel_ranges(n, [&](unsigned t, uint64_t lo, uint64_t hi) {
    std::vector<T> v(d + lo, d + hi);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    part[t] = std::move(v);
  });
  std::vector<T> all;
  for (auto& p : part) all.insert(all.end(), p.begin(), p.end());
  std::sort(all.begin(), all.end());
  all.erase(std::unique(all.begin(), all.end()), all.end());
  return std::vector<uint64_t>(all.begin(), all.end());
}

// data[i] = index of the first alphabet entry >= data[i] (lower bound)
template <typename T>
void remap(T* d, uint64_t n, const std::vector<uint64_t>& alpha) {
  if constexpr (sizeof(T) <= 2) {
    const size_t R = size_t(1) << (8 * sizeof(T));
    std::vector<T> lut(R);
    for (size_t v = 0; v < R; ++v)
      lut[v] = static_cast<T>(std::lower_bound(alpha.begin(), alpha.end(), v) - alpha.begin());
    parallel_ranges(n, [&](unsigned, uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; ++i) d[i] = lut[d[i]];
    });
  } else {
    parallel_ranges(n, [&](unsigned, uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; ++i)
        d[i] = static_cast<T>(
            std::lower_bound(alpha.begin(), alpha.end(), static_cast<uint64_t>(d[i])) -
            alpha.begin());
    });
  }
}

template <typename T>
uint64_t convert(T* d, uint64_t n, const void* alphabet, size_t alphabet_size, void* alpha_out) {
  std::vector<uint64_t> alpha;
  if (alphabet != nullptr) {
    const T* a = static_cast<const T*>(alphabet);
    alpha.assign(a, a + alphabet_size);
    require(std::is_sorted(alpha.begin(), alpha.end()), "the alphabet must be sorted");
  } else {
    if constexpr (sizeof(T) <= 2) alpha = distinct_small(d, n);
    else alpha = distinct_wide(d, n);
  }
  remap(d, n, alpha);
  if (alpha_out != nullptr) {
    T* o = static_cast<T*>(alpha_out);
    for (size_t i = 0; i < alpha.size(); ++i) o[i] = static_cast<T>(alpha[i]);
  }
  return alpha.size();
}

}  // namespace

namespace ecl {
namespace lib {
// readDataFromFile (test_benchmark_utils.cuh:588-619)
void read_text_file(const char* path, int esize, uint64_t num_symbols, void* out) {
  require(path != nullptr, "path is NULL");
  const std::string name(path);
  std::FILE* f = std::fopen(path, "rb");
  if (!f) throw Error(ECL_ERROR_INVALID_ARGUMENT, "Failed to open file: " + name);
  struct Closer {
    std::FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  if (std::fseek(f, 0, SEEK_END) != 0)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "Failed to get file size: " + name);
  const long long size = std::ftell(f);
  if (size < 0) throw Error(ECL_ERROR_INVALID_ARGUMENT, "Failed to get file size: " + name);
  if (num_symbols * static_cast<uint64_t>(esize) > static_cast<uint64_t>(size))
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "Data size is larger than file size");
  if (static_cast<uint64_t>(size) % static_cast<uint64_t>(esize) != 0)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "File size is not a multiple of the type size");
  if (std::fseek(f, 0, SEEK_SET) != 0)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "Failed to read file: " + name);
  const uint64_t bytes = num_symbols * static_cast<uint64_t>(esize);
  if (bytes && std::fread(out, 1, bytes, f) != bytes)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "Failed to read file: " + name);
}

uint64_t convert_to_min_alphabet(int esize, void* data, uint64_t n, const void* alphabet,
                                 size_t alphabet_size, void* alphabet_out) {
  switch (esize) {
    case 1: return convert(static_cast<uint8_t*>(data), n, alphabet, alphabet_size, alphabet_out);
    case 2: return convert(static_cast<uint16_t*>(data), n, alphabet, alphabet_size, alphabet_out);
    case 4: return convert(static_cast<uint32_t*>(data), n, alphabet, alphabet_size, alphabet_out);
    case 8: return convert(static_cast<uint64_t*>(data), n, alphabet, alphabet_size, alphabet_out);
    default: throw Error(ECL_ERROR_INVALID_ARGUMENT, "invalid element type");
  }
}
}  // namespace lib
}  // namespace ecl

extern "C" {

ecl_status ecl_read_text_file(const char* path, ecl_type type, size_t num_symbols, void* out) {
  return guard([&] {
    require(out != nullptr || num_symbols == 0, "out is NULL");
    read_text_file(path, static_cast<int>(type), num_symbols, out);
  });
}

ecl_status ecl_convert_to_min_alphabet(ecl_type type, void* data, size_t n, const void* alphabet,
                                       size_t alphabet_size, void* alphabet_out,
                                       size_t* sigma) {
  return guard([&] {
    require(data != nullptr || n == 0, "data is NULL");
    require(sigma != nullptr, "sigma is NULL");
    *sigma = convert_to_min_alphabet(static_cast<int>(type), data, n, alphabet, alphabet_size,
                                     alphabet_out);
  });
}

ecl_status ecl_wavelet_tree_create_from_file(ecl_type type, const char* path, size_t num_symbols,
                                             const void* alphabet, size_t alphabet_size,
                                             int device, ecl_wavelet_tree* out) {
  return guard([&] {
    require(out != nullptr, "out is NULL");
    const int esize = static_cast<int>(type);
    require(esize == 1 || esize == 2 || esize == 4 || esize == 8, "invalid element type");
    std::vector<uint8_t> text(num_symbols * static_cast<size_t>(esize));
    read_text_file(path, esize, num_symbols, text.data());
    const ecl_status st =
        ecl_wavelet_tree_create(type, text.data(), num_symbols, alphabet, alphabet_size, device, out);
    if (st != ECL_OK) throw Error(st, last_error());
  });
}

ecl_status ecl_server_add_tree_text_file(ecl_server s, const char* path, size_t num_symbols,
                                         uint8_t* alphabet_out, uint64_t* sigma,
                                         uint32_t* tree) {
  return guard([&] {
    require(s != nullptr && tree != nullptr && sigma != nullptr, "NULL argument");
    std::vector<uint8_t> text(num_symbols);
    read_text_file(path, 1, num_symbols, text.data());
    uint8_t alpha[256];
    const uint64_t sg = convert_to_min_alphabet(1, text.data(), num_symbols, nullptr, 0, alpha);
    require(sg >= 2, "the text needs at least 2 distinct symbols");
    const ecl_status st = ecl_server_add_tree(s, text.data(), 0, num_symbols, sg, tree);
    if (st != ECL_OK) throw Error(st, last_error());
    if (alphabet_out) std::memcpy(alphabet_out, alpha, sg);
    *sigma = sg;
  });
}

}  // extern "C"
