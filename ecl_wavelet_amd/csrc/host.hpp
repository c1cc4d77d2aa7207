// host.hpp -- host-side helpers shared by the C ABI translation units:
// error guard, copies, and the shape / metadata of a wavelet tree.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"

namespace ecl {
namespace lib {

// message of the last failed call on this thread (ecl_last_error)
std::string& last_error();

// Runs f and converts exceptions into an ecl_status.
template <typename F>
ecl_status guard(F&& f) {
  try {
    f();
    return ECL_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    last_error() = "host allocation failed";
    return ECL_ERROR_INTERNAL;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return ECL_ERROR_INTERNAL;
  }
}

inline void require(bool ok, const char* msg) {
  if (!ok) throw Error(ECL_ERROR_INVALID_ARGUMENT, msg);
}
void set_device(int device);
size_t free_device_memory();

template <typename T>
void h2d(void* dst, const T* src, size_t count) {
  if (count) ECL_HIP(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
}
template <typename T>
void d2h(T* dst, const void* src, size_t count) {
  if (count) ECL_HIP(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDeviceToHost));
}

// ---- tree shape: minimal codes and the internal nodes of every level
struct Node { uint64_t lo, hi; };
struct TreeShape {
  uint64_t sigma = 0;
  uint32_t L = 0;
  bool pow2 = true;
  uint64_t codes_start = 0;                // first symbol with its own code entry
  std::vector<uint64_t> code;              // L-bit left-aligned code per symbol
  std::vector<uint8_t> len;                // code length per symbol
  std::vector<std::vector<Node>> nodes;    // internal nodes per level
};
TreeShape tree_shape(uint64_t sigma);

// ---- metadata from the symbol counts: exclusive prefix C, node ranks
// (ordered by level, then symbol), node starts and levels, nodes per level
struct TreeMeta {
  std::vector<uint64_t> counts, ranks, node_start, nodes_at_level;
  std::vector<uint8_t> node_level;
  uint64_t nodes_until_last_level = 0;
};
TreeMeta tree_meta(const TreeShape& sh, const std::vector<uint64_t>& sym_count, uint64_t n);
// level l holds the codes longer than l (wavelet_tree.cuh:776-790)
std::vector<uint64_t> level_sizes(const TreeShape& sh, const std::vector<uint64_t>& sym_count);

}  // namespace lib
}  // namespace ecl
