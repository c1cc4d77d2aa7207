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

// Collects small host->device copies and issues them on a stream as
// kernel-argument payloads (launch_upload): the caller's host data may go
// away after add(), and nothing waits for earlier work of the stream.
// Copies above 64 KiB go through hipMemcpy after the pending ones.
class Uploader {
 public:
  explicit Uploader(hipStream_t s) : s_(s) { clear(); }
  template <typename T>
  void add(void* dst, const T* src, size_t count) {
    add_bytes(dst, src, count * sizeof(T));
  }
  void add_bytes(void* dst, const void* src, size_t bytes);
  void flush();  // launches what is pending

 private:
  void clear();
  hipStream_t s_;
  UploadBlob b_;
};

// Device scratch of a build, taken from a per-device cache that keeps the
// buffer between builds (unmapping a large buffer costs about as much as a
// construction kernel); a build that finds the cache in use by another
// thread allocates its own.  Sub-buffers are carved in 256-byte units.  The
// destructor waits for the stream before giving the buffer back.
class Scratch {
 public:
  Scratch(int device, hipStream_t s, size_t bytes);
  ~Scratch();
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  template <typename T>
  T* carve(size_t count) {
    const size_t b = (count * sizeof(T) + 255) & ~size_t(255);
    if (used_ + b > cap_) throw Error(ECL_ERROR_INTERNAL, "scratch layout exceeded");
    T* p = reinterpret_cast<T*>(base_ + used_);
    used_ += b;
    return p;
  }
  static size_t units(size_t bytes) { return (bytes + 255) & ~size_t(255); }

 private:
  int device_;
  hipStream_t s_;
  bool cached_ = false;
  DevBuf own_;
  uint8_t* base_ = nullptr;
  size_t used_ = 0, cap_ = 0;
};
// bytes the device's scratch cache holds (free for a build to reuse)
size_t cached_scratch_bytes(int device);
// frees the device's scratch cache (waits for a build that uses it)
void release_scratch_cache(int device);

// memcpy of large host buffers split over a small pool of host threads
// (one copy thread reaches a fraction of the memory bandwidth that PCIe
// transfers of pinned buffers need); small copies run on the caller.
void parallel_memcpy(void* dst, const void* src, size_t bytes);
// Host -> device copy of a large pageable buffer through two pinned staging
// buffers (kept for the process), overlapping the host copies with the
// transfers; returns after the data is on the device.
void upload_staged(void* d_dst, const void* src, size_t bytes, hipStream_t s);

// Element size, size and alphabet size of a tree file written by
// ecl_wavelet_tree_save (its header only; throws for other files).
void read_tree_file_header(const char* path, int* esize, uint64_t* n, uint64_t* sigma);

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
