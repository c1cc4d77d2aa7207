// ecl_wavelet/bitarray/bit_array.hpp -- drop-in for bitarray/bit_array.cuh
// (mfranzreb/ECL_wavelet).  Storage is owned by libecl_wavelet_hip.so through
// the C ABI; this class is the same shallow view the reference passes by
// value into kernels: copies share the device memory and never free it.
#pragma once

#include <memory>
#include <vector>

#include "ecl_wavelet/utils/utils.hpp"

namespace ecl {

class RankSelect;

class BitArray {
 public:
  BitArray() = default;
  BitArray& operator=(BitArray const&) = delete;

  // BitArray(sizes) (bit_array.cu:40-85): payload left uninitialised
  __host__ explicit BitArray(std::vector<size_t> const& sizes) { create(sizes, -1); }
  // BitArray(sizes, init_value) (bit_array.cu:87-92)
  __host__ BitArray(std::vector<size_t> const& sizes, bool const init_value) {
    create(sizes, init_value ? 1 : 0);
  }
  // copy = shallow view (bit_array.cu:94-101)
  __host__ BitArray(BitArray const& o)
      : handle_(o.handle_), host_(o.host_), d_data_(o.d_data_), d_offsets_(o.d_offsets_),
        d_bit_sizes_(o.d_bit_sizes_), num_arrays_(o.num_arrays_), is_copy_(true),
        owns_host_(false) {}
  __host__ BitArray(BitArray&& o) noexcept { *this = std::move(o); }
  __host__ BitArray& operator=(BitArray&& o) noexcept {
    release();
    handle_ = o.handle_; host_ = o.host_; d_data_ = o.d_data_; d_offsets_ = o.d_offsets_;
    d_bit_sizes_ = o.d_bit_sizes_; num_arrays_ = o.num_arrays_; is_copy_ = o.is_copy_;
    owns_host_ = o.owns_host_;
    o.is_copy_ = true;
    o.owns_host_ = false;
    return *this;
  }
  __host__ ~BitArray() { release(); }

  // ---- device accessors (bit_array.cuh:96-326)
  __device__ __forceinline__ bool access(size_t const array_index, size_t const index) const {
    return (d_data_[d_offsets_[array_index] + (index >> 5)] >> (index & 31)) & 1u;
  }
  __device__ __forceinline__ bool access(size_t const, size_t const index,
                                         size_t const offset) const {
    return (d_data_[offset + (index >> 5)] >> (index & 31)) & 1u;
  }
  __device__ __forceinline__ void writeWord(size_t const array_index, size_t const index,
                                            uint32_t const value) {
    d_data_[d_offsets_[array_index] + index] = value;
  }
  __device__ __forceinline__ void writeWordAtBit(size_t const array_index, size_t const index,
                                                 uint32_t const value) {
    d_data_[d_offsets_[array_index] + (index >> 5)] = value;
  }
  __device__ __forceinline__ void writeWordAtBit(size_t const, size_t const index,
                                                 uint32_t const value, size_t const offset) {
    d_data_[offset + (index >> 5)] = value;
  }
  __device__ __forceinline__ size_t getOffset(size_t const array_index) const {
    return d_offsets_[array_index];
  }
  __device__ __forceinline__ uint32_t word(size_t const array_index, size_t const index) const {
    return d_data_[d_offsets_[array_index] + index];
  }
  __device__ __forceinline__ uint32_t word(size_t const, size_t const index,
                                           size_t const offset) const {
    return d_data_[offset + index];
  }
  // two consecutive words as one little-endian u64, index even
  __device__ __forceinline__ uint64_t twoWords(size_t const array_index,
                                               size_t const index) const {
    return reinterpret_cast<const uint64_t*>(d_data_)[(d_offsets_[array_index] + index) / 2];
  }
  __device__ __forceinline__ uint64_t twoWords(size_t const, size_t const index,
                                               size_t const offset) const {
    return reinterpret_cast<const uint64_t*>(d_data_)[(offset + index) / 2];
  }
  __device__ __forceinline__ uint32_t wordAtBit(size_t const array_index,
                                                size_t const index) const {
    return d_data_[d_offsets_[array_index] + (index >> 5)];
  }
  __host__ __device__ __forceinline__ uint32_t partialWord(uint32_t w, uint8_t const bit_index) const {
    return bit_index >= 32 ? w : (w & ((1u << bit_index) - 1));
  }
  __host__ __device__ __forceinline__ uint64_t partialTwoWords(uint64_t w,
                                                               uint8_t const bit_index) const {
    return bit_index >= 64 ? w : (w & ((1ull << bit_index) - 1));
  }
  __host__ __device__ __forceinline__ size_t size(size_t const array_index) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return d_bit_sizes_[array_index];
#else
    return host_->sizes[array_index];
#endif
  }
  __device__ __forceinline__ size_t sizeInWords(size_t const array_index) const {
    return (d_bit_sizes_[array_index] + 31) / 32;
  }
  __host__ __device__ __forceinline__ uint8_t numArrays() const { return num_arrays_; }

  // BitArray::getNeededGPUMemory (bit_array.cu:140-151)
  __host__ static size_t getNeededGPUMemory(size_t const size, uint8_t const num_arrays) {
    return ecl_bit_array_needed_gpu_memory(size, num_arrays);
  }

  // host helpers (new): whole-word copies of one array
  __host__ void writeWordsFromHost(size_t array_index, size_t first, uint32_t const* words,
                                   size_t count) {
    utils::throwOnError(ecl_bit_array_write_words(handle_, array_index, first, words, count));
  }
  __host__ void readWordsToHost(size_t array_index, size_t first, uint32_t* words,
                                size_t count) const {
    utils::throwOnError(ecl_bit_array_read_words(handle_, array_index, first, words, count));
  }
  __host__ ecl_bit_array handle() const { return handle_; }

 private:
  friend class RankSelect;
  struct Host { std::vector<size_t> sizes; };

  __host__ void release() {
    if (!is_copy_ && handle_ != nullptr) ecl_bit_array_destroy(handle_);
    if (owns_host_) delete host_;
    handle_ = nullptr;
    host_ = nullptr;
    is_copy_ = true;
    owns_host_ = false;
  }
  __host__ void create(std::vector<size_t> const& sizes, int init) {
    int dev = 0;
    utils::hipThrow(hipGetDevice(&dev), "hipGetDevice");
    utils::throwOnError(ecl_bit_array_create(sizes.data(), sizes.size(), init, dev, &handle_));
    host_ = new Host{sizes};
    owns_host_ = true;
    num_arrays_ = static_cast<uint8_t>(sizes.size());
    is_copy_ = false;
    bind();
  }
  __host__ void bind() {
    uint64_t* offs = nullptr;
    uint64_t* sz = nullptr;
    utils::throwOnError(ecl_bit_array_device_pointers(handle_, &d_data_, &offs, &sz));
    static_assert(sizeof(size_t) == sizeof(uint64_t), "64-bit size_t");
    d_offsets_ = reinterpret_cast<size_t*>(offs);
    d_bit_sizes_ = reinterpret_cast<size_t*>(sz);
  }
  // view a handle owned elsewhere (the RankSelect / WaveletTree keeps it)
  __host__ void adopt(ecl_bit_array h) {
    release();
    handle_ = h;
    is_copy_ = true;
    owns_host_ = true;
    size_t na = 0;
    utils::throwOnError(ecl_bit_array_num_arrays(h, &na));
    host_ = new Host{};
    host_->sizes.resize(na);
    for (size_t i = 0; i < na; ++i) utils::throwOnError(ecl_bit_array_size(h, i, &host_->sizes[i]));
    num_arrays_ = static_cast<uint8_t>(na);
    bind();
  }

  ecl_bit_array handle_ = nullptr;
  Host* host_ = nullptr;
  uint32_t* d_data_ = nullptr;
  size_t* d_offsets_ = nullptr;
  size_t* d_bit_sizes_ = nullptr;
  uint8_t num_arrays_ = 0;
  bool is_copy_ = true;     // a view: never destroys the library object
  bool owns_host_ = false;  // frees the host-side size table
};

}  // namespace ecl
