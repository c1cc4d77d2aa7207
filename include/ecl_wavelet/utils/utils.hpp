// ecl_wavelet/utils/utils.hpp -- drop-in for utils/utils.cuh of
// mfranzreb/ECL_wavelet on HIP / gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "ecl_wavelet/detail/device_core.hpp"
#include "ecl_wavelet_c.h"

// Wavefront width of the target (utils.cuh:41 defines 32 for CUDA warps).
#define WS 64

namespace ecl {

// RankSelectQuery<T> (utils.cuh:70-74): 16 bytes for every T
template <typename T>
struct RankSelectQuery {
  size_t index_;
  T symbol_;
};

namespace utils {

static constexpr uint8_t kMinAlphabetSize = 2;  // utils.cuh:61

// isPowTwo / getPrevPowTwo / powTwo / ceilLog2 / getBit (utils.cuh:123-240)
template <typename T>
__host__ __device__ inline bool isPowTwo(T const n) {
  static_assert(std::is_integral<T>::value && std::is_unsigned<T>::value,
                "T must be an unsigned integral type.");
  return (n & (n - 1)) == 0;
}
template <typename T>
__host__ __device__ inline T getPrevPowTwo(T n) {
  return static_cast<T>(detail::prev_pow2(static_cast<uint64_t>(n)));
}
template <typename T>
__host__ __device__ inline T powTwo(T n) {
  return static_cast<T>(1ull << n);
}
template <typename T>
__host__ __device__ inline uint8_t ceilLog2(T n) {
  return static_cast<uint8_t>(detail::ceil_log2(static_cast<uint64_t>(n)));
}
template <typename T>
__host__ __device__ inline bool getBit(uint8_t const i, T const c) {
  return (static_cast<uint64_t>(c) >> i) & 1u;
}

// Error reporting: the reference prints and exits on CUDA errors
// (utils.cuh:83-95); the drop-in throws std::runtime_error instead so that a
// failing call cannot take the host process down.
inline void throwOnError(ecl_status st) {
  if (st != ECL_OK) throw std::runtime_error(ecl_last_error());
}
inline void hipThrow(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// checkWarpSize (utils.cu:101-120): the wave64 target needs no check; kept
// for source compatibility, it only selects the device.
inline void checkWarpSize(uint32_t const GPU_index) {
  hipThrow(hipSetDevice(static_cast<int>(GPU_index)), "hipSetDevice");
}

}  // namespace utils
}  // namespace ecl

#define gpuErrchk(ans) ::ecl::utils::hipThrow((ans), #ans)
#define kernelCheck()                                               \
  do {                                                              \
    ::ecl::utils::hipThrow(hipGetLastError(), "kernel launch");     \
    ::ecl::utils::hipThrow(hipDeviceSynchronize(), "kernel");       \
  } while (0)
