// probe.hip -- the device's rate of random, dependent 128-byte line reads,
// the bound of the tree queries: every probe query reads `steps` lines one
// after the other, each at a position derived from the previous line's
// contents, as a descent reads one line per level.  With `lanes` lanes per
// query each lane fetches 128 / lanes bytes of the line (the query kernels
// use two).  No counterpart in the reference; it prices the query kernels.
#include <hip/hip_runtime.h>

#include "host.hpp"
#include "internal.hpp"

namespace ecl {
namespace lib {
namespace {

template <int G>
__global__ __launch_bounds__(256) void random_lines_kernel(const uint4* __restrict__ buf,
                                                           uint64_t nlines, uint64_t queries,
                                                           uint32_t steps,
                                                           uint64_t* __restrict__ out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256 / G;
  const uint32_t t = threadIdx.x % G;
  for (uint64_t q = (static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x) / G; q < queries;
       q += stride) {
    uint64_t x = detail::hash64(0x5EEDull, q);
    for (uint32_t l = 0; l < steps; ++l) {
      const uint64_t line = __umul64hi(x, nlines);  // < nlines
      const uint4* p = buf + line * 8 + t * (8 / G);
      uint32_t s = 0;
#pragma unroll
      for (int u = 0; u < 8 / G; ++u) {
        const uint4 v = p[u];
        s ^= v.x ^ v.y ^ v.z ^ v.w;
      }
#pragma unroll
      for (int m = 1; m < G; m <<= 1) s ^= __shfl_xor(s, m, G);
      x = detail::hash64(x ^ s, l);
    }
    if (t == 0) out[q] = x;
  }
}

template <int G>
void launch_random_lines(const uint4* buf, uint64_t nlines, uint64_t queries, uint32_t steps,
                         uint64_t* out, hipStream_t s) {
  uint64_t g = (queries * G + 255) / 256;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(random_lines_kernel<G>, dim3(static_cast<uint32_t>(g)), dim3(256), 0, s, buf,
                     nlines, queries, steps, out);
  ECL_HIP(hipGetLastError());
}

}  // namespace
}  // namespace lib
}  // namespace ecl

using namespace ecl::lib;

extern "C" ecl_status ecl_probe_random_lines(int device, uint64_t bytes, uint64_t queries,
                                             uint32_t steps, uint32_t lanes,
                                             double* lines_per_s) {
  return guard([&] {
    require(lines_per_s, "NULL argument");
    require(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8, "lanes must be 1, 2, 4 or 8");
    require(bytes >= 128 && queries >= 1 && steps >= 1, "invalid probe size");
    set_device(device);
    const uint64_t nlines = bytes / 128;
    if (nlines * 128 + queries * 8 + (64u << 20) > free_device_memory())
      throw Error(ECL_ERROR_OUT_OF_MEMORY_QUERIES, "Not enough GPU memory available for the probe.");
    DevBuf buf(nlines * 128), out(queries * 8);
    hipStream_t s = nullptr;
    ECL_HIP(hipMemsetAsync(buf.ptr, 0x5A, nlines * 128, s));
    auto run = [&] {
      const uint4* b = buf.as<uint4>();
      uint64_t* o = out.as<uint64_t>();
      switch (lanes) {
        case 1: launch_random_lines<1>(b, nlines, queries, steps, o, s); break;
        case 2: launch_random_lines<2>(b, nlines, queries, steps, o, s); break;
        case 4: launch_random_lines<4>(b, nlines, queries, steps, o, s); break;
        default: launch_random_lines<8>(b, nlines, queries, steps, o, s); break;
      }
    };
    run();  // warm
    hipEvent_t e0, e1;
    ECL_HIP(hipEventCreate(&e0));
    ECL_HIP(hipEventCreate(&e1));
    float ms = 0.f;
    hipError_t err = hipEventRecord(e0, s);
    if (err == hipSuccess) {
      run();
      err = hipEventRecord(e1, s);
    }
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ECL_HIP(err);
    *lines_per_s = static_cast<double>(queries) * steps / (ms / 1e3);
  });
}
