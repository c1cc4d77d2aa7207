// noalloc.hip -- the RankSelect directory build for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// ecl_rank_select_build_directory_noalloc computes exactly what
// ecl_rank_select_create computes for one bit array (L1, L2, select samples
// and the total, rank_select.cu:50-304), with the same kernels, into device
// arrays the caller owns.  The build's scratch -- the array's descriptor and
// two u32 per 65536-bit L1 block (the block's array, the block's ones) --
// lives in the file-scope array `scratch` below instead of a per-call or
// cached allocation.  Nothing in it has to start at zero: the kernels write
// every slot they read within the call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "host.hpp"

namespace {

using namespace ecl::lib;
namespace detail = ecl::detail;

constexpr uint64_t kMaxBlocks = ECL_NOALLOC_MAX_BITS / detail::kL1Bits;  // 2^19
constexpr size_t kDescBytes = 256;  // one RsArrayDesc, rounded up
constexpr size_t kScratchBytes = kDescBytes + 2 * kMaxBlocks * sizeof(uint32_t);
static_assert(sizeof(RsArrayDesc) <= kDescBytes, "descriptor slot too small");

// 4 MiB + 256 B per device, held until the process ends.  In host code this
// name is a stand-in, not a device address: the kernels get the address
// hipGetSymbolAddress returns for the current device (scratch_of).
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on `device`, looked up once per device
// with that device current
uint8_t* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[device] == nullptr) {
    void* p = nullptr;
    ECL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = static_cast<uint8_t*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

ecl_status ecl_rank_select_directory_sizes(uint64_t size_bits, uint64_t* num_l1,
                                           uint64_t* num_l2, uint64_t* sample_slots) {
  return guard([&] {
    require(num_l1 != nullptr && num_l2 != nullptr && sample_slots != nullptr,
            "NULL output");
    const uint64_t m = size_bits;
    *num_l1 = (m + detail::kL1Bits - 1) / detail::kL1Bits;
    uint64_t last = ((m % detail::kL1Bits) + detail::kL2Bits - 1) / detail::kL2Bits;
    if (last == 0 && m) last = detail::kL2PerL1;
    *num_l2 = m ? (*num_l1 - 1) * detail::kL2PerL1 + last : 0;
    *sample_slots = m / detail::kSampleRate;
  });
}

ecl_status ecl_rank_select_build_directory_noalloc(int device, const uint32_t* d_words,
                                                   uint64_t size_bits, uint64_t* d_l1,
                                                   uint16_t* d_l2, uint64_t* d_samples,
                                                   uint64_t* d_total_ones, void* stream) {
  return guard([&] {
    if (size_bits > ECL_NOALLOC_MAX_BITS)
      throw Error(ECL_ERROR_INVALID_ARGUMENT,
                  "ecl_rank_select_build_directory_noalloc: the bit array exceeds "
                  "ECL_NOALLOC_MAX_BITS (2^35 bits), the size its fixed scratch holds");
    require(device >= 0 && device < kMaxDevices, "invalid device");
    require(d_total_ones != nullptr, "d_total_ones is NULL");
    set_device(device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t* base = scratch_of(device);
    if (size_bits == 0) {
      launch_fill_u64(d_total_ones, 1, 0, s);
      ECL_HIP(hipGetLastError());
      return;
    }
    require(d_words != nullptr && d_l1 != nullptr && d_l2 != nullptr, "NULL device array");
    uint64_t num_l1 = 0, num_l2 = 0, slots = 0;
    (void)ecl_rank_select_directory_sizes(size_bits, &num_l1, &num_l2, &slots);
    require(slots == 0 || d_samples != nullptr, "d_samples is NULL");
    RsArrayDesc d{};
    d.words64 = reinterpret_cast<const uint64_t*>(d_words);
    d.m = size_bits;
    d.l1 = d_l1;
    d.l2 = d_l2;
    d.num_l1 = num_l1;
    d.num_l2 = num_l2;
    d.samples = d_samples;
    d.cap = slots;
    d.total_ones = d_total_ones;
    d.first_tile = 0;
    d.first_chunk = 0;
    RsArrayDesc* d_desc = reinterpret_cast<RsArrayDesc*>(base);
    uint32_t* d_blk = reinterpret_cast<uint32_t*>(base + kDescBytes);
    // the descriptor travels as a kernel argument (no staging allocation)
    Uploader up(s);
    up.add(d_desc, &d, 1);
    up.flush();
    launch_rs_build(d_desc, 1, num_l1, rs_scan_chunks(num_l1), d_blk, s);
  });
}

}  // extern "C"
