// internal.hpp -- declarations shared by the translation units of
// libecl_wavelet_hip.so (not installed; the public surface is
// include/ecl_wavelet_c.h and include/ecl_wavelet/).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ecl_wavelet/detail/device_core.hpp"
#include "ecl_wavelet_c.h"

namespace ecl {
namespace lib {

using detail::LevelView;
using detail::TreeView;

// An exception carrying an ecl_status; converted at the C boundary.
struct Error : std::runtime_error {
  ecl_status status;
  Error(ecl_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

void hip_check(hipError_t e, const char* what);
#define ECL_HIP(call) ::ecl::lib::hip_check((call), #call)

// Device allocation that records itself for freeing (RAII).
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b);
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept;
  void reset();
  template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

// ---- kernel profiling (events around selected launches) -------------------
enum KernelId : int {
  K_HIST = 0, K_LEVEL, K_RS_BUILD, K_WT_ACCESS, K_WT_RANK, K_WT_SELECT,
  K_WT_MIXED, K_RS_RANK, K_RS_SELECT, K_GEN, K_FILL, K_COUNT
};
const char* kernel_name(int id);
struct ProfScope {
  int id; hipStream_t stream; hipEvent_t e0 = nullptr, e1 = nullptr;
  ProfScope(int id, hipStream_t s);
  ~ProfScope();
};

// ---- launchers (kernels_build.hip) ------------------------------------------
void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

struct HistParams {
  const void* text; uint64_t n; int esize;
  uint64_t sigma;          // alphabet size
  int map_kind;            // 0 identity, 1 LUT (u8/u16), 2 binary search
  const uint32_t* lut;     // value -> index (map_kind 1)
  const void* alphabet;    // sorted alphabet (map_kind 2), element type
  uint64_t codes_start; int pow2;
  const uint64_t* codes;   // code of symbol codes_start + k
  void* codes_out;         // remapped text (nullable)
  uint64_t* hist;          // [sigma] u64 counts (zeroed by caller)
};
void launch_histogram(const HistParams& p, hipStream_t s);

// One level pass: bits of the level, its rank/select directory (L1, L2,
// select samples, total ones) and, unless it is the last level, the stable
// partition of the symbols into the next level's order.
struct LevelParams {
  const void* src; void* dst; int esize;   // dst == nullptr: last level
  uint64_t n, n_next;
  uint32_t bit;            // code bit of this level (L-1-l)
  uint32_t pshift;         // code >> pshift = node prefix (L-l)
  const uint64_t* table;   // [2 * prefixes] child bases (kSkip = leaf)
  uint64_t num_prefixes;   // 2^l
  uint64_t* words64;       // level bit words (u64 view)
  uint64_t nwords64;       // u64 words this level writes
  uint64_t* l1; uint16_t* l2; uint64_t num_l1, num_l2;
  uint64_t* l2_global;     // scratch: ones before each 512-bit block
  uint64_t* samples; uint64_t cap;
  uint64_t* total_ones;
  uint64_t* status; uint64_t* counter;  // lookback state (zeroed)
};
// tiles of a level pass; the look-back state needs 2 * tiles + 2 u64
uint64_t level_tiles(uint64_t n, int esize);
void launch_level(const LevelParams& p, hipStream_t s);

struct RsBuildParams {
  const uint64_t* words64; uint64_t m;
  uint64_t* l1; uint16_t* l2; uint64_t num_l1, num_l2;
  uint64_t* samples; uint64_t cap;
  uint64_t* total_ones;
  uint64_t* status; uint64_t* counter;  // lookback state (zeroed)
};
void launch_rs_build(const RsBuildParams& p, hipStream_t s);

void launch_gen_text(void* out, int esize, uint64_t n, uint64_t seed,
                     uint64_t sigma, const uint64_t* thresholds,
                     hipStream_t s);
void launch_gen_bits(uint32_t* words, uint64_t nbits, uint64_t nwords_total,
                     uint64_t seed, uint64_t threshold, hipStream_t s);
void launch_presence(const void* text, int esize, uint64_t n, uint32_t* flags,
                     hipStream_t s);

// ---- launchers (kernels_query.hip) -------------------------------------------
struct TreeQueryParams {
  TreeView view;
  int esize;
  int symbols_minimal;       // 1: query symbols already minimal indices
  int map_kind; const uint32_t* lut; const void* alphabet;  // symbol map
  const void* alphabet_out;  // access: map result back (nullable = identity)
  uint64_t max_lds_counts;   // unused by callers; tuning knob
};
void launch_tree_access(const TreeQueryParams& p, const uint64_t* idx,
                        uint64_t m, void* out, hipStream_t s);
void launch_tree_rank(const TreeQueryParams& p, const void* q, uint64_t m,
                      uint64_t* out, hipStream_t s);
void launch_tree_select(const TreeQueryParams& p, const void* q, uint64_t m,
                        uint64_t* out, hipStream_t s);
void launch_tree_mixed(const TreeQueryParams& p, const ecl_query* q,
                       uint64_t m, uint64_t* out, hipStream_t s);
void launch_tree_gen_queries(const TreeView& v, const uint64_t* hcounts_dev,
                             int kind, uint64_t seed, uint64_t m, void* out,
                             int esize, hipStream_t s);
void launch_rs_rank(const LevelView* lv, int value, const uint64_t* idx,
                    uint64_t m, uint64_t* out, int generated, uint64_t seed,
                    hipStream_t s);
void launch_rs_select(const LevelView* lv, int value, const uint64_t* k,
                      uint64_t m, uint64_t* out, int generated, uint64_t seed,
                      hipStream_t s);

}  // namespace lib
}  // namespace ecl
