// device_core.hpp -- wave64 (gfx950) device primitives shared by the HIP
// library's kernels and the header-only device API of the drop-in classes.
//
// Layout conventions kept from the reference so that every exposed value is
// identical (mfranzreb/ECL_wavelet):
//   * bit j of an array is (word[j/32] >> (j%32)) & 1, u32 words
//     (bitarray/bit_array.cuh:39-46); each array is padded to a multiple of
//     32 words = 128 B (src/bitarray/bit_array.cu:57-72);
//   * L2 = u16 ones count per 512-bit block relative to its 65536-bit L1
//     block, L1 = u64 ones before the L1 block (bitarray/rank_select.cuh:
//     44-64);
//   * select samples every 4096 ones / zeros (rank_select.cuh:59).
// Select samples of one array share one buffer of floor(m/4096) slots: the
// ones samples grow from the front, the zeros samples from the back.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ecl {
namespace detail {

constexpr uint64_t kL2Bits = 512;
constexpr uint64_t kL2PerL1 = 128;
constexpr uint64_t kL1Bits = kL2Bits * kL2PerL1;
constexpr uint64_t kSampleRate = 4096;
constexpr uint64_t kSkip = ~0ull;  // node-table entry of a leaf child

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot64(bool p) {
  return static_cast<uint64_t>(__ballot(p));
}

// number of set bits of `mask` in lanes below the calling lane
__device__ __forceinline__ uint32_t mbcnt64(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(
      static_cast<uint32_t>(mask >> 32),
      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Inclusive wave64 scan with DPP: shifts by 1, 2, 4, 8 lanes inside each
// row of 16 (lanes that would read across the row start get 0), then the
// last lane of row 0 / rows 0-1 is broadcast into the later rows
// (row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3).
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
  v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xF, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xF, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xF, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xF, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xA, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xC, 0xF, false));
  return v;
}

// the same scan over two independent 32-bit halves of a u64 (no carry
// between the halves)
__device__ __forceinline__ uint64_t wave_incl_scan_2x32(uint64_t v) {
  const uint32_t lo = wave_incl_scan_u32(static_cast<uint32_t>(v));
  const uint32_t hi = wave_incl_scan_u32(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// value of lane 63
__device__ __forceinline__ uint32_t wave_last_u32(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// position (0-based) of the r-th (1-based) set bit of x; r <= popcount(x)
__device__ __forceinline__ uint32_t select_in_word(uint64_t x, uint32_t r) {
  uint32_t pos = 0;
  uint32_t c = __popc(static_cast<uint32_t>(x));
  if (r > c) {
    r -= c;
    x >>= 32;
    pos = 32;
  }
  uint32_t y = static_cast<uint32_t>(x);
  c = __popc(y & 0xFFFFu);
  if (r > c) { r -= c; y >>= 16; pos += 16; }
  c = __popc(y & 0xFFu);
  if (r > c) { r -= c; y >>= 8; pos += 8; }
  c = __popc(y & 0xFu);
  if (r > c) { r -= c; y >>= 4; pos += 4; }
  c = __popc(y & 0x3u);
  if (r > c) { r -= c; y >>= 2; pos += 2; }
  if (r > (y & 1u)) pos += 1;
  return pos;
}

__host__ __device__ __forceinline__ bool is_pow2(uint64_t n) {
  return (n & (n - 1)) == 0;
}
// largest power of two strictly below n, 0 for n == 0
// (utils/utils.cuh:138-158)
__host__ __device__ __forceinline__ uint64_t prev_pow2(uint64_t n) {
  if (n == 0) return 0;
  if (is_pow2(n)) return n >> 1;
#if defined(__HIP_DEVICE_COMPILE__)
  return 1ull << (63 - __clzll(static_cast<long long>(n)));
#else
  return 1ull << (63 - __builtin_clzll(n));
#endif
}
__host__ __device__ __forceinline__ uint32_t ceil_log2(uint64_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t hb = 63 - __clzll(static_cast<long long>(n));
#else
  uint32_t hb = 63 - __builtin_clzll(n);
#endif
  return is_pow2(n) ? hb : hb + 1;
}
__host__ __device__ __forceinline__ uint64_t shr64(uint64_t v, uint32_t s) {
  return s >= 64 ? 0 : (v >> s);
}

// counter-based generator shared with the oracle: splitmix64 finaliser over
// seed + (i+1)*golden, Lemire's multiply-high reduction to [0, m)
__host__ __device__ __forceinline__ uint64_t hash64(uint64_t seed,
                                                    uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t bounded64(uint64_t h,
                                                       uint64_t m) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(h, m);
#else
  return static_cast<uint64_t>((static_cast<unsigned __int128>(h) * m) >> 64);
#endif
}

// Element i of an array in device memory, read with a global load.  The
// level views sit in LDS in the query kernels, so the pointers read from
// them are generic to the compiler, which would issue flat loads: a flat
// load counts on both the vector-memory and the LDS/scalar counters, and
// the compiler waits for every outstanding LDS and flat access before using
// one.  Every array a LevelView points to is device memory the library
// allocated, so these pointers are always global.
template <typename T>
__device__ __forceinline__ T gld(const T* p, uint64_t i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}
// uint4 is a class type without a copy from another address space: load
// the 16 bytes as a plain vector
__device__ __forceinline__ uint4 gld(const uint4* p, uint64_t i) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 v = ((const __attribute__((address_space(1))) u32x4*)p)[i];
  return make_uint4(v.x, v.y, v.z, v.w);
}

// ---------------------------------------------------------------------------
// One bit array with its rank/select directory.  Trivially copyable; the
// library keeps one per array in device memory.
// ---------------------------------------------------------------------------
struct LevelView {
  const uint32_t* words;    // word 0 of this array (128 B aligned)
  const uint64_t* l1;       // L1 entries of this array
  const uint16_t* l2;       // L2 entries of this array
  const uint64_t* samples;  // ones samples from the front, zeros from back
  uint64_t size;            // bits
  uint64_t ones;            // total ones
  uint64_t num_l1;
  uint64_t sample_cap;      // floor(size / 4096)
  // Optional query lines (nullptr when not built): 128-byte lines, each the
  // u64 count of ones before the line followed by the line's 960 bits, so a
  // rank reads one line instead of the bit block, its L2 and its L1 entry.
  const uint4* lines;
  // With the lines: one 64-byte select hint per sample bracket, ones then
  // zeros (hint j of a Value covers the bracket from its j-th sample, or bit
  // 0, to the next sample, or the last bit):
  //   u64 position of sample j (0 for j = 0), u64 Values before the 512-bit
  //   block holding it (block b0), then 24 u16: the bracket's last block
  //   minus b0 (0xFFFF above kHintBlocks), and for u = 1..23 the Values
  //   before block b0 + u minus the u64 count (0xFFFF past the bracket).
  // A select then reads its hint and the one block that holds the answer.
  const uint4* sel_hints;
  // bit v: a select of Value v reads the hints (set where at least half of
  // that Value's brackets fit one hint; otherwise every select would read a
  // hint only to fall back to the directory search)
  uint32_t hint_values;
  // Optional select lines (nullptr when not built): for Value v with rate
  // r = sel_rate[v] > 0, line j (j = 0 .. total_v / r; the zeros' lines from
  // line sel_zero_line on) is the u64 position s_j of the (j r)-th Value (0
  // for j = 0) followed by the 960 bits from bit s_j & ~63.  r follows the
  // Value's density, so the next r Values lie in those bits; a select of
  // Value v then reads one line.
  const uint4* sel_lines;
  uint64_t sel_zero_line;
  uint32_t sel_rate[2];
  // floor((2^64 - 1) / r) + 1 for r = sel_rate[v] > 1: i / r is then the
  // high word of i * sel_magic[v] for every i < 2^64 / r (the counts here are
  // below 2^44), which spares a 64-bit division per level
  uint64_t sel_magic[2];
};
constexpr uint32_t kLineBits = 960;  // bits per query line (30 words)
constexpr uint32_t kHintBlocks = 23;  // blocks after b0 one hint covers

__device__ __forceinline__ bool level_bit(const LevelView& v, uint64_t i) {
  return (gld(v.words, i >> 5) >> (i & 31)) & 1u;
}

// Ones in [0, index) for index in (0, size): L1 + L2 + popcount of the
// 512-bit block prefix.  The 64 B block lies inside the array's 128 B
// padding, so the full-block load never leaves the allocation.
__device__ __forceinline__ uint64_t rank1_inner(const LevelView& v,
                                                uint64_t index, bool* bit) {
  const uint64_t blk = index >> 9;
  const uint4* p = reinterpret_cast<const uint4*>(v.words + (blk << 4));
  uint4 q0 = gld(p, 0), q1 = gld(p, 1), q2 = gld(p, 2), q3 = gld(p, 3);
  uint64_t r = gld(v.l1, index >> 16) + gld(v.l2, blk);
  uint64_t w[8] = {
      (uint64_t)q0.x | ((uint64_t)q0.y << 32), (uint64_t)q0.z | ((uint64_t)q0.w << 32),
      (uint64_t)q1.x | ((uint64_t)q1.y << 32), (uint64_t)q1.z | ((uint64_t)q1.w << 32),
      (uint64_t)q2.x | ((uint64_t)q2.y << 32), (uint64_t)q2.z | ((uint64_t)q2.w << 32),
      (uint64_t)q3.x | ((uint64_t)q3.y << 32), (uint64_t)q3.z | ((uint64_t)q3.w << 32)};
  const uint32_t wi = static_cast<uint32_t>((index >> 6) & 7);
  const uint32_t bi = static_cast<uint32_t>(index & 63);
  uint32_t acc = 0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    uint64_t x = w[k];
    if (k == wi) {
      *bit = (x >> bi) & 1u;
      x &= (bi == 0) ? 0ull : (~0ull >> (64 - bi));
    }
    acc += (k <= wi) ? static_cast<uint32_t>(__popcll(x)) : 0u;
  }
  return r + acc;
}

// Ones in [0, index) for index in (0, size) from the query line holding
// the index: the line's header plus the ones of its bits before the index.
__device__ __forceinline__ uint64_t rank1_lines(const LevelView& v, uint64_t index, bool* bit) {
  const uint64_t line = index / kLineBits;
  const uint32_t off = static_cast<uint32_t>(index - line * kLineBits);
  const uint4* p = v.lines + line * 8;
  uint4 q[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] = gld(p, k);
  const uint64_t head = (uint64_t)q[0].x | ((uint64_t)q[0].y << 32);
  // bits as 15 u64 words: (q0.z, q0.w), (q1.x, q1.y), (q1.z, q1.w), ...
  uint64_t w[15];
  w[0] = (uint64_t)q[0].z | ((uint64_t)q[0].w << 32);
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    w[2 * k - 1] = (uint64_t)q[k].x | ((uint64_t)q[k].y << 32);
    if (2 * k < 15) w[2 * k] = (uint64_t)q[k].z | ((uint64_t)q[k].w << 32);
  }
  const uint32_t wi = off >> 6, bi = off & 63;
  uint32_t acc = 0;
  bool b = false;
#pragma unroll
  for (uint32_t k = 0; k < 15; ++k) {
    uint64_t x = w[k];
    if (k == wi) {
      b = (x >> bi) & 1u;
      x &= (bi == 0) ? 0ull : (~0ull >> (64 - bi));
    }
    acc += (k <= wi) ? static_cast<uint32_t>(__popcll(x)) : 0u;
  }
  *bit = b;
  return head + acc;
}

// rank_Value(index) over [0, index), index in [0, size]; rank_select.cuh:
// 125-208 semantics (index == size returns the total, bit = false).
template <int Value>
__device__ __forceinline__ uint64_t level_rank(const LevelView& v,
                                               uint64_t index,
                                               bool* bit = nullptr) {
  bool b = false;
  uint64_t ones;
  if (index == 0) {
    ones = 0;
    if (bit) b = level_bit(v, 0);
  } else if (index >= v.size) {
    ones = v.ones;
    index = v.size;
  } else {
    ones = v.lines != nullptr ? rank1_lines(v, index, &b) : rank1_inner(v, index, &b);
  }
  if (bit) *bit = b;
  return Value ? ones : index - ones;
}

// Ones in [0, index) for index in (0, size), computed by the two lanes of
// an adjacent pair (lanes 2k and 2k+1 of a wave, both active, same
// arguments), each fetching half of the bits: with query lines, lane 0 the
// line's header and first 448 bits, lane 1 its other 512 bits; otherwise
// each lane 256 bits of the 64 B block, and both the L1/L2 entries (the
// same addresses, one fetch).  A random line read by one lane as eight 16 B
// loads moves far fewer lines per second than the same line split over two
// lanes.  Both lanes return the count and the bit at index.
__device__ __forceinline__ uint64_t rank1_pair(const LevelView& v, uint64_t index, bool* bit) {
  const uint32_t t = threadIdx.x & 1u;
  const bool lines = v.lines != nullptr;
  uint32_t d[16];    // this lane's fetched dwords
  uint32_t first;    // d[first] is bit 0 of this lane's half
  uint32_t o, split; // index's offset in the line / block; first bit of lane 1's half
  uint64_t base;     // ones before the line / block
  if (lines) {
    const uint64_t line = index / kLineBits;
    o = static_cast<uint32_t>(index - line * kLineBits);
    split = 448;
    const uint4* p = v.lines + line * 8 + 4 * t;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 q = gld(p, k);
      d[4 * k] = q.x; d[4 * k + 1] = q.y; d[4 * k + 2] = q.z; d[4 * k + 3] = q.w;
    }
    // lane 0's first two dwords are the line's header
    const unsigned long long h = (uint64_t)d[0] | ((uint64_t)d[1] << 32);
    const unsigned long long ho = __shfl_xor(h, 1, 2);
    base = t ? ho : h;
    first = t ? 0u : 2u;
  } else {
    const uint64_t blk = index >> 9;
    o = static_cast<uint32_t>(index & 511);
    split = 256;
    const uint4* p = reinterpret_cast<const uint4*>(v.words + (blk << 4)) + 2 * t;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint4 q = gld(p, k);
      d[4 * k] = q.x; d[4 * k + 1] = q.y; d[4 * k + 2] = q.z; d[4 * k + 3] = q.w;
    }
#pragma unroll
    for (int k = 8; k < 16; ++k) d[k] = 0;
    base = gld(v.l1, index >> 16) + gld(v.l2, blk);
    first = 0;
  }
  // this lane counts its bits before index and reads the bit at index when
  // index lies in its half (lane 0: o < split, lane 1: o >= split)
  const bool mine = (t == 0) == (o < split);
  const uint32_t off = t == 0 ? (o < split ? o : split) : (mine ? o - split : 0u);
  const uint32_t wi = off >> 5, bi = off & 31;
  uint32_t acc = 0, b = 0;
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) {
    if (k >= first) {
      const uint32_t j = k - first;
      const uint32_t w = d[k];
      if (j < wi) acc += __popc(w);
      else if (j == wi) { acc += __popc(w & ((1u << bi) - 1u)); b = (w >> bi) & 1u; }
    }
  }
  b = mine ? b : 0u;
  acc += __shfl_xor(acc, 1, 2);
  b |= __shfl_xor(b, 1, 2);
  *bit = b != 0;
  return base + acc;
}

// level_rank with the pair fetch (rank1_pair): both lanes of the pair call it
// with the same arguments and get the same result.
template <int Value>
__device__ __forceinline__ uint64_t level_rank_pair(const LevelView& v, uint64_t index,
                                                    bool* bit = nullptr) {
  bool b = false;
  uint64_t ones;
  if (index == 0) {
    ones = 0;
    if (bit) b = level_bit(v, 0);
  } else if (index >= v.size) {
    ones = v.ones;
    index = v.size;
  } else {
    ones = rank1_pair(v, index, &b);
  }
  if (bit) *bit = b;
  return Value ? ones : index - ones;
}

// Position of the i-th (1-based) Value; size when i exceeds the total
// (rank_select.cuh:224-523 semantics).  Search: the sample pair brackets
// the answer; the L1 block is found by binary search inside the bracket
// (none when the bracket lies in one block); the L2 block by binary search,
// or with WideL2 by loading all entries of a bracket of up to 16 blocks at
// once and counting those below the target (counts before a block are
// monotone); then the 64 B block is loaded at once and the bit selected
// in-word.  WideL2 shortens the dependent chain of one select (binary
// rank/select, one level per query) but issues more instructions, which
// costs the tree descents and mixed batches more than it saves.
// Position of the r-th (1-based) Value inside 512-bit block blk, which
// holds it: the 64 B block is loaded at once and the bit selected in-word.
template <int Value>
__device__ __forceinline__ uint64_t select_in_block(const LevelView& v, uint64_t blk2,
                                                   uint64_t r) {
  // the 64 B block lies inside the array's 128 B padding
  const uint4* p = reinterpret_cast<const uint4*>(v.words + (blk2 << 4));
  const uint4 q0 = gld(p, 0), q1 = gld(p, 1), q2 = gld(p, 2), q3 = gld(p, 3);
  const uint64_t w[8] = {
      (uint64_t)q0.x | ((uint64_t)q0.y << 32), (uint64_t)q0.z | ((uint64_t)q0.w << 32),
      (uint64_t)q1.x | ((uint64_t)q1.y << 32), (uint64_t)q1.z | ((uint64_t)q1.w << 32),
      (uint64_t)q2.x | ((uint64_t)q2.y << 32), (uint64_t)q2.z | ((uint64_t)q2.w << 32),
      (uint64_t)q3.x | ((uint64_t)q3.y << 32), (uint64_t)q3.z | ((uint64_t)q3.w << 32)};
  // (selecting once after the loop, as select_from_line does, made the hint
  // and directory searches slower: 130 instead of 116 VGPRs)
  uint64_t pos = blk2 << 9, res = v.size;
  bool found = false;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    const uint64_t x = Value ? w[k] : ~w[k];
    const uint32_t c = static_cast<uint32_t>(__popcll(x));
    if (!found) {
      if (r <= c) {
        res = pos + select_in_word(x, static_cast<uint32_t>(r));
        found = true;
      } else {
        r -= c;
        pos += 64;
      }
    }
  }
  return res;  // size only for an inconsistent directory
}

// The i-th Value (0 < i <= total) from its select line; size when the
// answer lies past the line's bits (the caller then searches otherwise).
template <int Value>
__device__ __forceinline__ uint64_t select_from_line(const LevelView& v, uint64_t i,
                                                     uint32_t rate) {
  const uint64_t jj = rate == 1 ? i : __umul64hi(i, v.sel_magic[Value]);  // i / rate
  const uint4* p = v.sel_lines + ((Value ? 0 : v.sel_zero_line) + jj) * 8;
  uint4 q[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] = gld(p, k);
  const uint64_t s = (uint64_t)q[0].x | ((uint64_t)q[0].y << 32);
  uint64_t w[15];
  w[0] = (uint64_t)q[0].z | ((uint64_t)q[0].w << 32);
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    w[2 * k - 1] = (uint64_t)q[k].x | ((uint64_t)q[k].y << 32);
    if (2 * k < 15) w[2 * k] = (uint64_t)q[k].z | ((uint64_t)q[k].w << 32);
  }
  // counted from s (the (jj rate)-th Value itself, or bit 0 for jj = 0)
  uint64_t r = jj ? i - jj * rate + 1 : i;
  const uint32_t off = static_cast<uint32_t>(s & 63);
  // the word first, then one in-word select: a select under each of the
  // unrolled word tests runs for every word that any lane of the wave stops at
  uint32_t kw = 15;
  uint64_t xw = 0;
#pragma unroll
  for (uint32_t k = 0; k < 15; ++k) {
    uint64_t x = Value ? w[k] : ~w[k];
    if (k == 0) x &= ~0ull << off;
    const uint32_t c = static_cast<uint32_t>(__popcll(x));
    if (kw == 15) {
      if (r <= c) {
        kw = k;
        xw = x;
      } else {
        r -= c;
      }
    }
  }
  if (kw == 15) return v.size;
  return (s & ~63ull) + 64ull * kw + select_in_word(xw, static_cast<uint32_t>(r));
}

// SelLines: consult the select lines when the array has them (the query
// kernels of trees without them compile the path out, which keeps their
// register use and occupancy).
template <int Value, bool WideL2 = false, bool SelLines = true, bool Hints = true>
__device__ __forceinline__ uint64_t level_select(const LevelView& v,
                                                 uint64_t i) {
  const uint64_t total = Value ? v.ones : v.size - v.ones;
  if (i == 0 || i > total) return v.size;
  if (SelLines && v.sel_lines != nullptr && v.sel_rate[Value] != 0) {
    const uint64_t r = select_from_line<Value>(v, i, v.sel_rate[Value]);
    if (r < v.size) return r;
  }
  const uint64_t j = i / kSampleRate;
  const uint64_t nsamp = total / kSampleRate;
  if (Hints && v.sel_hints != nullptr && ((v.hint_values >> Value) & 1u)) {
    // the bracket's hint: sample position, Values before its first block,
    // and the counts before the next kHintBlocks blocks
    const uint4* h = v.sel_hints + (Value ? j : v.ones / kSampleRate + 1 + j) * 4;
    const uint4 h0 = gld(h, 0), h1 = gld(h, 1), h2 = gld(h, 2), h3 = gld(h, 3);
    const uint64_t lo_pos = (uint64_t)h0.x | ((uint64_t)h0.y << 32);
    if (j > 0 && i == j * kSampleRate) return lo_pos;
    const uint32_t span = h1.x & 0xFFFFu;
    if (span <= kHintBlocks) {
      const uint64_t c0 = (uint64_t)h0.z | ((uint64_t)h0.w << 32);
      const uint32_t hw[12] = {h1.x, h1.y, h1.z, h1.w, h2.x, h2.y, h2.z, h2.w,
                               h3.x, h3.y, h3.z, h3.w};
      // blocks b0 + u (u <= span) with fewer than i Values before them; the
      // counts grow with u, so the last such block holds the answer
      uint32_t below = 0, eb = 0;
#pragma unroll
      for (uint32_t u = 1; u <= kHintBlocks; ++u) {
        const uint32_t e = (hw[u >> 1] >> ((u & 1) * 16)) & 0xFFFFu;
        if (u <= span && c0 + e < i) {
          below = u;
          eb = e;
        }
      }
      return select_in_block<Value>(v, (lo_pos >> 9) + below, i - c0 - eb);
    }
  }
  uint64_t lo_pos = 0, hi_pos = v.size - 1;
  if (j > 0) {
    lo_pos = gld(v.samples, Value ? j - 1 : v.sample_cap - j);
    if (i == j * kSampleRate) return lo_pos;
  }
  if (j < nsamp) hi_pos = gld(v.samples, Value ? j : v.sample_cap - 1 - j);
  // L1: largest block b in [lo_pos>>16, hi_pos>>16] with count_before(b) < i
  uint64_t b_lo = lo_pos >> 16, b_hi = hi_pos >> 16;
  while (b_lo < b_hi) {
    uint64_t mid = (b_lo + b_hi + 1) >> 1;
    const uint64_t e = gld(v.l1, mid);
    uint64_t c = Value ? e : (mid << 16) - e;
    if (c < i) b_lo = mid; else b_hi = mid - 1;
  }
  const uint64_t blk1 = b_lo;
  const uint64_t e1 = gld(v.l1, blk1);
  const uint64_t c1 = Value ? e1 : (blk1 << 16) - e1;
  uint64_t r = i - c1;  // rank inside the L1 block (1-based)
  // L2: largest k in [max(lo_pos>>9, blk1*128), min(hi_pos>>9, end)] with
  // count_before(k) < r
  uint64_t k_lo = lo_pos >> 9, k_hi = hi_pos >> 9;
  const uint64_t first = blk1 << 7, last = first + kL2PerL1 - 1;
  if (k_lo < first) k_lo = first;
  if (k_hi > last) k_hi = last;
  uint64_t blk2;
  if (WideL2 && k_hi - k_lo < 16) {
    const uint32_t span = static_cast<uint32_t>(k_hi - k_lo);
    uint32_t below = 0;
#pragma unroll
    for (uint32_t u = 1; u < 16; ++u) {
      if (u <= span) {
        const uint64_t k = k_lo + u;
        const uint64_t e = gld(v.l2, k);
        const uint64_t c = Value ? e : ((k - first) << 9) - e;
        below += c < r ? 1u : 0u;
      }
    }
    blk2 = k_lo + below;
  } else {
    while (k_lo < k_hi) {
      uint64_t mid = (k_lo + k_hi + 1) >> 1;
      const uint64_t e = gld(v.l2, mid);
      uint64_t c = Value ? e : ((mid - first) << 9) - e;
      if (c < r) k_lo = mid; else k_hi = mid - 1;
    }
    blk2 = k_lo;
  }
  const uint64_t e2 = gld(v.l2, blk2);
  r -= Value ? e2 : ((blk2 - first) << 9) - e2;
  return select_in_block<Value>(v, blk2, r);
}

// ---------------------------------------------------------------------------
// Cooperative forms for groups of G lanes: G is a power of two up to the
// wave width, a group is G consecutive lanes of one wave (threadIdx.x % G
// numbers a lane inside it), as the reference's NumThreads groups
// (rank_select.cuh:125-208).  Every lane of a group calls with the same
// arguments and gets the same result; each lane reads only its share of the
// 64-byte block or 128-byte line, and the group combines the counts.
// ---------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ uint32_t group_lane() {
  return static_cast<uint32_t>(threadIdx.x) & (G - 1);
}
template <int G, typename V>
__device__ __forceinline__ V group_sum(V v) {
#pragma unroll
  for (int off = 1; off < G; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}
template <int G>
__device__ __forceinline__ uint32_t group_excl_scan(uint32_t v) {
  const uint32_t t = group_lane<G>();
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, G);
    if (t >= static_cast<uint32_t>(off)) incl += y;
  }
  return incl - v;
}
// the value of the group's lane where `owner` holds (the first one), and
// whether there is one
template <int G>
__device__ __forceinline__ uint64_t group_from(bool owner, uint64_t val, bool* any) {
  const uint64_t m = static_cast<uint64_t>(__ballot(owner));
  const uint32_t base = (static_cast<uint32_t>(threadIdx.x) & 63u) & ~static_cast<uint32_t>(G - 1);
  const uint64_t gm = G == 64 ? m : (m >> base) & ((1ull << (G & 63)) - 1ull);
  *any = gm != 0;
  const int src = static_cast<int>(base) + (gm ? __ffsll(static_cast<long long>(gm)) - 1 : 0);
  return __shfl(val, src, 64);
}

// ones in [0, index) for index in (0, size), lanes splitting the words
template <int G>
__device__ __forceinline__ uint64_t rank1_coop(const LevelView& v, uint64_t index, bool* bit) {
  const uint32_t t = group_lane<G>();
  const uint64_t* w;
  uint32_t nw, wi, bi;
  uint64_t base;
  if (v.lines != nullptr) {
    const uint64_t line = index / kLineBits;
    const uint32_t off = static_cast<uint32_t>(index - line * kLineBits);
    const uint64_t* lp = reinterpret_cast<const uint64_t*>(v.lines + line * 8);
    base = gld(lp, 0);
    w = lp + 1;
    nw = 15;
    wi = off >> 6;
    bi = off & 63;
  } else {
    const uint64_t blk = index >> 9;
    w = reinterpret_cast<const uint64_t*>(v.words + (blk << 4));
    nw = 8;
    wi = static_cast<uint32_t>((index >> 6) & 7);
    bi = static_cast<uint32_t>(index & 63);
    base = gld(v.l1, index >> 16) + gld(v.l2, blk);
  }
  uint32_t acc = 0, b = 0;
  for (uint32_t k = t; k < nw && k <= wi; k += G) {
    uint64_t x = gld(w, k);
    if (k == wi) {
      b = static_cast<uint32_t>((x >> bi) & 1u);
      x &= bi == 0 ? 0ull : (~0ull >> (64 - bi));
    }
    acc += static_cast<uint32_t>(__popcll(x));
  }
  acc = group_sum<G>(acc);
  b = group_sum<G>(b);
  *bit = b != 0;
  return base + acc;
}

// level_rank<Value> by a group of G lanes
template <int Value, int G>
__device__ __forceinline__ uint64_t level_rank_coop(const LevelView& v, uint64_t index,
                                                    bool* bit = nullptr) {
  if constexpr (G == 1) return level_rank<Value>(v, index, bit);
  bool b = false;
  uint64_t ones;
  if (index == 0) {
    ones = 0;
    if (bit) b = level_bit(v, 0);
  } else if (index >= v.size) {
    ones = v.ones;
    index = v.size;
  } else {
    ones = rank1_coop<G>(v, index, &b);
  }
  if (bit) *bit = b;
  return Value ? ones : index - ones;
}

// position of the r-th (1-based) Value in the nw words from w (the first
// word's bits below off dropped), the words split over the group in
// contiguous shares; `none` when the words hold fewer than r
template <int Value, int G>
__device__ __forceinline__ uint64_t group_select_words(const uint64_t* w, uint32_t nw,
                                                       uint32_t off, uint64_t r,
                                                       uint64_t pos0, uint64_t none) {
  const uint32_t t = group_lane<G>();
  const uint32_t k0 = t * nw / G, k1 = (t + 1) * nw / G;
  uint32_t cnt = 0;
  for (uint32_t k = k0; k < k1; ++k) {
    uint64_t x = Value ? gld(w, k) : ~gld(w, k);
    if (k == 0) x &= ~0ull << off;
    cnt += static_cast<uint32_t>(__popcll(x));
  }
  const uint32_t ex = group_excl_scan<G>(cnt);
  const bool own = ex < r && r <= static_cast<uint64_t>(ex) + cnt;
  uint64_t pos = none;
  if (own) {
    uint64_t rr = r - ex;
    for (uint32_t k = k0; k < k1; ++k) {
      uint64_t x = Value ? gld(w, k) : ~gld(w, k);
      if (k == 0) x &= ~0ull << off;
      const uint32_t c = static_cast<uint32_t>(__popcll(x));
      if (rr <= c) {
        pos = pos0 + 64ull * k + select_in_word(x, static_cast<uint32_t>(rr));
        break;
      }
      rr -= c;
    }
  }
  bool any = false;
  pos = group_from<G>(own, pos, &any);
  return any ? pos : none;
}

// level_select<Value> by a group of G lanes: the select line when the array
// has one (its 960 bits split over the lanes), otherwise the directory: the
// sample bracket and the L1 search in every lane, the bracket's L2 entries
// and the final 64-byte block split over the lanes
template <int Value, int G>
__device__ __forceinline__ uint64_t level_select_coop(const LevelView& v, uint64_t i) {
  if constexpr (G == 1) return level_select<Value, true>(v, i);
  const uint64_t total = Value ? v.ones : v.size - v.ones;
  if (i == 0 || i > total) return v.size;
  const uint32_t rate = v.sel_rate[Value];
  if (v.sel_lines != nullptr && rate != 0) {
    const uint64_t jj = rate == 1 ? i : __umul64hi(i, v.sel_magic[Value]);
    const uint64_t* lp =
        reinterpret_cast<const uint64_t*>(v.sel_lines + ((Value ? 0 : v.sel_zero_line) + jj) * 8);
    const uint64_t s = gld(lp, 0);
    const uint64_t r = jj ? i - jj * rate + 1 : i;
    const uint64_t res = group_select_words<Value, G>(lp + 1, 15, static_cast<uint32_t>(s & 63),
                                                      r, s & ~63ull, v.size);
    if (res < v.size) return res;
  }
  const uint64_t j = i / kSampleRate;
  const uint64_t nsamp = total / kSampleRate;
  uint64_t lo_pos = 0, hi_pos = v.size - 1;
  if (j > 0) {
    lo_pos = gld(v.samples, Value ? j - 1 : v.sample_cap - j);
    if (i == j * kSampleRate) return lo_pos;
  }
  if (j < nsamp) hi_pos = gld(v.samples, Value ? j : v.sample_cap - 1 - j);
  uint64_t b_lo = lo_pos >> 16, b_hi = hi_pos >> 16;
  while (b_lo < b_hi) {
    const uint64_t mid = (b_lo + b_hi + 1) >> 1;
    const uint64_t e = gld(v.l1, mid);
    const uint64_t c = Value ? e : (mid << 16) - e;
    if (c < i) b_lo = mid; else b_hi = mid - 1;
  }
  const uint64_t blk1 = b_lo;
  const uint64_t e1 = gld(v.l1, blk1);
  const uint64_t r = i - (Value ? e1 : (blk1 << 16) - e1);
  uint64_t k_lo = lo_pos >> 9, k_hi = hi_pos >> 9;
  const uint64_t first = blk1 << 7, last = first + kL2PerL1 - 1;
  if (k_lo < first) k_lo = first;
  if (k_hi > last) k_hi = last;
  // blocks k_lo + u (0 < u <= span) with fewer than r Values before them
  // (the counts grow with u): counted by the lanes together
  uint32_t below = 0;
  for (uint64_t u = 1 + group_lane<G>(); u <= k_hi - k_lo; u += G) {
    const uint64_t k = k_lo + u;
    const uint64_t e = gld(v.l2, k);
    below += (Value ? e : ((k - first) << 9) - e) < r ? 1u : 0u;
  }
  const uint64_t blk2 = k_lo + group_sum<G>(below);
  const uint64_t e2 = gld(v.l2, blk2);
  const uint64_t rr = r - (Value ? e2 : ((blk2 - first) << 9) - e2);
  return group_select_words<Value, G>(reinterpret_cast<const uint64_t*>(v.words + (blk2 << 4)),
                                      8, 0, rr, blk2 << 9, v.size);
}

// ---------------------------------------------------------------------------
// Wavelet tree view (minimal-alphabet symbols), trivially copyable.
// ---------------------------------------------------------------------------
struct TreeView {
  const LevelView* levels;      // [num_levels]
  const uint64_t* counts;       // C[c]: occurrences of symbols < c
  const uint64_t* ranks;        // node ranks ordered by (level, symbol)
  const uint64_t* nodes_at_level;  // [num_levels - 1] node starts per level
  const uint64_t* codes;        // code of symbol codes_start + k
  const uint8_t* code_lens;     // length of that code
  uint64_t n;
  uint64_t sigma;
  uint64_t codes_start;
  uint64_t nodes_until_last_level;
  uint64_t num_ranks;
  uint32_t num_levels;
  uint32_t pow2;
};

__device__ __forceinline__ uint64_t tree_total(const TreeView& t,
                                               const uint64_t* counts,
                                               uint64_t c) {
  return c == t.sigma - 1 ? t.n - counts[c] : counts[c + 1] - counts[c];
}

__device__ __forceinline__ uint64_t tree_nodes_at(const TreeView& t,
                                                  const uint64_t* nal,
                                                  uint32_t l) {
  return t.pow2 ? (1ull << l) - 1 : nal[l];
}

// x / d for a divisor d that is a power of two (a shift; d == 0 keeps the
// division, which valid trees never reach)
__device__ __forceinline__ uint64_t div_pow2(uint64_t x, uint64_t d) {
  return d != 0 ? x >> __builtin_ctzll(d) : x / d;
}

// getNodePosAtLevel (tree/wavelet_tree.cuh:2241-2268); the node lengths are
// powers of two (a power of two shifted right), so the divisions are shifts
__device__ __forceinline__ uint64_t tree_node_pos(const TreeView& t,
                                                  uint64_t symbol,
                                                  uint32_t level) {
  if (level == 1) return 0;
  uint64_t prev = 0, node_len = 1;
  if (t.pow2) {
    node_len = 1ull << (t.num_levels - level);
  } else {
    uint64_t remaining = t.sigma;
    for (int32_t i = static_cast<int32_t>(level) - 1; i >= 0; --i) {
      uint64_t sub = prev_pow2(remaining);
      node_len = sub >> i;
      if (symbol <= sub) break;
      remaining -= sub;
      prev += div_pow2(sub, node_len);
      symbol -= sub;
    }
  }
  return div_pow2(symbol, node_len) + prev - 1;
}

// access: top-down descent (tree/wavelet_tree.cuh:1815-1877).  Pair: the
// two lanes of an adjacent pair descend for the same index together, each
// fetching half of every level's bits (rank1_pair).
template <bool Pair = false, int G = 1>
__device__ __forceinline__ uint64_t tree_access(const TreeView& t,
                                                const LevelView* lv,
                                                const uint64_t* counts,
                                                const uint64_t* ranks,
                                                const uint64_t* nal,
                                                uint64_t index) {
  uint64_t lo = 0, hi = t.sigma - 1, rs = 0;
  for (uint32_t l = 0; l < t.num_levels; ++l) {
    const LevelView v = lv[l];
    const uint64_t cc = counts[lo];
    if (hi - lo < 2) {
      if (hi - lo > 0 && level_bit(v, cc + index)) lo++;
      break;
    }
    const uint64_t start = lo == 0 ? 0 : ranks[rs + tree_node_pos(t, lo, l)];
    bool bit;
    const uint64_t pos = Pair ? level_rank_pair<0>(v, cc + index, &bit)
                              : level_rank_coop<0, G>(v, cc + index, &bit);
    const uint64_t diff = prev_pow2(hi - lo + 1);
    if (!bit) {
      index = pos - start;
      hi = lo + diff - 1;
    } else {
      index -= pos - start;
      lo += diff;
    }
    rs += tree_nodes_at(t, nal, l);
  }
  return lo;
}

// rank: rankKernel early-out + top-down descent (wavelet_tree.cuh:481-486,
// 1896-1956).  Pair as for tree_access.
template <bool Pair = false, int G = 1>
__device__ __forceinline__ uint64_t tree_rank(const TreeView& t,
                                              const LevelView* lv,
                                              const uint64_t* counts,
                                              const uint64_t* ranks,
                                              const uint64_t* nal,
                                              uint64_t symbol,
                                              uint64_t index) {
  if (symbol >= t.sigma || tree_total(t, counts, symbol) == 0) return 0;
  uint64_t lo = 0, hi = t.sigma - 1, rs = 0, result = index;
  for (uint32_t l = 0; l < t.num_levels; ++l) {
    if (hi == lo) break;
    const uint64_t cc = counts[lo];
    const uint64_t start = lo == 0 ? 0 : ranks[rs + tree_node_pos(t, lo, l)];
    const uint64_t pos = Pair ? level_rank_pair<0>(lv[l], cc + result)
                              : level_rank_coop<0, G>(lv[l], cc + result);
    const uint64_t split = lo + prev_pow2(hi - lo + 1);
    if (symbol < split) {
      result = pos - start;
      hi = split - 1;
    } else {
      result -= pos - start;
      lo = split;
    }
    if (l + 1 < t.num_levels) rs += tree_nodes_at(t, nal, l);
  }
  return result;
}

// getPrevCharStart (wavelet_tree.cuh:1970-1992)
__device__ __forceinline__ uint64_t tree_prev_char_start(
    const TreeView& t, uint64_t cs, bool rightmost, uint32_t level,
    uint32_t code_len) {
  if (t.pow2) {
    const uint64_t node_len = 1ull << (t.num_levels - level);
    return cs & ~(node_len - 1);
  }
  if (rightmost) {
    uint64_t s = 0;
    for (uint32_t l = 0; l < level; ++l) s += prev_pow2(t.sigma - s);
    return s;
  }
  return cs - (1ull << (code_len - 1 - level));
}

// select: selectKernel early-out + bottom-up ascent (wavelet_tree.cuh:
// 532-535, 2007-2079).  SelLines as for level_select.
template <bool SelLines = true, bool Hints = true, int G = 1>
__device__ __forceinline__ uint64_t tree_select(const TreeView& t,
                                                const LevelView* lv,
                                                const uint64_t* counts,
                                                const uint64_t* ranks,
                                                const uint64_t* nal,
                                                uint64_t symbol, uint64_t k) {
  if (symbol >= t.sigma || k == 0 || k > tree_total(t, counts, symbol))
    return t.n;
  const uint32_t L = t.num_levels;
  uint64_t code = symbol;
  uint32_t clen = L;
  if (!t.pow2 && symbol >= t.codes_start) {
    code = t.codes[symbol - t.codes_start];
    clen = t.code_lens[symbol - t.codes_start];
  }
  uint64_t cs = symbol, result = k;
  uint64_t rs = t.nodes_until_last_level;
  int32_t l = static_cast<int32_t>(clen) - 1;
  for (int32_t lv2 = static_cast<int32_t>(L) - 2; lv2 >= l; --lv2)
    rs -= tree_nodes_at(t, nal, static_cast<uint32_t>(lv2));
  for (; l >= 0; --l) {
    const LevelView v = lv[l];
    uint64_t start;
    if ((code >> (L - 1 - l)) & 1) {
      const bool rightmost =
          static_cast<uint32_t>(__popcll(shr64(code, L - l))) ==
          static_cast<uint32_t>(l);
      cs = tree_prev_char_start(t, cs, rightmost, static_cast<uint32_t>(l),
                                clen);
      start = cs == 0 ? 0
                      : counts[cs] - ranks[rs + tree_node_pos(
                                                    t, cs, static_cast<uint32_t>(l))];
      if constexpr (G > 1) result = level_select_coop<1, G>(v, start + result) + 1;
      else result = level_select<1, false, SelLines, Hints>(v, start + result) + 1;
    } else {
      start = cs == 0
                  ? 0
                  : ranks[rs + tree_node_pos(t, cs, static_cast<uint32_t>(l))];
      if constexpr (G > 1) result = level_select_coop<0, G>(v, start + result) + 1;
      else result = level_select<0, false, SelLines, Hints>(v, start + result) + 1;
    }
    if (static_cast<uint32_t>(l) == clen - 1 && result > v.size) {
      result = t.n + 1;
      break;
    }
    result -= counts[cs];
    if (l > 1) rs -= tree_nodes_at(t, nal, static_cast<uint32_t>(l - 1));
  }
  return result - 1;
}

}  // namespace detail
}  // namespace ecl
