// ecl_wavelet/detail/quad_core.hpp -- the quad and hex line layouts and
// their descents (no counterpart in the reference).
//
// Quad lines pack the bits of two consecutive levels as 2-bit digits, hex
// lines those of four levels as 4-bit digits, in 128-byte lines that carry
// the counts of the digits before them; quad select lines sample the
// positions of every r-th occurrence of each digit.  They are opt-in
// structures of power-of-two alphabets built by the library
// (ecl_wavelet_tree_build_quad_lines / _quad_select_lines / _hex_lines).
// The descents run on a pair of adjacent lanes of a wave (threadIdx.x 2k
// and 2k + 1, as the cooperative groups of the device API; both active, both
// with the same query): each lane reads one half of a line.  They take any parameter type P with members `view` (TreeView),
// `quad` (QuadView) and `hex` (HexView): the library's kernel parameters, or
// LayoutRef for the device API of the C++ drop-in.
#pragma once

#include "ecl_wavelet/detail/device_core.hpp"

namespace ecl {
namespace detail {

// Quad lines (opt-in; power-of-two alphabets with an even number of levels,
// at most 16): quad level q holds the 2-bit digits of the codes (bits 2q and
// 2q+1 from the top, the first one high) in the order of binary level 2q,
// in 128-byte lines: the counts of digits 0, 1 and 2 before the line as
// three 40-bit fields in the first 16 bytes, then 448 digits, digit i at
// bits 2i..2i+1 of the 112 bytes.  With them access and rank read one line
// per two levels.  Descending from position p in node N of quad level q by
// digit d leads to position base[off[q] + 4N + d] + (d's before p) of quad
// level q + 1 (base = the child's start minus the d's before node N).
constexpr uint32_t kQuadDigits = 448;
// matches of digit d among the 32 digits of w, one bit per digit (bit 2i)
__device__ __forceinline__ uint64_t quad_match(uint64_t w, uint32_t d) {
  const uint64_t y = w ^ (0x5555555555555555ull * d);
  return ~(y | (y >> 1)) & 0x5555555555555555ull;
}
constexpr uint32_t kMaxQuadLevels = 8;
struct QuadView {
  const uint4* lines[kMaxQuadLevels];
  const uint64_t* base;          // every level's bases
  uint64_t off[kMaxQuadLevels];  // level q's bases start at base + off[q]
  uint64_t nbase;                // entries of base
  uint32_t levels;               // 0: no quad lines
  // Node tables of every level (for select), in one array of ntab u64:
  // tab[s_off[q] + N] = start of node N of level 2q (N in [0, 4^q], the
  // last one n) and tab[cd_off[q] + 4N + d] = the d's before node N.
  const uint64_t* tab;
  uint64_t ntab;
  uint64_t s_off[kMaxQuadLevels], cd_off[kMaxQuadLevels];
  // Optional quad select lines: for level q and digit d with rate
  // r = sel_rate[q][d] > 0, line j (from sel_first[q][d] in sel_lines[q])
  // is the position s_j of the (j r)-th d of the level (0 for j = 0), 8
  // unused bytes, then the 448 digits from the 32-digit word holding s_j.
  const uint4* sel_lines[kMaxQuadLevels];
  uint64_t sel_first[kMaxQuadLevels][4];
  uint32_t sel_rate[kMaxQuadLevels][4];
  uint64_t sel_magic[kMaxQuadLevels][4];
  uint32_t has_sel;
};

// Hex lines (opt-in; power-of-two alphabets whose level count is a multiple
// of 4, at most 16): hex level h holds the 4-bit digits of the codes (bits
// 4h .. 4h+3 from the top, the first one high) in the order of binary level
// 4h, in 128-byte lines of 128 digits.  A line is four 32-byte parts: the
// counts of digits 0..7 before the line (u32 each), digits 0..63, the
// counts of digits 8..15, digits 64..127 (digit i at bits 4(i mod 16) of
// its 64-bit word), so that each lane of a pair reads one half.  A count is
// stored modulo 2^32: with sb = the full count at the start of the line's
// superblock (2^24 lines, 2^31 digits), the count is sb + (u32)(stored - sb).
// Descending from position p in node N of hex level h by digit x leads to
// position base[16N + x] + (x's before p) of hex level h + 1 (base = the
// child's start minus the x's before node N).  With them access and rank
// read one line per four levels.
constexpr uint32_t kHexDigits = 128;
constexpr uint32_t kMaxHexLevels = 4;
constexpr uint32_t kHexSbShift = 24;  // lines per superblock: 2^24
// matches of digit x among the 16 digits of w, one bit per digit (bit 4i)
__device__ __forceinline__ uint64_t hex_match(uint64_t w, uint32_t x) {
  const uint64_t y = w ^ (0x1111111111111111ull * x);
  return ~(y | (y >> 1) | (y >> 2) | (y >> 3)) & 0x1111111111111111ull;
}
struct HexView {
  const uint4* lines[kMaxHexLevels];  // n / 128 + 1 lines per level
  // one array of ntab u64: tab[base_off[h] + 16N + x] = base, and
  // tab[sb_off[h] + 16 s + x] = the x's before superblock s
  const uint64_t* tab;
  uint64_t ntab;
  uint64_t base_off[kMaxHexLevels], sb_off[kMaxHexLevels];
  uint32_t levels;  // 0: no hex lines
};

// The parameters of the descents as references to a tree's views (the
// C++ drop-in's device API reads the quad and hex views from device memory).
struct LayoutRef {
  const TreeView& view;
  const QuadView& quad;
  const HexView& hex;
};

// A node table read from LDS (staged by the workgroup) or from global
// memory, fixed at compile time so that the reads are ds_read or global
// loads rather than generic (flat) loads, which wait on both counters.
template <bool Lds>
struct QuadTable {
  const uint64_t* t;
  __device__ __forceinline__ uint64_t operator[](uint64_t i) const {
    if constexpr (Lds) return t[i];
    else return gld(t, i);
  }
};

// The partner's value within a pair of lanes (2k, 2k + 1): a DPP quad_perm
// [1, 0, 3, 2] move, one VALU instruction instead of an LDS permute.  Both
// lanes of a pair are active wherever these are used.
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ uint64_t pair_swap64(uint64_t v) {
  return static_cast<uint64_t>(pair_swap(static_cast<uint32_t>(v))) |
         (static_cast<uint64_t>(pair_swap(static_cast<uint32_t>(v >> 32))) << 32);
}

// one quad level for the pair: the digit at offset o of line `line` (when
// want_digit; otherwise d is given) and the d's before position line*448+o
struct QuadStep {
  uint32_t d;
  uint64_t rank;
};
__device__ __forceinline__ QuadStep quad_step(const uint4* lines, uint64_t p, bool want_digit,
                                              uint32_t d_given) {
  const uint32_t h = threadIdx.x & 1u;
  // p / 448 = (p / 64) / 7, the division by 7 as a multiply-high
  const uint64_t line = __umul64hi(p >> 6, 0x2492492492492493ull);
  const uint32_t o = static_cast<uint32_t>(p - line * kQuadDigits);
  const uint4* lp = lines + line * 8 + 4 * h;
  const uint4 x0 = gld(lp, 0), x1 = gld(lp, 1), x2 = gld(lp, 2),
              x3 = gld(lp, 3);
  // this lane's digit words: lane 0 words 0..5 (x1..x3), lane 1 words 6..13
  uint64_t w[8];
  w[0] = h ? ((uint64_t)x0.x | ((uint64_t)x0.y << 32)) : ((uint64_t)x1.x | ((uint64_t)x1.y << 32));
  w[1] = h ? ((uint64_t)x0.z | ((uint64_t)x0.w << 32)) : ((uint64_t)x1.z | ((uint64_t)x1.w << 32));
  w[2] = h ? ((uint64_t)x1.x | ((uint64_t)x1.y << 32)) : ((uint64_t)x2.x | ((uint64_t)x2.y << 32));
  w[3] = h ? ((uint64_t)x1.z | ((uint64_t)x1.w << 32)) : ((uint64_t)x2.z | ((uint64_t)x2.w << 32));
  w[4] = h ? ((uint64_t)x2.x | ((uint64_t)x2.y << 32)) : ((uint64_t)x3.x | ((uint64_t)x3.y << 32));
  w[5] = h ? ((uint64_t)x2.z | ((uint64_t)x2.w << 32)) : ((uint64_t)x3.z | ((uint64_t)x3.w << 32));
  w[6] = h ? ((uint64_t)x3.x | ((uint64_t)x3.y << 32)) : 0ull;
  w[7] = h ? ((uint64_t)x3.z | ((uint64_t)x3.w << 32)) : 0ull;
  const uint32_t kb = h ? 6u : 0u;   // first word of this lane
  const uint32_t ko = o >> 5;        // word holding offset o
  const uint32_t sh = 2 * (o & 31);  // its bit
  uint32_t d = d_given;
  if (want_digit) {
    uint64_t sel = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
      if (kb + j == ko) sel = w[j];
    const uint32_t mine = static_cast<uint32_t>((sel >> sh) & 3u);
    const uint32_t owner = ko < 6 ? 0u : 1u;
    // the owner's digit to both lanes of the pair
    const uint32_t other = pair_swap(mine);
    d = owner == h ? mine : other;
  }
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    const uint32_t k = kb + j;
    if (k <= ko && (h || j < 6)) {
      uint64_t m = quad_match(w[j], d);
      if (k == ko) m &= sh == 0 ? 0ull : (~0ull >> (64 - sh));
      cnt += static_cast<uint32_t>(__popcll(m));
    }
  }
  // lane 0 holds the header: counts of digits 0, 1, 2 before the line; each
  // lane's part (lane 0: header + count, lane 1: count) plus the partner's
  uint64_t part = cnt;
  if (h == 0) {
    const uint64_t h0 = (uint64_t)x0.x | ((uint64_t)x0.y << 32);
    const uint64_t h1 = (uint64_t)x0.z | ((uint64_t)x0.w << 32);
    const uint64_t m40 = (1ull << 40) - 1;
    const uint64_t c0 = h0 & m40, c1 = (h0 >> 40) | ((h1 & 0xFFFFull) << 24),
                   c2 = (h1 >> 16) & m40;
    part += d == 0 ? c0 : d == 1 ? c1 : d == 2 ? c2 : line * kQuadDigits - c0 - c1 - c2;
  }
  return QuadStep{d, part + pair_swap64(part)};
}

// access(i) over the quad lines: the symbol (minimal), for the pair
template <typename P, typename Table>
__device__ __forceinline__ uint64_t quad_access(const P& p, Table base,
                                                uint64_t i) {
  uint64_t pos = i, N = 0;
  for (uint32_t l = 0; l < p.quad.levels; ++l) {
    const QuadStep st = quad_step(p.quad.lines[l], pos, true, 0);
    pos = base[p.quad.off[l] + 4 * N + st.d] + st.rank;
    N = 4 * N + st.d;
  }
  return N;
}
// rank_c(i) over the quad lines, c < sigma and i <= n, for the pair
template <typename P, typename Table>
__device__ __forceinline__ uint64_t quad_rank(const P& p, Table base,
                                              uint64_t c, uint64_t i) {
  const uint32_t Q = p.quad.levels;
  const uint64_t cc = gld(p.view.counts, c);  // issued ahead of the descent
  uint64_t pos = i, N = 0;
  for (uint32_t l = 0; l < Q; ++l) {
    const uint32_t d = static_cast<uint32_t>(c >> (2 * (Q - 1 - l))) & 3u;
    const QuadStep st = quad_step(p.quad.lines[l], pos, false, d);
    pos = base[p.quad.off[l] + 4 * N + d] + st.rank;
    N = 4 * N + d;
  }
  return pos - cc;
}

// one hex level for the pair: the digit at position p (when want_digit;
// otherwise d is given) and the d's before p
template <typename Table>
__device__ __forceinline__ QuadStep hex_step(const uint4* lines, Table tab, uint64_t sb_off,
                                             uint64_t p, bool want_digit, uint32_t d_given) {
  const uint32_t h = threadIdx.x & 1u;
  const uint64_t line = p >> 7;
  const uint32_t o = static_cast<uint32_t>(p & 127);
  const uint4* lp = lines + line * 8 + 4 * h;
  const uint4 c0 = gld(lp, 0), c1 = gld(lp, 1), x0 = gld(lp, 2),
              x1 = gld(lp, 3);
  const uint64_t w[4] = {(uint64_t)x0.x | ((uint64_t)x0.y << 32),
                         (uint64_t)x0.z | ((uint64_t)x0.w << 32),
                         (uint64_t)x1.x | ((uint64_t)x1.y << 32),
                         (uint64_t)x1.z | ((uint64_t)x1.w << 32)};
  uint32_t d = d_given;
  if (want_digit) {
    const uint32_t k = (o >> 4) & 3u, sh = 4 * (o & 15);
    const uint64_t sel = k == 0 ? w[0] : k == 1 ? w[1] : k == 2 ? w[2] : w[3];
    const uint32_t mine = static_cast<uint32_t>(sel >> sh) & 15u;
    const uint32_t other = pair_swap(mine);
    d = (o >> 6) == h ? mine : other;  // the owner's digit to both lanes
  }
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t g0 = 64 * h + 16 * k;  // the word's first digit
    if (g0 < o) {
      uint64_t m = hex_match(w[k], d);
      if (g0 + 16 > o) m &= ~0ull >> (64 - 4 * (o - g0));
      cnt += static_cast<uint32_t>(__popcll(m));
    }
  }
  // the stored count of d (mod 2^32): lane d >> 3, entry d & 7
  const uint32_t e = d & 7u;
  const uint4 cc = e < 4 ? c0 : c1;
  const uint32_t cv = (e & 3u) == 0 ? cc.x : (e & 3u) == 1 ? cc.y : (e & 3u) == 2 ? cc.z : cc.w;
  const uint32_t part = cnt + ((d >> 3) == h ? cv : 0u);
  const uint32_t tot = part + pair_swap(part);
  const uint64_t sbv = tab[sb_off + (line >> kHexSbShift) * 16 + d];
  return QuadStep{d, sbv + static_cast<uint32_t>(tot - static_cast<uint32_t>(sbv))};
}

// access(i) over the hex lines: the symbol (minimal), for the pair
template <typename P, typename Table>
__device__ __forceinline__ uint64_t hex_access(const P& p, Table tab, uint64_t i) {
  uint64_t pos = i, N = 0;
  for (uint32_t l = 0; l < p.hex.levels; ++l) {
    const QuadStep st = hex_step(p.hex.lines[l], tab, p.hex.sb_off[l], pos, true, 0);
    pos = tab[p.hex.base_off[l] + 16 * N + st.d] + st.rank;
    N = 16 * N + st.d;
  }
  return N;
}
// rank_c(i) over the hex lines, c < sigma and i <= n, for the pair
template <typename P, typename Table>
__device__ __forceinline__ uint64_t hex_rank(const P& p, Table tab, uint64_t c,
                                             uint64_t i) {
  const uint32_t H = p.hex.levels;
  const uint64_t cc = gld(p.view.counts, c);  // issued ahead of the descent
  uint64_t pos = i, N = 0;
  for (uint32_t l = 0; l < H; ++l) {
    const uint32_t d = static_cast<uint32_t>(c >> (4 * (H - 1 - l))) & 15u;
    const QuadStep st = hex_step(p.hex.lines[l], tab, p.hex.sb_off[l], pos, false, d);
    pos = tab[p.hex.base_off[l] + 16 * N + d] + st.rank;
    N = 16 * N + d;
  }
  return pos - cc;
}

// One 128-byte line of 448 digits for the pair (a quad line, or a quad
// select line, which has the same layout): this lane's digit words (lane 0:
// words 0..5, lane 1: words 6..13, zeros after) and the line's first 16
// bytes (lane 0's header).
struct PairLine {
  uint64_t w[8];
  uint64_t h0, h1;
};
__device__ __forceinline__ PairLine pair_line(const uint4* line) {
  const uint32_t h = threadIdx.x & 1u;
  const uint4* lp = line + 4 * h;
  const uint4 x0 = gld(lp, 0), x1 = gld(lp, 1), x2 = gld(lp, 2),
              x3 = gld(lp, 3);
  PairLine L;
  L.w[0] = h ? ((uint64_t)x0.x | ((uint64_t)x0.y << 32)) : ((uint64_t)x1.x | ((uint64_t)x1.y << 32));
  L.w[1] = h ? ((uint64_t)x0.z | ((uint64_t)x0.w << 32)) : ((uint64_t)x1.z | ((uint64_t)x1.w << 32));
  L.w[2] = h ? ((uint64_t)x1.x | ((uint64_t)x1.y << 32)) : ((uint64_t)x2.x | ((uint64_t)x2.y << 32));
  L.w[3] = h ? ((uint64_t)x1.z | ((uint64_t)x1.w << 32)) : ((uint64_t)x2.z | ((uint64_t)x2.w << 32));
  L.w[4] = h ? ((uint64_t)x2.x | ((uint64_t)x2.y << 32)) : ((uint64_t)x3.x | ((uint64_t)x3.y << 32));
  L.w[5] = h ? ((uint64_t)x2.z | ((uint64_t)x2.w << 32)) : ((uint64_t)x3.z | ((uint64_t)x3.w << 32));
  L.w[6] = h ? ((uint64_t)x3.x | ((uint64_t)x3.y << 32)) : 0ull;
  L.w[7] = h ? ((uint64_t)x3.z | ((uint64_t)x3.w << 32)) : 0ull;
  L.h0 = (uint64_t)x0.x | ((uint64_t)x0.y << 32);
  L.h1 = (uint64_t)x0.z | ((uint64_t)x0.w << 32);
  return L;
}

// The digit index in the line (0..447) of the need-th (1-based) match among
// the pair's masks m, where c0 is lane 0's count of matches; ~0u to both
// lanes when the line holds fewer than `need`.  The word is picked first
// and selected in once: a select under each word test would run, on wave64,
// for every word that any lane stops at.
__device__ __forceinline__ uint32_t pair_select_in_line(const uint64_t (&m)[8], uint32_t c0,
                                                        uint64_t need) {
  const uint32_t h = threadIdx.x & 1u;
  const uint32_t nw = h ? 8u : 6u;
  const uint32_t owner = need <= c0 ? 0u : 1u;
  uint64_t nd = h ? need - c0 : need;
  uint64_t word = 0;
  uint32_t kw = 0;
  bool found = false;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t c = static_cast<uint32_t>(__popcll(m[k]));
    const bool hit = !found && k < nw && nd <= c;
    word = hit ? m[k] : word;
    kw = hit ? k : kw;
    nd = found || hit || k >= nw ? nd : nd - c;
    found = found || hit;
  }
  const uint32_t idx =
      owner == h && found
          ? 32u * (kw + (h ? 6u : 0u)) + (select_in_word(word, static_cast<uint32_t>(nd)) >> 1)
          : ~0u;
  // the owner's answer to both lanes
  const uint32_t theirs = pair_swap(idx);
  return owner == h ? idx : theirs;
}

// position of the g-th (1-based) d of quad level ql, for the pair, from its
// quad select line; *past is set (and n returned) when the d lies beyond
// the line's digits, which then end before digit *end
template <typename P>
__device__ __forceinline__ uint64_t quad_select_line(const P& p, uint32_t ql,
                                                     uint32_t d, uint64_t g, bool* past,
                                                     uint64_t* end) {
  const uint32_t h = threadIdx.x & 1u;
  // the level's four digits' tables (the level is the same for the whole
  // wave), then this lane's digit picked by selects: no per-lane indexing
  // into the kernel arguments
  const uint32_t* rr = p.quad.sel_rate[ql];
  const uint64_t* mg = p.quad.sel_magic[ql];
  const uint64_t* fs = p.quad.sel_first[ql];
  const uint32_t r = d < 2 ? (d == 0 ? rr[0] : rr[1]) : (d == 2 ? rr[2] : rr[3]);
  const uint64_t magic = d < 2 ? (d == 0 ? mg[0] : mg[1]) : (d == 2 ? mg[2] : mg[3]);
  const uint64_t first = d < 2 ? (d == 0 ? fs[0] : fs[1]) : (d == 2 ? fs[2] : fs[3]);
  const uint64_t j = r == 1 ? g : __umul64hi(g, magic);
  const PairLine L = pair_line(p.quad.sel_lines[ql] + (first + j) * 8);
  const uint64_t hp = pair_swap64(L.h0);
  const uint64_t s = h ? hp : L.h0;  // lane 0's header: the line's first position
  const uint64_t rem = j ? g - j * r + 1 : g;  // counted from s itself
  const uint32_t nw = h ? 8u : 6u;
  uint64_t m[8];
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    m[k] = k < nw ? quad_match(L.w[k], d) : 0ull;
    if (h == 0 && k == 0) m[k] &= ~0ull << (2 * (s & 31));  // digits before s
    cnt += static_cast<uint32_t>(__popcll(m[k]));
  }
  const uint32_t cp = pair_swap(cnt);
  const uint32_t c0 = h ? cp : cnt;
  const uint32_t idx = pair_select_in_line(m, c0, rem);
  *past = idx == ~0u;
  *end = (s & ~31ull) + kQuadDigits;
  return (s & ~31ull) + idx;
}

// position of the g-th (1-based) d of quad level ql at or after digit `from`
// (a line start; fewer than g d's lie before it), for the pair: a forward
// walk over the quad lines, each header giving the d's before its line.  The
// walk ends within the level's lines because g is at most the level's count
// of d; n is returned if it does not.
template <typename P>
__device__ __forceinline__ uint64_t quad_select_walk(const P& p, uint32_t ql,
                                                         uint32_t d, uint64_t g, uint64_t from) {
  const uint32_t h = threadIdx.x & 1u;
  const uint64_t n = p.view.n;
  const uint64_t nl = __umul64hi(n >> 6, 0x2492492492492493ull) + 1;  // n / 448 + 1 lines
  const uint32_t nw = h ? 8u : 6u;
  for (uint64_t l = __umul64hi(from >> 6, 0x2492492492492493ull); l < nl; ++l) {
    const PairLine L = pair_line(p.quad.lines[ql] + l * 8);
    uint64_t m[8];
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      m[k] = k < nw ? quad_match(L.w[k], d) : 0ull;
      cnt += static_cast<uint32_t>(__popcll(m[k]));
    }
    // lane 0: the header's count of d before the line
    const uint64_t m40 = (1ull << 40) - 1;
    const uint64_t c0h = L.h0 & m40, c1h = (L.h0 >> 40) | ((L.h1 & 0xFFFFull) << 24),
                   c2h = (L.h1 >> 16) & m40;
    const uint64_t bd =
        d == 0 ? c0h : d == 1 ? c1h : d == 2 ? c2h : l * kQuadDigits - c0h - c1h - c2h;
    const uint64_t bp = pair_swap64(bd);
    const uint64_t before = h ? bp : bd;
    const uint32_t cp = pair_swap(cnt);
    const uint32_t c0 = h ? cp : cnt;
    const uint64_t total = before + cnt + cp;  // both lanes: the d's up to the line's end
    if (g <= total) return l * kQuadDigits + pair_select_in_line(m, c0, g - before);
  }
  return n;
}

// select_c(k) over the quad select lines (n when absent), for the pair
template <typename P, typename Table>
__device__ __forceinline__ uint64_t quad_select(const P& p, Table tab,
                                                uint64_t c, uint64_t k) {
  const uint32_t Q = p.quad.levels;
  const uint64_t n = p.view.n, sigma = p.view.sigma;
  if (c >= sigma || k == 0) return n;
  // c's count from the node tables of the last quad level (c's node there
  // is c >> 2, its digit c & 3): the d's before the next node minus those
  // before this one
  {
    const uint64_t cd = p.quad.cd_off[Q - 1];
    const uint64_t N = c >> 2, d = c & 3;
    if (k > tab[cd + 4 * (N + 1) + d] - tab[cd + 4 * N + d]) return n;
  }
  // the node-table entries of a level depend only on c: the next level's
  // are read while this level's line is in flight
  uint64_t kk = k;
  int32_t ql = static_cast<int32_t>(Q) - 1;
  uint64_t N = c >> 2;
  uint32_t d = static_cast<uint32_t>(c) & 3u;
  uint64_t cdv = tab[p.quad.cd_off[ql] + 4 * N + d], sv = tab[p.quad.s_off[ql] + N];
  while (true) {
    const uint64_t g = cdv + kk;
    const int32_t qn = ql - 1;
    uint64_t Nn = 0, cdn = 0, sn = 0;
    uint32_t dn = 0;
    if (qn >= 0) {
      Nn = c >> (2 * (Q - qn));
      dn = static_cast<uint32_t>(c >> (2 * (Q - 1 - qn))) & 3u;
      cdn = tab[p.quad.cd_off[qn] + 4 * Nn + dn];
      sn = tab[p.quad.s_off[qn] + Nn];
    }
    bool past = false;
    uint64_t end = 0;
    uint64_t pos = quad_select_line(p, static_cast<uint32_t>(ql), d, g, &past, &end);
    // past the line's digits (a sparse stretch of d): the quad lines from
    // the one holding digit `end` on
    if (past) pos = quad_select_walk(p, static_cast<uint32_t>(ql), d, g, end);
    kk = pos - sv + 1;
    if (qn < 0) break;
    ql = qn;
    N = Nn;
    d = dn;
    cdv = cdn;
    sv = sn;
  }
  return kk - 1;
}

}  // namespace detail
}  // namespace ecl
