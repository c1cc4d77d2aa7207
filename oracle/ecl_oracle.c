/*
 * ecl_oracle.c -- CPU restatement of the reference's wavelet-tree and
 * rank/select semantics.  TEST INFRASTRUCTURE ONLY (see ecl_oracle.h): the
 * product library never links this file.
 *
 * Paths below are relative to the reference repository mfranzreb/ECL_wavelet.
 */
#include "ecl_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------
 * Counter-based generator.  h(seed, i) = splitmix64(seed + (i+1)*golden);
 * bounded(h, m) = floor(h*m / 2^64) (Lemire).  The GPU library implements
 * the same function independently; tests compare the two.
 * ---------------------------------------------------------------------- */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
uint64_t orc_hash(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
uint64_t orc_bounded(uint64_t h, uint64_t m) {
  return (uint64_t)(((unsigned __int128)h * m) >> 64);
}

static inline void put_elem(void* out, int esize, uint64_t i, uint64_t v) {
  switch (esize) {
    case 1: ((uint8_t*)out)[i] = (uint8_t)v; break;
    case 2: ((uint16_t*)out)[i] = (uint16_t)v; break;
    case 4: ((uint32_t*)out)[i] = (uint32_t)v; break;
    default: ((uint64_t*)out)[i] = v; break;
  }
}
static inline uint64_t get_elem(const void* in, int esize, uint64_t i) {
  switch (esize) {
    case 1: return ((const uint8_t*)in)[i];
    case 2: return ((const uint16_t*)in)[i];
    case 4: return ((const uint32_t*)in)[i];
    default: return ((const uint64_t*)in)[i];
  }
}

void orc_gen_uniform(uint64_t seed, uint64_t n, uint64_t sigma, int esize,
                     void* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    put_elem(out, esize, (uint64_t)i, orc_bounded(orc_hash(seed, i), sigma));
}

void orc_gen_cdf(uint64_t seed, uint64_t n, const uint64_t* thr,
                 uint64_t sigma, int esize, void* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t h = orc_hash(seed, i);
    uint64_t lo = 0, hi = sigma - 1; /* first s with h < thr[s] */
    while (lo < hi) {
      uint64_t mid = (lo + hi) / 2;
      if (h < thr[mid]) hi = mid; else lo = mid + 1;
    }
    put_elem(out, esize, (uint64_t)i, lo);
  }
}

void orc_gen_bits(uint64_t seed, uint64_t nbits, uint64_t threshold,
                  uint32_t* words) {
  uint64_t nw = (nbits + 31) / 32;
#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < (int64_t)nw; ++w) {
    uint32_t v = 0;
    for (int b = 0; b < 32; ++b) {
      uint64_t i = (uint64_t)w * 32 + b;
      if (i < nbits && orc_hash(seed, i) < threshold) v |= 1u << b;
    }
    words[w] = v;
  }
}

/* ------------------------------------------------------------------------
 * utils.cuh:123-240
 * ---------------------------------------------------------------------- */
/* isPowTwo: (n & (n-1)) == 0 (utils.cuh:123-129) */
int orc_is_pow2(uint64_t n) { return (n & (n - 1)) == 0; }
/* getPrevPowTwo: largest power of two strictly below n; 0 for n == 0
 * (utils.cuh:138-158) */
uint64_t orc_prev_pow2(uint64_t n) {
  if (n == 0) return 0;
  if (orc_is_pow2(n)) return n >> 1;
  return 1ULL << (63 - __builtin_clzll(n));
}
/* ceilLog2 (utils.cuh:206-240) */
uint32_t orc_ceil_log2(uint64_t n) {
  uint32_t hb = 63 - __builtin_clzll(n);
  return orc_is_pow2(n) ? hb : hb + 1;
}

/* ------------------------------------------------------------------------
 * createMinimalCodes, wavelet_tree.cuh:2100-2159.
 * ---------------------------------------------------------------------- */
uint64_t orc_minimal_codes(uint64_t sigma, uint64_t* code, uint8_t* len) {
  uint32_t L = orc_ceil_log2(sigma);
  for (uint64_t i = 0; i < sigma; ++i) {
    code[i] = i;
    len[i] = (uint8_t)L;
  }
  if (orc_is_pow2(sigma)) return 0;
  uint64_t total = 0;
  uint32_t start_bit = 0;
  uint64_t start_i = 0;
  uint32_t code_len = L;
  uint64_t num = sigma;
  do {
    for (uint32_t i = code_len - 1; i > 0; --i) {
      uint64_t p = 1ULL << i;
      if (num <= p) break;
      num -= p;
      start_i += p;
      start_bit++;
    }
    if (code_len == L) total = num; /* first round */
    if (num == 1) {
      code_len = 1;
      len[sigma - 1] = (uint8_t)start_bit;
      code[sigma - 1] = ((1ULL << start_bit) - 1) << (L - start_bit);
    } else {
      code_len = orc_ceil_log2(num);
      for (uint64_t i = sigma - num; i < sigma; ++i) {
        uint64_t c = (i - start_i) << (L - start_bit - code_len);
        c += (~((1ULL << (L - start_bit)) - 1)) & code[i];
        code[i] = c;
        len[i] = (uint8_t)(start_bit + code_len);
      }
    }
  } while (code_len > 1);
  return total;
}

/* getNodeInfos, wavelet_tree.cuh:2501-2532: a node of level l >= 1 starts at
 * symbol i when len(i) > l, bit (L-l-1) of code(i) is 0 and that bit of
 * code(i-1) is 1. */
uint64_t orc_node_infos(uint64_t sigma, const uint64_t* code,
                        const uint8_t* len, uint64_t* start, uint8_t* level) {
  uint32_t L = orc_ceil_log2(sigma);
  uint64_t cnt = 0;
  for (uint32_t l = 1; l < L; ++l) {
    uint32_t b = L - l - 1;
    for (uint64_t i = 1; i < sigma; ++i) {
      if (len[i] > l && ((code[i] >> b) & 1) == 0 &&
          ((code[i - 1] >> b) & 1) == 1) {
        if (start) start[cnt] = i;
        if (level) level[cnt] = (uint8_t)l;
        cnt++;
      }
    }
  }
  return cnt;
}

/* getNodePosAtLevel, wavelet_tree.cuh:2241-2268 */
uint64_t orc_node_pos_at_level(uint64_t sigma, uint32_t L, uint64_t symbol,
                               uint32_t level) {
  if (level == 1) return 0;
  uint64_t prev = 0, node_len = 1;
  if (orc_is_pow2(sigma)) {
    node_len = 1ULL << (L - level);
  } else {
    uint64_t remaining = sigma;
    for (int32_t i = (int32_t)level - 1; i >= 0; --i) {
      uint64_t sub = orc_prev_pow2(remaining);
      node_len = sub >> i;
      if (symbol <= sub) break;
      remaining -= sub;
      prev += sub / node_len;
      symbol -= sub;
    }
  }
  return symbol / node_len + prev - 1;
}

/* num_nodes_at_level loop, wavelet_tree.cuh:704-715 (array of L-1) */
void orc_num_nodes_at_level(const uint8_t* lv, uint64_t num, uint32_t L,
                            uint64_t* out) {
  for (uint32_t i = 0; i + 1 < L; ++i) out[i] = 0;
  uint64_t counter = 0;
  for (uint64_t i = 0; i < num; ++i) {
    if (i > 0 && lv[i] > lv[i - 1]) {
      if ((uint32_t)(lv[i] - 1) < L - 1) out[lv[i] - 1] = counter;
      counter = 0;
    }
    counter++;
  }
}

/* ------------------------------------------------------------------------
 * Binary rank/select, RankSelectHelper semantics (test_rank_select.cu:
 * 113-175): rank counts [0, i); select is 1-based, 0-based result, returns
 * m when the k-th value does not exist.
 * ---------------------------------------------------------------------- */
static inline int bit_at(const uint32_t* w, uint64_t i) {
  return (w[i >> 5] >> (i & 31)) & 1;
}
uint64_t orc_bits_rank(const uint32_t* w, uint64_t m, int value, uint64_t i) {
  if (i > m) i = m;
  uint64_t ones = 0, full = i >> 5;
  for (uint64_t j = 0; j < full; ++j) ones += __builtin_popcount(w[j]);
  if (i & 31) ones += __builtin_popcount(w[full] & ((1u << (i & 31)) - 1));
  return value ? ones : i - ones;
}
uint64_t orc_bits_select(const uint32_t* w, uint64_t m, int value,
                         uint64_t k) {
  if (k == 0) return m;
  uint64_t seen = 0;
  uint64_t nw = (m + 31) / 32;
  for (uint64_t j = 0; j < nw; ++j) {
    uint32_t x = value ? w[j] : ~w[j];
    if (j == nw - 1 && (m & 31)) x &= (1u << (m & 31)) - 1;
    uint32_t c = __builtin_popcount(x);
    if (seen + c >= k) {
      for (int b = 0; b < 32; ++b) {
        if ((x >> b) & 1) {
          if (++seen == k) return j * 32 + b;
        }
      }
    }
    seen += c;
  }
  return m;
}

/* RankSelect constructor sizes, rank_select.cu:57-117 */
uint64_t orc_rs_num_l1(uint64_t m) { return (m + 65535) / 65536; }
uint64_t orc_rs_num_last_l2(uint64_t m) {
  uint64_t last = ((m % 65536) + 511) / 512;
  return last == 0 ? 128 : last;
}
uint64_t orc_rs_num_l2(uint64_t m) {
  return (orc_rs_num_l1(m) - 1) * 128 + orc_rs_num_last_l2(m);
}

/* Index content: L1[k] = rank1(65536k), L2[j] = rank1(512j) -
 * rank1(65536*floor(j/128)) (test_rank_select.cu:452-476), select samples
 * S_v[j] = select_v((j+1)*4096) (rank_select.cu:554-578). */
uint64_t orc_rs_index(const uint32_t* w, uint64_t m, uint64_t* l1,
                      uint16_t* l2, uint64_t* s1, uint64_t* s0) {
  uint64_t n1 = orc_rs_num_l1(m), n2 = orc_rs_num_l2(m);
  uint64_t nw = (m + 31) / 32;
  uint64_t ones = 0, zeros = 0, cur_l1 = 0;
  uint64_t ns1 = 0, ns0 = 0;
  for (uint64_t j = 0; j < nw; ++j) {
    uint64_t bitpos = j * 32;
    if (bitpos % 65536 == 0) {
      uint64_t k = bitpos / 65536;
      if (k < n1 && l1) l1[k] = ones;
      cur_l1 = ones;
    }
    if (bitpos % 512 == 0) {
      uint64_t k = bitpos / 512;
      if (k < n2 && l2) l2[k] = (uint16_t)(ones - cur_l1);
    }
    uint32_t x = w[j];
    int valid = (j == nw - 1 && (m & 31)) ? (int)(m & 31) : 32;
    for (int b = 0; b < valid; ++b) {
      if ((x >> b) & 1) {
        ones++;
        if (ones % 4096 == 0) {
          if (s1) s1[ns1] = bitpos + b;
          ns1++;
        }
      } else {
        zeros++;
        if (zeros % 4096 == 0) {
          if (s0) s0[ns0] = bitpos + b;
          ns0++;
        }
      }
    }
  }
  return ones;
}

/* ------------------------------------------------------------------------
 * Whole tree.  Level sizes (wavelet_tree.cuh:776-790), level order
 * (:806-818: stable sort of the text by the top l code bits, first n_l
 * elements), level bits (fillLevelKernel :241-323: bit L-1-l of the code,
 * LSB-first in u32 words), node ranks (precomputeRanksKernel :332-352).
 * ---------------------------------------------------------------------- */
void orc_level_order(const uint64_t* sym, uint64_t n, const uint64_t* code,
                     const uint8_t* len, uint32_t L, uint32_t level,
                     uint64_t* order, uint64_t* count_out) {
  /* counting sort of present elements by the top `level` bits */
  uint64_t buckets = 1ULL << level;
  uint64_t* cnt = (uint64_t*)calloc(buckets + 1, sizeof(uint64_t));
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t c = sym[i];
    if (len[c] > level) cnt[(code[c] >> (L - level)) + 1]++;
  }
  for (uint64_t b = 0; b < buckets; ++b) cnt[b + 1] += cnt[b];
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t c = sym[i];
    if (len[c] > level) order[cnt[code[c] >> (L - level)]++] = i;
  }
  *count_out = cnt[buckets - 1];
  free(cnt);
}

int orc_wt_build(orc_wt* wt, const uint64_t* sym, uint64_t n,
                 uint64_t sigma) {
  memset(wt, 0, sizeof(*wt));
  wt->n = n;
  wt->sigma = sigma;
  uint32_t L = orc_ceil_log2(sigma);
  wt->num_levels = L;
  wt->code = (uint64_t*)malloc(sigma * sizeof(uint64_t));
  wt->len = (uint8_t*)malloc(sigma);
  uint64_t total_codes = orc_minimal_codes(sigma, wt->code, wt->len);
  wt->codes_start = sigma - total_codes;

  uint64_t* hist = (uint64_t*)calloc(sigma, sizeof(uint64_t));
  for (uint64_t i = 0; i < n; ++i) hist[sym[i]]++;
  wt->counts = (uint64_t*)malloc(sigma * sizeof(uint64_t));
  uint64_t acc = 0;
  for (uint64_t c = 0; c < sigma; ++c) {
    wt->counts[c] = acc;
    acc += hist[c];
  }
  wt->level_size = (uint64_t*)malloc(L * sizeof(uint64_t));
  for (uint32_t l = 0; l < L; ++l) {
    uint64_t s = n;
    for (uint64_t c = 0; c < sigma; ++c)
      if (wt->len[c] <= l) s -= hist[c];
    wt->level_size[l] = s;
  }
  wt->level_words = (uint32_t**)malloc(L * sizeof(uint32_t*));
  uint64_t* order = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
  for (uint32_t l = 0; l < L; ++l) {
    uint64_t cnt;
    orc_level_order(sym, n, wt->code, wt->len, L, l, order, &cnt);
    uint64_t nw = (cnt + 31) / 32;
    uint32_t* w = (uint32_t*)calloc(nw ? nw : 1, sizeof(uint32_t));
    for (uint64_t j = 0; j < cnt; ++j) {
      uint64_t c = sym[order[j]];
      if ((wt->code[c] >> (L - 1 - l)) & 1) w[j >> 5] |= 1u << (j & 31);
    }
    wt->level_words[l] = w;
    if (cnt != wt->level_size[l]) return -1;
  }
  free(order);
  free(hist);

  wt->num_ranks = orc_node_infos(sigma, wt->code, wt->len, NULL, NULL);
  uint64_t nr = wt->num_ranks ? wt->num_ranks : 1;
  wt->node_start = (uint64_t*)malloc(nr * sizeof(uint64_t));
  wt->node_level = (uint8_t*)malloc(nr);
  wt->ranks = (uint64_t*)malloc(nr * sizeof(uint64_t));
  orc_node_infos(sigma, wt->code, wt->len, wt->node_start, wt->node_level);
  for (uint64_t k = 0; k < wt->num_ranks; ++k) {
    uint32_t l = wt->node_level[k];
    wt->ranks[k] = orc_bits_rank(wt->level_words[l], wt->level_size[l], 0,
                                 wt->counts[wt->node_start[k]]);
  }
  wt->nodes_at_level =
      (uint64_t*)calloc(L > 1 ? L - 1 : 1, sizeof(uint64_t));
  if (orc_is_pow2(sigma)) {
    for (uint32_t l = 0; l + 1 < L; ++l)
      wt->nodes_at_level[l] = (1ULL << l) - 1;
    wt->nodes_until_last_level =
        wt->num_ranks - ((L >= 1 ? (1ULL << (L - 1)) : 1) - 1);
  } else {
    orc_num_nodes_at_level(wt->node_level, wt->num_ranks, L,
                           wt->nodes_at_level);
    wt->nodes_until_last_level = 0;
    for (uint32_t l = 0; l + 1 < L; ++l)
      wt->nodes_until_last_level += wt->nodes_at_level[l];
  }
  return 0;
}

void orc_wt_free(orc_wt* wt) {
  if (wt->level_words)
    for (uint32_t l = 0; l < wt->num_levels; ++l) free(wt->level_words[l]);
  free(wt->level_words);
  free(wt->code);
  free(wt->len);
  free(wt->counts);
  free(wt->level_size);
  free(wt->node_start);
  free(wt->node_level);
  free(wt->ranks);
  free(wt->nodes_at_level);
  memset(wt, 0, sizeof(*wt));
}

static uint64_t wt_total(const orc_wt* wt, uint64_t c) {
  /* getTotalAppearances, wavelet_tree.cuh:2207-2215 */
  return c == wt->sigma - 1 ? wt->n - wt->counts[c]
                            : wt->counts[c + 1] - wt->counts[c];
}
static uint64_t lvl_nodes(const orc_wt* wt, uint32_t l) {
  return orc_is_pow2(wt->sigma) ? (1ULL << l) - 1 : wt->nodes_at_level[l];
}

/* deviceAccess, wavelet_tree.cuh:1815-1877 */
uint64_t orc_wt_access(const orc_wt* wt, uint64_t index) {
  uint64_t lo = 0, hi = wt->sigma - 1, rs = 0;
  uint32_t L = wt->num_levels;
  for (uint32_t l = 0; l < L; ++l) {
    uint64_t cc = wt->counts[lo];
    const uint32_t* w = wt->level_words[l];
    if (hi - lo < 2) {
      if (hi - lo > 0 && bit_at(w, cc + index)) lo++;
      break;
    }
    uint64_t start =
        lo == 0 ? 0
                : wt->ranks[rs + orc_node_pos_at_level(wt->sigma, L, lo, l)];
    uint64_t pos = orc_bits_rank(w, wt->level_size[l], 0, cc + index);
    int bit = bit_at(w, cc + index);
    uint64_t diff = orc_prev_pow2(hi - lo + 1);
    if (!bit) {
      index = pos - start;
      hi = lo + diff - 1;
    } else {
      index -= pos - start;
      lo += diff;
    }
    rs += lvl_nodes(wt, l);
  }
  return lo;
}

/* rankKernel early-out (wavelet_tree.cuh:481-486) + deviceRank (:1896-1956) */
uint64_t orc_wt_rank(const orc_wt* wt, uint64_t c, uint64_t index) {
  if (wt_total(wt, c) == 0) return 0;
  uint64_t lo = 0, hi = wt->sigma - 1, rs = 0, result = index;
  uint32_t L = wt->num_levels;
  for (uint32_t l = 0; l < L; ++l) {
    if (hi == lo) break;
    uint64_t cc = wt->counts[lo];
    uint64_t start =
        lo == 0 ? 0
                : wt->ranks[rs + orc_node_pos_at_level(wt->sigma, L, lo, l)];
    uint64_t pos =
        orc_bits_rank(wt->level_words[l], wt->level_size[l], 0, cc + result);
    uint64_t split = lo + orc_prev_pow2(hi - lo + 1);
    if (c < split) {
      result = pos - start;
      hi = split - 1;
    } else {
      result -= pos - start;
      lo = split;
    }
    if (l < L - 1) rs += lvl_nodes(wt, l);
  }
  return result;
}

/* getPrevCharStart, wavelet_tree.cuh:1970-1992 */
static uint64_t prev_char_start(const orc_wt* wt, uint64_t cs, int rightmost,
                                uint32_t level, uint32_t code_len) {
  uint32_t L = wt->num_levels;
  if (orc_is_pow2(wt->sigma)) {
    uint64_t node_lens = 1ULL << (L - level);
    return cs & ~(node_lens - 1);
  }
  if (rightmost) {
    uint64_t s = 0;
    for (uint32_t l = 0; l < level; ++l) s += orc_prev_pow2(wt->sigma - s);
    return s;
  }
  return cs - (1ULL << (code_len - 1 - level));
}

/* selectKernel early-out (:532-535) + deviceSelect (:2007-2079) */
uint64_t orc_wt_select(const orc_wt* wt, uint64_t c, uint64_t k) {
  if (k > wt_total(wt, c)) return wt->n;
  uint32_t L = wt->num_levels;
  int p2 = orc_is_pow2(wt->sigma);
  uint64_t code = c, clen = L;
  if (!p2 && c >= wt->codes_start) {
    code = wt->code[c];
    clen = wt->len[c];
  }
  uint64_t cs = c, result = k;
  uint64_t rs = wt->nodes_until_last_level;
  int32_t l = (int32_t)clen - 1;
  for (int32_t lv = (int32_t)L - 2; lv >= l; --lv) rs -= lvl_nodes(wt, lv);
  for (; l >= 0; --l) {
    const uint32_t* w = wt->level_words[l];
    uint64_t m = wt->level_size[l];
    uint64_t start;
    if ((code >> (L - 1 - l)) & 1) {
      int rightmost =
          (uint32_t)__builtin_popcountll(l == 0 ? 0 : (code >> (L - l))) ==
          (uint32_t)l;
      cs = prev_char_start(wt, cs, rightmost, (uint32_t)l, (uint32_t)clen);
      start = cs == 0 ? 0
                      : wt->counts[cs] -
                            wt->ranks[rs + orc_node_pos_at_level(
                                               wt->sigma, L, cs, l)];
      result = orc_bits_select(w, m, 1, start + result) + 1;
    } else {
      start = cs == 0
                  ? 0
                  : wt->ranks[rs + orc_node_pos_at_level(wt->sigma, L, cs, l)];
      result = orc_bits_select(w, m, 0, start + result) + 1;
    }
    if ((uint64_t)l == clen - 1 && result > m) {
      result = wt->n + 1;
      break;
    }
    result -= wt->counts[cs];
    if (l > 1) rs -= lvl_nodes(wt, l - 1);
  }
  return result - 1;
}

/* ------------------------------------------------------------------------
 * Direct scans: rank = std::count over [0,i) (test_wavelet_tree.cu:617-619),
 * select = position of the k-th occurrence via find_if (test_utils.cu:
 * 81-91), n when absent (test_wavelet_tree.cu:674-678).
 * ---------------------------------------------------------------------- */
/* occurrences of byte c in t[0, len): 16-byte compares accumulated in byte
 * lanes, 255 rounds of 64 bytes at a time (the count of one lane stays
 * below 256), then a byte loop for the tail -- a vectorizable form of the
 * same count */
typedef uint8_t orc_v16 __attribute__((vector_size(16)));
#define ORC_COUNT_BLOCK (64u * 255u)
static uint64_t count_u8(const uint8_t* t, uint64_t len, uint8_t c) {
  orc_v16 cv;
  memset(&cv, c, sizeof(cv));
  uint64_t r = 0, j = 0;
  while (j + ORC_COUNT_BLOCK <= len) {
    orc_v16 a0 = {0}, a1 = {0}, a2 = {0}, a3 = {0};
    for (unsigned k = 0; k < 255; ++k, j += 64) {
      orc_v16 x0, x1, x2, x3;
      memcpy(&x0, t + j, 16);
      memcpy(&x1, t + j + 16, 16);
      memcpy(&x2, t + j + 32, 16);
      memcpy(&x3, t + j + 48, 16);
      a0 -= (orc_v16)(x0 == cv);
      a1 -= (orc_v16)(x1 == cv);
      a2 -= (orc_v16)(x2 == cv);
      a3 -= (orc_v16)(x3 == cv);
    }
    for (int k = 0; k < 16; ++k) r += (uint64_t)a0[k] + a1[k] + a2[k] + a3[k];
  }
  for (; j < len; ++j) r += t[j] == c;
  return r;
}

uint64_t orc_scan_rank(const void* text, int esize, uint64_t n, uint64_t c,
                       uint64_t i) {
  uint64_t r = 0;
  if (i > n) i = n;
  switch (esize) {
    case 1: {
      if (c < 256) r = count_u8((const uint8_t*)text, i, (uint8_t)c);
      break;
    }
    case 2: {
      const uint16_t* t = (const uint16_t*)text;
      for (uint64_t j = 0; j < i; ++j) r += t[j] == c;
      break;
    }
    default:
      for (uint64_t j = 0; j < i; ++j) r += get_elem(text, esize, j) == c;
  }
  return r;
}
uint64_t orc_scan_select(const void* text, int esize, uint64_t n, uint64_t c,
                         uint64_t k) {
  uint64_t seen = 0;
  if (k == 0) return n;
  switch (esize) {
    case 1: {
      /* whole blocks are counted while the k-th occurrence lies past them;
       * the block that holds it is walked byte by byte */
      const uint8_t* t = (const uint8_t*)text;
      if (c >= 256) break;
      uint64_t j = 0;
      while (j + ORC_COUNT_BLOCK <= n) {
        const uint64_t b = count_u8(t + j, ORC_COUNT_BLOCK, (uint8_t)c);
        if (seen + b >= k) break;
        seen += b;
        j += ORC_COUNT_BLOCK;
      }
      for (; j < n; ++j)
        if (t[j] == c && ++seen == k) return j;
      break;
    }
    case 2: {
      const uint16_t* t = (const uint16_t*)text;
      for (uint64_t j = 0; j < n; ++j)
        if (t[j] == c && ++seen == k) return j;
      break;
    }
    default:
      for (uint64_t j = 0; j < n; ++j)
        if (get_elem(text, esize, j) == c && ++seen == k) return j;
  }
  return n;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_scan_access_batch(const void* text, int esize, uint64_t n,
                          const uint64_t* idx, uint64_t m, uint64_t* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < (int64_t)m; ++q)
    out[q] = idx[q] < n ? get_elem(text, esize, idx[q]) : 0;
  return orc_max_threads();
}
int orc_scan_rank_batch(const void* text, int esize, uint64_t n,
                        const uint64_t* idx, const uint64_t* sym, uint64_t m,
                        uint64_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < (int64_t)m; ++q)
    out[q] = orc_scan_rank(text, esize, n, sym[q], idx[q]);
  return orc_max_threads();
}
int orc_scan_select_batch(const void* text, int esize, uint64_t n,
                          const uint64_t* k, const uint64_t* sym, uint64_t m,
                          uint64_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < (int64_t)m; ++q)
    out[q] = orc_scan_select(text, esize, n, sym[q], k[q]);
  return orc_max_threads();
}
/* op: 0 = rank0, 1 = rank1, 2 = select0, 3 = select1 */
int orc_scan_bits_batch(const uint32_t* w, uint64_t m, int op,
                        const uint64_t* arg, uint64_t q, uint64_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t j = 0; j < (int64_t)q; ++j)
    out[j] = op < 2 ? orc_bits_rank(w, m, op & 1, arg[j])
                    : orc_bits_select(w, m, op & 1, arg[j]);
  return orc_max_threads();
}

void orc_histogram(const void* text, int esize, uint64_t n, uint64_t sigma,
                   uint64_t* hist) {
  memset(hist, 0, sigma * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t c = get_elem(text, esize, i);
    if (c < sigma) hist[c]++;
  }
}

/* ------------------------------------------------------------------------
 * Sweep rank: the answers of many rank queries from one pass over the text.
 * out[j] = count of sym[j] in [0, pos[j]) -- the direct scan's rank
 * (test_wavelet_tree.cu:617-619) -- for positions given in ascending order
 * (pos[j] <= n).  The text is cut into one range per thread; each range's
 * histogram is counted, the histograms are prefixed over the ranges, and
 * each thread then walks its range with a running histogram that starts at
 * that prefix, answering its queries as it reaches their positions.  This
 * is how the by-construction queries of test_benchmark_utils.cuh:417-528
 * get their answers (a running histogram while the text is generated),
 * here for arbitrary (symbol, position) pairs.  O(n + m) time.
 * ---------------------------------------------------------------------- */
static uint64_t lower_pos(const uint64_t* pos, uint64_t m, uint64_t x) {
  uint64_t lo = 0, hi = m;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (pos[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

int orc_sweep_rank(const void* text, int esize, uint64_t n, uint64_t sigma,
                   const uint64_t* pos, const uint64_t* sym, uint64_t m,
                   uint64_t* out) {
  int T = orc_max_threads();
  if ((uint64_t)T > n / 4096 + 1) T = (int)(n / 4096 + 1);
  uint64_t* hist = (uint64_t*)calloc((size_t)(T + 1) * sigma, sizeof(uint64_t));
  if (!hist) return -1;
  /* range t covers [t*n/T, (t+1)*n/T); row t+1 of hist first gets range t's
   * histogram, then the prefix turns row t into the counts before range t */
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; ++t) {
    uint64_t lo = (uint64_t)t * n / T, hi = (uint64_t)(t + 1) * n / T;
    uint64_t* h = hist + (size_t)(t + 1) * sigma;
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t c = get_elem(text, esize, i);
      if (c < sigma) h[c]++;
    }
  }
  for (int t = 1; t <= T; ++t)
    for (uint64_t c = 0; c < sigma; ++c)
      hist[(size_t)t * sigma + c] += hist[(size_t)(t - 1) * sigma + c];
  /* queries at position n take the totals (row T) */
  uint64_t first_n = lower_pos(pos, m, n);
  for (uint64_t j = first_n; j < m; ++j)
    out[j] = sym[j] < sigma ? hist[(size_t)T * sigma + sym[j]] : 0;
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; ++t) {
    uint64_t lo = (uint64_t)t * n / T, hi = (uint64_t)(t + 1) * n / T;
    uint64_t* h = hist + (size_t)t * sigma; /* counts before lo */
    uint64_t q = lower_pos(pos, m, lo), qe = lower_pos(pos, m, hi);
    uint64_t i = lo;
    for (; q < qe; ++q) {
      for (; i < pos[q]; ++i) {
        uint64_t c = get_elem(text, esize, i);
        if (c < sigma) h[c]++;
      }
      out[q] = sym[q] < sigma ? h[sym[q]] : 0;
    }
  }
  free(hist);
  return T;
}

/* ------------------------------------------------------------------------
 * Blocked oracle (CPU baseline, variant B): the direct scans' answers
 * (test_wavelet_tree.cu:617-678, test_utils.cu:81-91) from per-symbol
 * prefix counts every 2^16 symbols plus a scan of at most one block.
 * table has (ceil(n / 2^16) + 1) rows of sigma u64: row b holds the counts
 * of every symbol in [0, min(b * 2^16, n)), so the last row holds the
 * totals.  Same answers as variant A, including n for a select past the
 * last occurrence.
 * ---------------------------------------------------------------------- */
#define ORC_BLOCK_LOG2 16
uint64_t orc_blocked_rows(uint64_t n) {
  return ((n + ((1ull << ORC_BLOCK_LOG2) - 1)) >> ORC_BLOCK_LOG2) + 1;
}

void orc_blocked_build(const void* text, int esize, uint64_t n, uint64_t sigma,
                       uint64_t* table) {
  const uint64_t rows = orc_blocked_rows(n);
  memset(table, 0, sigma * sizeof(uint64_t));
  /* row b + 1 first holds block b's own histogram */
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t b = 0; b < (int64_t)(rows - 1); ++b) {
    uint64_t* h = table + (size_t)(b + 1) * sigma;
    memset(h, 0, sigma * sizeof(uint64_t));
    uint64_t lo = (uint64_t)b << ORC_BLOCK_LOG2;
    uint64_t hi = lo + (1ull << ORC_BLOCK_LOG2);
    if (hi > n) hi = n;
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t c = get_elem(text, esize, i);
      if (c < sigma) h[c]++;
    }
  }
  /* prefix over the rows, one symbol column per iteration */
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)sigma; ++c)
    for (uint64_t b = 1; b < rows; ++b)
      table[b * sigma + c] += table[(b - 1) * sigma + c];
}

uint64_t orc_blocked_rank(const void* text, int esize, uint64_t n, uint64_t sigma,
                          const uint64_t* table, uint64_t c, uint64_t i) {
  if (c >= sigma) return 0;
  if (i > n) i = n;
  const uint64_t b = i >> ORC_BLOCK_LOG2;
  uint64_t r = table[b * sigma + c];
  for (uint64_t j = b << ORC_BLOCK_LOG2; j < i; ++j) r += get_elem(text, esize, j) == c;
  return r;
}

uint64_t orc_blocked_select(const void* text, int esize, uint64_t n, uint64_t sigma,
                            const uint64_t* table, uint64_t c, uint64_t k) {
  const uint64_t rows = orc_blocked_rows(n);
  if (k == 0 || c >= sigma || k > table[(rows - 1) * sigma + c]) return n;
  /* largest block b with fewer than k occurrences before it */
  uint64_t lo = 0, hi = rows - 2;
  while (lo < hi) {
    uint64_t mid = (lo + hi + 1) / 2;
    if (table[mid * sigma + c] < k) lo = mid; else hi = mid - 1;
  }
  uint64_t seen = table[lo * sigma + c];
  uint64_t end = (lo + 1) << ORC_BLOCK_LOG2;
  if (end > n) end = n;
  for (uint64_t j = lo << ORC_BLOCK_LOG2; j < end; ++j)
    if (get_elem(text, esize, j) == c && ++seen == k) return j;
  return n; /* unreachable for a consistent table */
}

int orc_blocked_rank_batch(const void* text, int esize, uint64_t n, uint64_t sigma,
                           const uint64_t* table, const uint64_t* idx,
                           const uint64_t* sym, uint64_t m, uint64_t* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < (int64_t)m; ++q)
    out[q] = orc_blocked_rank(text, esize, n, sigma, table, sym[q], idx[q]);
  return orc_max_threads();
}

int orc_blocked_select_batch(const void* text, int esize, uint64_t n, uint64_t sigma,
                             const uint64_t* table, const uint64_t* k,
                             const uint64_t* sym, uint64_t m, uint64_t* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < (int64_t)m; ++q)
    out[q] = orc_blocked_select(text, esize, n, sigma, table, sym[q], k[q]);
  return orc_max_threads();
}
