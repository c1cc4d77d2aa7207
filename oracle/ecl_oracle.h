/*
 * ecl_oracle.h -- CPU oracle for the wavelet-tree / rank-select hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (ecl_wavelet_amd, libecl_wavelet_hip.so) never links or calls it.
 *
 * Every function restates the behaviour of the reference library
 * (mfranzreb/ECL_wavelet); the file:line it follows is cited at each
 * definition in ecl_oracle.c.  The checks that pin each function are in
 * tests/test_oracle.py.
 */
#ifndef ECL_ORACLE_H
#define ECL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- deterministic counter-based generator (shared with the GPU side) --- */
uint64_t orc_hash(uint64_t seed, uint64_t i);
uint64_t orc_bounded(uint64_t h, uint64_t m);
/* text[i] = bounded(hash(seed, i), sigma) */
void orc_gen_uniform(uint64_t seed, uint64_t n, uint64_t sigma, int esize,
                     void* out);
/* text[i] = first s with hash(seed, i) < thresholds[s] (thresholds[sigma-1]
 * must be UINT64_MAX) */
void orc_gen_cdf(uint64_t seed, uint64_t n, const uint64_t* thresholds,
                 uint64_t sigma, int esize, void* out);
/* bit i = hash(seed, i) < threshold; words are LSB-first u32 */
void orc_gen_bits(uint64_t seed, uint64_t nbits, uint64_t threshold,
                  uint32_t* words);

/* ---- integer helpers (utils.cuh:123-240) ------------------------------- */
int orc_is_pow2(uint64_t n);
uint64_t orc_prev_pow2(uint64_t n);
uint32_t orc_ceil_log2(uint64_t n);

/* ---- alphabet codes / node layout (wavelet_tree.cuh) ------------------- */
/* Fills code[c], len[c] for every symbol 0..sigma-1 (symbols below the
 * codes start keep code=c, len=L).  Returns the number of codes the
 * reference keeps (0 for a power-of-two alphabet). */
uint64_t orc_minimal_codes(uint64_t sigma, uint64_t* code, uint8_t* len);
/* Node starts ordered by (level, symbol); returns the count.  Output arrays
 * may be NULL to only count. */
uint64_t orc_node_infos(uint64_t sigma, const uint64_t* code,
                        const uint8_t* len, uint64_t* start, uint8_t* level);
uint64_t orc_node_pos_at_level(uint64_t sigma, uint32_t num_levels,
                               uint64_t symbol, uint32_t level);
/* the reference's num_nodes_at_level loop (size L-1) */
void orc_num_nodes_at_level(const uint8_t* node_level, uint64_t num_nodes,
                            uint32_t num_levels, uint64_t* out);

/* ---- binary rank/select over LSB-first u32 words (test_rank_select.cu) -- */
uint64_t orc_bits_rank(const uint32_t* words, uint64_t m, int value,
                       uint64_t i);
uint64_t orc_bits_select(const uint32_t* words, uint64_t m, int value,
                         uint64_t k);
/* RankSelect index content (rank_select.cu:57-143, 412-591) */
uint64_t orc_rs_num_l1(uint64_t m);
uint64_t orc_rs_num_last_l2(uint64_t m);
uint64_t orc_rs_num_l2(uint64_t m);
/* writes l1[num_l1], l2[num_l2], s1[ones/4096], s0[zeros/4096];
 * returns total ones.  Any output may be NULL. */
uint64_t orc_rs_index(const uint32_t* words, uint64_t m, uint64_t* l1,
                      uint16_t* l2, uint64_t* s1, uint64_t* s0);

/* ---- whole wavelet tree (minimal-alphabet symbols) ---------------------- */
typedef struct {
  uint64_t n, sigma;
  uint32_t num_levels;
  uint64_t codes_start;
  uint64_t* code; /* [sigma] */
  uint8_t* len;   /* [sigma] */
  uint64_t* counts;      /* [sigma] exclusive prefix (C array) */
  uint64_t* level_size;  /* [L] */
  uint32_t** level_words; /* [L][ceil(size/32)] */
  uint64_t num_ranks;
  uint64_t* node_start;  /* [num_ranks] */
  uint8_t* node_level;   /* [num_ranks] */
  uint64_t* ranks;       /* [num_ranks] precomputed node ranks */
  uint64_t* nodes_at_level; /* [L-1] */
  uint64_t nodes_until_last_level;
} orc_wt;

int orc_wt_build(orc_wt* wt, const uint64_t* sym, uint64_t n, uint64_t sigma);
void orc_wt_free(orc_wt* wt);
uint64_t orc_wt_access(const orc_wt* wt, uint64_t i);
uint64_t orc_wt_rank(const orc_wt* wt, uint64_t c, uint64_t i);
uint64_t orc_wt_select(const orc_wt* wt, uint64_t c, uint64_t k);
/* stable order of the text by the top l code bits, restricted to the
 * elements present at level l (for the level-order pin) */
void orc_level_order(const uint64_t* sym, uint64_t n, const uint64_t* code,
                     const uint8_t* len, uint32_t num_levels, uint32_t level,
                     uint64_t* order_out, uint64_t* count_out);

/* ---- direct scans (test_wavelet_tree.cu:617-678, test_utils.cu:81-91) -- */
uint64_t orc_scan_rank(const void* text, int esize, uint64_t n, uint64_t c,
                       uint64_t i);
uint64_t orc_scan_select(const void* text, int esize, uint64_t n, uint64_t c,
                         uint64_t k);
/* OpenMP-parallel batches (CPU baseline, variant A).  Returns threads used. */
int orc_scan_access_batch(const void* text, int esize, uint64_t n,
                          const uint64_t* idx, uint64_t m, uint64_t* out);
int orc_scan_rank_batch(const void* text, int esize, uint64_t n,
                        const uint64_t* idx, const uint64_t* sym, uint64_t m,
                        uint64_t* out);
int orc_scan_select_batch(const void* text, int esize, uint64_t n,
                          const uint64_t* k, const uint64_t* sym, uint64_t m,
                          uint64_t* out);
int orc_scan_bits_batch(const uint32_t* words, uint64_t m, int op,
                        const uint64_t* arg, uint64_t q, uint64_t* out);
int orc_max_threads(void);
/* histogram of min-alphabet symbols (counts, not prefix) */
void orc_histogram(const void* text, int esize, uint64_t n, uint64_t sigma,
                   uint64_t* hist);

/* Sweep rank (by-construction answers, test_benchmark_utils.cuh:417-528):
 * out[j] = count of sym[j] in [0, pos[j]) for positions in ascending order,
 * pos[j] <= n; one pass over the text.  Returns threads used (-1: no
 * memory). */
int orc_sweep_rank(const void* text, int esize, uint64_t n, uint64_t sigma,
                   const uint64_t* pos, const uint64_t* sym, uint64_t m,
                   uint64_t* out);
/* Blocked oracle (CPU baseline, variant B): per-symbol counts before every
 * 2^16-symbol block, orc_blocked_rows(n) rows of sigma u64, plus a scan of
 * at most one block per query.  Same answers as the direct scans. */
uint64_t orc_blocked_rows(uint64_t n);
void orc_blocked_build(const void* text, int esize, uint64_t n, uint64_t sigma,
                       uint64_t* table);
uint64_t orc_blocked_rank(const void* text, int esize, uint64_t n, uint64_t sigma,
                          const uint64_t* table, uint64_t c, uint64_t i);
uint64_t orc_blocked_select(const void* text, int esize, uint64_t n, uint64_t sigma,
                            const uint64_t* table, uint64_t c, uint64_t k);
int orc_blocked_rank_batch(const void* text, int esize, uint64_t n, uint64_t sigma,
                           const uint64_t* table, const uint64_t* idx,
                           const uint64_t* sym, uint64_t m, uint64_t* out);
int orc_blocked_select_batch(const void* text, int esize, uint64_t n, uint64_t sigma,
                             const uint64_t* table, const uint64_t* k,
                             const uint64_t* sym, uint64_t m, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif
