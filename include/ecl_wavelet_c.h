/*
 * ecl_wavelet_c.h -- C ABI of libecl_wavelet_hip.so, the MI355X (gfx950)
 * wavelet-tree / rank-select library.
 *
 * Plain pointers and sizes only.  Every entry point returns an ecl_status;
 * ecl_last_error() gives the message of the calling thread's last failure.
 * The out-of-memory messages are the reference's exact strings, so C++
 * wrappers can rethrow them unchanged.
 *
 * Each declaration names the reference (mfranzreb/ECL_wavelet) interface it
 * replaces.  Entry points marked "new" have no reference counterpart.
 */
#ifndef ECL_WAVELET_C_H
#define ECL_WAVELET_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ECL_WAVELET_ABI_VERSION 1

typedef enum ecl_status {
  ECL_OK = 0,
  ECL_ERROR_INVALID_ARGUMENT = 1,
  /* "Not enough memory available for the wavelet tree."
   * (tree/wavelet_tree.cuh:652-653, 686-687) */
  ECL_ERROR_OUT_OF_MEMORY_TREE = 2,
  /* "Not enough GPU memory available for the {access,rank,select} queries."
   * (tree/wavelet_tree.cuh:972-973, 1268-1269, 1587-1588) */
  ECL_ERROR_OUT_OF_MEMORY_QUERIES = 3,
  ECL_ERROR_HIP = 4,
  ECL_ERROR_NO_DEVICE = 5,
  ECL_ERROR_INTERNAL = 6
} ecl_status;

/* element type of a text: sizeof(T) of WaveletTree<T> */
typedef enum ecl_type { ECL_U8 = 1, ECL_U16 = 2, ECL_U32 = 4, ECL_U64 = 8 } ecl_type;

/* query flags */
#define ECL_SYMBOLS_MINIMAL 1u /* query symbols are already minimal-alphabet
                                  indices (as after the reference's in-place
                                  remap, wavelet_tree.cuh:1426-1436) */
#define ECL_RESULTS_MINIMAL 2u /* access returns minimal-alphabet indices
                                  (deviceAccess semantics) */

typedef struct ecl_bit_array_t* ecl_bit_array;
typedef struct ecl_rank_select_t* ecl_rank_select;
typedef struct ecl_wavelet_tree_t* ecl_wavelet_tree;

/* RankSelectQuery<T> (utils/utils.cuh:70-74): 16 bytes for every T, the
 * symbol occupies the first sizeof(T) bytes of `symbol`. */
typedef struct ecl_rank_select_query {
  uint64_t index;
  uint64_t symbol;
} ecl_rank_select_query;

/* new: one query of a mixed batch.  kind 0 = access(arg), 1 = rank(symbol,
 * arg), 2 = select(symbol, arg); symbol is a minimal-alphabet index. */
typedef struct ecl_query {
  uint64_t arg;
  uint32_t symbol;
  uint32_t kind;
} ecl_query;

/* new: one query of a batch over many trees of a forest (same kinds as
 * ecl_query; `tree` is the tree's index in the forest). */
typedef struct ecl_forest_query {
  uint64_t arg;
  uint32_t symbol;
  uint16_t kind;
  uint16_t reserved;
  uint32_t tree;
  uint32_t reserved2;
} ecl_forest_query;

typedef struct ecl_rank_select_info {
  uint64_t size_bits;     /* BitArray::size (bit_array.cuh:274-283) */
  uint64_t offset_words;  /* BitArray::getOffset (bit_array.cuh:160-164) */
  uint64_t num_l1_blocks; /* getNumL1Blocks (rank_select.cuh:530-535) */
  uint64_t num_l2_blocks; /* getNumL2Blocks (rank_select.cuh:542-550) */
  uint64_t num_last_l2_blocks; /* getNumLastL2Blocks (:557-561) */
  uint64_t total_ones;    /* getTotalNumVals<1> (:569-578) */
  uint64_t total_zeros;   /* getTotalNumVals<0> */
  uint64_t num_samples_1; /* floor(ones / 4096) */
  uint64_t num_samples_0; /* floor(zeros / 4096) */
} ecl_rank_select_info;

typedef struct ecl_wavelet_tree_info {
  uint64_t data_size;       /* rank_select_.bit_array_.size(0) */
  uint64_t alphabet_size;   /* getAlphabetSize (wavelet_tree.cuh:2165) */
  uint32_t num_levels;      /* getNumLevels (:2181) */
  uint32_t is_min_alphabet; /* isMinAlphabet (:2220) */
  uint64_t codes_start;     /* getCodesStart (:2228) */
  uint64_t num_ranks;       /* number of precomputed node ranks */
  uint64_t nodes_until_last_level;
  uint32_t element_size;    /* sizeof(T) */
  uint32_t is_pow_two;
} ecl_wavelet_tree_info;

/* ---- library / device helpers (new) ------------------------------------ */
const char* ecl_last_error(void);
int ecl_abi_version(void);
ecl_status ecl_device_count(int* count);
ecl_status ecl_device_alloc(int device, size_t bytes, void** ptr);
ecl_status ecl_device_free(int device, void* ptr);
ecl_status ecl_memcpy_h2d(int device, void* dst, const void* src, size_t bytes);
ecl_status ecl_memcpy_d2h(int device, void* dst, const void* src, size_t bytes);
ecl_status ecl_device_synchronize(int device);
/* free bytes: the runtime's free memory plus the library's cached blocks
 * (below), which the library reuses and frees when an allocation needs them */
ecl_status ecl_device_mem_info(int device, size_t* free_bytes, size_t* total_bytes);
/* The library keeps freed device blocks of 1 MiB or more per device and
 * hands them out again for requests of the same size class (up to 40% of
 * the device's memory; freed when an allocation would fail), and keeps the
 * construction scratch (tile histograms and their scan, about 0.1 bytes per
 * symbol) for the next build.  This frees both.  No counterpart in the
 * reference, which allocates and frees per call; for callers that need the
 * memory back outside the library. */
ecl_status ecl_release_cached_memory(int device);
/* Caps the free device memory the library's admission checks assume on a
 * device (tree builds, query batches, server budgets): they see at most
 * `bytes` free; 0 removes the cap.  No counterpart in the reference; for
 * sharing a device with other work, and to exercise the host-memory
 * fallback of the build (below) on a device that has room. */
ecl_status ecl_set_device_memory_cap(int device, uint64_t bytes);
/* Memory a tree build needs (upper bounds, bytes): `resident` must fit the
 * device -- bits, directory, construction scratch, metadata and the
 * reference's 10M-query headroom (wavelet_tree.cuh:644-648); otherwise the
 * build fails with ECL_ERROR_OUT_OF_MEMORY_TREE.  `staging` is a host text's
 * copy and the code buffers of the passes: the build places each on the
 * device while it fits after `resident`, and otherwise in pinned host memory
 * that the kernels read and write over the host link, as the reference
 * moves its data, sorted copy and sort scratch to unified memory
 * (wavelet_tree.cuh:655-689, 735-802); the host part must fit the host's
 * available memory (MemAvailable).  Results do not depend on the placement;
 * a build that stages on the host runs at the host link's bandwidth. */
ecl_status ecl_wavelet_tree_build_memory(ecl_type type, size_t n, size_t sigma, int host_text,
                                         uint64_t* resident, uint64_t* staging);
/* Bytes of the tree's build buffers that were placed in host memory (0 when
 * the whole build ran in device memory). */
ecl_status ecl_wavelet_tree_build_host_bytes(ecl_wavelet_tree wt, uint64_t* bytes);
/* Tiles of 32768 codes each construction workgroup takes in turn (0:
 * automatic, up to 16 for large texts).  A tuning knob with no effect on
 * results; no counterpart in the reference. */
ecl_status ecl_set_build_span(uint32_t tiles_per_workgroup);
/* Lanes per access/rank query in the tree query kernels: 0 (default,
 * automatic: 2 over query lines, 1 otherwise), 1, or 2 (the two lanes of a
 * pair each fetch half of every level's bits).  A tuning knob with no effect
 * on results; no counterpart in the reference. */
ecl_status ecl_set_query_lanes(uint32_t lanes);
ecl_status ecl_device_name(int device, char* buf, size_t len);
/* events on `stream` (NULL = default stream); elapsed milliseconds */
ecl_status ecl_event_create(void** event);
ecl_status ecl_event_destroy(void* event);
ecl_status ecl_event_record(void* event, void* stream);
ecl_status ecl_event_elapsed(void* start, void* stop, float* ms);
/* per-kernel timing: events around every launch of the library's kernels
 * while enabled; query by kernel name ("wt_level", "rs_build", ...) */
ecl_status ecl_profile_enable(int on);
ecl_status ecl_profile_reset(void);
ecl_status ecl_profile_get(const char* kernel, uint64_t* launches, double* total_ms);
/* The device's rate of random, dependent 128-byte line reads (lines per
 * second): `queries` probe queries over a buffer of `bytes`, each reading
 * `steps` lines one after the other at positions derived from the previous
 * line, with `lanes` (1, 2, 4 or 8) lanes per line.  The bound of the tree
 * queries, which read one such line per level (two for select); no
 * counterpart in the reference. */
ecl_status ecl_probe_random_lines(int device, uint64_t bytes, uint64_t queries, uint32_t steps,
                                  uint32_t lanes, double* lines_per_s);
/* deterministic synthetic inputs on the device: text[i] = bounded(hash(seed,
 * i), sigma), or the inverse CDF of `thresholds` (host array of sigma u64)
 * when non-NULL; bits i.i.d. with P(1) = threshold / 2^64 */
ecl_status ecl_generate_text(int device, ecl_type type, uint64_t seed, size_t n,
                             uint64_t sigma, const uint64_t* thresholds, void* d_out);

/* ---- BitArray (bitarray/bit_array.cuh, src/bitarray/bit_array.cu) ------- */
/* BitArray(sizes) / BitArray(sizes, init_value) (bit_array.cu:40-92);
 * init_value: -1 = uninitialised payload, 0 / 1 = every word set. */
ecl_status ecl_bit_array_create(const size_t* sizes, size_t num_arrays, int init_value,
                                int device, ecl_bit_array* out);
/* ~BitArray (bit_array.cu:127-133) */
ecl_status ecl_bit_array_destroy(ecl_bit_array ba);
/* numArrays / size / getOffset / sizeInWords (bit_array.cuh:160-320) */
ecl_status ecl_bit_array_num_arrays(ecl_bit_array ba, size_t* out);
ecl_status ecl_bit_array_size(ecl_bit_array ba, size_t array_index, size_t* bits);
ecl_status ecl_bit_array_offset(ecl_bit_array ba, size_t array_index, size_t* word_offset);
/* device pointers of d_data_, d_offsets_, d_bit_sizes_ (bit_array.cuh:336-345) */
ecl_status ecl_bit_array_device_pointers(ecl_bit_array ba, uint32_t** data,
                                         uint64_t** offsets, uint64_t** sizes);
/* writeWord / word over a host range (bit_array.cuh:117-220) */
ecl_status ecl_bit_array_write_words(ecl_bit_array ba, size_t array_index, size_t first_word,
                                     const uint32_t* words, size_t num_words);
ecl_status ecl_bit_array_read_words(ecl_bit_array ba, size_t array_index, size_t first_word,
                                    uint32_t* words, size_t num_words);
/* new: fill an array on the device with i.i.d. bits (hash(seed, i) < threshold) */
ecl_status ecl_bit_array_fill_random(ecl_bit_array ba, size_t array_index, uint64_t seed,
                                     uint64_t threshold);
/* BitArray::getNeededGPUMemory (bit_array.cu:140-151) */
size_t ecl_bit_array_needed_gpu_memory(size_t size, uint8_t num_arrays);

/* ---- RankSelect (bitarray/rank_select.cuh, src/bitarray/rank_select.cu) - */
/* RankSelect(BitArray&&, GPU_index) (rank_select.cu:50-304): takes
 * ownership of *ba and sets it to NULL. */
ecl_status ecl_rank_select_create(ecl_bit_array* ba, int device, ecl_rank_select* out);
ecl_status ecl_rank_select_destroy(ecl_rank_select rs);
/* the public member bit_array_ (rank_select.cuh:81), borrowed */
ecl_status ecl_rank_select_bit_array(ecl_rank_select rs, ecl_bit_array* out);
ecl_status ecl_rank_select_info_get(ecl_rank_select rs, size_t array_index,
                                    ecl_rank_select_info* out);
/* getL1Entry / getL2Entry / getSelectSample over a range
 * (rank_select.cuh:569-701) */
ecl_status ecl_rank_select_get_l1(ecl_rank_select rs, size_t array_index, size_t first,
                                  size_t count, uint64_t* out);
ecl_status ecl_rank_select_get_l2(ecl_rank_select rs, size_t array_index, size_t first,
                                  size_t count, uint16_t* out);
ecl_status ecl_rank_select_get_samples(ecl_rank_select rs, size_t array_index, int value,
                                       size_t first, size_t count, uint64_t* out);
/* Device pointers of the reference-shaped directory, for the header-only
 * device API (rank_select.cuh:827-853): l1, l2, l1_offsets, l2_offsets,
 * num_l1_blocks, num_last_l2_blocks (u16), samples, sample_offsets,
 * sample_caps, total_ones, and the per-array LevelView table. */
typedef struct ecl_rank_select_device {
  uint64_t* l1; uint16_t* l2;
  uint64_t* l1_offsets; uint64_t* l2_offsets;
  uint64_t* num_l1_blocks; uint16_t* num_last_l2_blocks;
  uint64_t* samples; uint64_t* sample_offsets; uint64_t* sample_caps;
  uint64_t* total_ones;
  void* levels;
  uint64_t total_num_l2_blocks;
} ecl_rank_select_device;
ecl_status ecl_rank_select_device_pointers(ecl_rank_select rs, ecl_rank_select_device* out);
/* batched rank<1,false,value>(array, i) / select<value,1>(array, k) over
 * device arrays (rank_select.cuh:125-523 semantics) -- new batch form of the
 * reference's per-thread device calls */
ecl_status ecl_rank_select_rank_device(ecl_rank_select rs, size_t array_index, int value,
                                       const uint64_t* d_index, size_t m, uint64_t* d_out,
                                       void* stream);
ecl_status ecl_rank_select_select_device(ecl_rank_select rs, size_t array_index, int value,
                                         const uint64_t* d_k, size_t m, uint64_t* d_out,
                                         void* stream);
/* host arrays in and out */
ecl_status ecl_rank_select_rank(ecl_rank_select rs, size_t array_index, int value,
                                const uint64_t* index, size_t m, uint64_t* out);
ecl_status ecl_rank_select_select(ecl_rank_select rs, size_t array_index, int value,
                                  const uint64_t* k, size_t m, uint64_t* out);
/* new: queries drawn on the device: rank i = bounded(hash(seed, q), size),
 * select k = 1 + bounded(hash(seed, q), total of `value`) */
ecl_status ecl_rank_select_rank_generated(ecl_rank_select rs, size_t array_index, int value,
                                          uint64_t seed, size_t m, uint64_t* d_out,
                                          void* stream);
ecl_status ecl_rank_select_select_generated(ecl_rank_select rs, size_t array_index,
                                            int value, uint64_t seed, size_t m,
                                            uint64_t* d_out, void* stream);
/* RankSelect::getNeededGPUMemory (rank_select.cu:382-409) */
size_t ecl_rank_select_needed_gpu_memory(size_t size, uint8_t num_arrays);
/* Entry counts of the directory of one bit array of size_bits bits, as the
 * RankSelect constructor lays it out (rank_select.cu:57-117; the same values
 * as getNumL1Blocks / getNumL2Blocks and floor(size / 4096) sample slots):
 * the sizes of the arrays ecl_rank_select_build_directory_noalloc writes. */
ecl_status ecl_rank_select_directory_sizes(uint64_t size_bits, uint64_t* num_l1,
                                           uint64_t* num_l2, uint64_t* sample_slots);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the directory that ecl_rank_select_create builds for one
 * bit array -- L1 entries (num_l1 u64), L2 entries (num_l2 u16), select
 * samples (sample_slots u64: ones samples from the front, zeros samples from
 * the back; slots between the two are left as they are) and the total of
 * ones (one u64) -- written into the caller's device arrays on `stream`
 * (NULL = default stream), with the same kernels and the same values.
 * d_words holds the bits LSB-first in u32 words and must be readable up to
 * the next multiple of 64 bytes (a BitArray's arrays are: they are padded to
 * 128 bytes); bits at or past size_bits are ignored.  The call allocates
 * nothing: its scratch (the array's descriptor and 8 bytes per 65536 bits)
 * is a fixed 4 MiB + 256 B device array of the library, one per GPU, held
 * until the process ends; size_bits above ECL_NOALLOC_MAX_BITS is rejected
 * with ECL_ERROR_INVALID_ARGUMENT.  Every call on a GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller. */
#define ECL_NOALLOC_MAX_BITS (1ull << 35)
ecl_status ecl_rank_select_build_directory_noalloc(int device, const uint32_t* d_words,
                                                   uint64_t size_bits, uint64_t* d_l1,
                                                   uint16_t* d_l2, uint64_t* d_samples,
                                                   uint64_t* d_total_ones, void* stream);

/* ---- WaveletTree<T> (tree/wavelet_tree.cuh) ----------------------------- */
/* WaveletTree(data, data_size, alphabet&&, GPU_index) (wavelet_tree.cuh:
 * 576-867).  alphabet_size == 0 discovers the alphabet; otherwise the
 * alphabet must be sorted and hold at least 2 symbols. */
ecl_status ecl_wavelet_tree_create(ecl_type type, const void* text, size_t n,
                                   const void* alphabet, size_t alphabet_size, int device,
                                   ecl_wavelet_tree* out);
/* new: the text is already in device memory (read only, not modified) */
ecl_status ecl_wavelet_tree_create_device(ecl_type type, const void* d_text, size_t n,
                                          const void* alphabet, size_t alphabet_size,
                                          int device, ecl_wavelet_tree* out);
/* readDataFromFile<T> (utils/test_benchmark_utils.cuh:588-619): the first
 * num_symbols elements of a raw little-endian file of T into host memory,
 * with the reference's checks and messages ("Failed to open file: <path>",
 * "Data size is larger than file size", "File size is not a multiple of the
 * type size", "Failed to read file: <path>"; ECL_ERROR_INVALID_ARGUMENT). */
ecl_status ecl_read_text_file(const char* path, ecl_type type, size_t num_symbols, void* out);
/* convertDataToMinAlphabet<T> (utils/test_benchmark_utils.cuh:628-679): every
 * element of the host array becomes its index in the sorted distinct
 * elements (alphabet == NULL) or, by lower bound, in the given sorted
 * alphabet; *sigma = the alphabet size.  alphabet_out (may be NULL)
 * receives the sorted alphabet (sigma elements of T). */
ecl_status ecl_convert_to_min_alphabet(ecl_type type, void* data, size_t n, const void* alphabet,
                                       size_t alphabet_size, void* alphabet_out, size_t* sigma);
/* new: WaveletTree over the first num_symbols elements of a raw text file
 * (read as ecl_read_text_file, then as ecl_wavelet_tree_create; an empty
 * alphabet discovers it) */
ecl_status ecl_wavelet_tree_create_from_file(ecl_type type, const char* path, size_t num_symbols,
                                             const void* alphabet, size_t alphabet_size,
                                             int device, ecl_wavelet_tree* out);
/* ~WaveletTree (wavelet_tree.cuh:931-943) */
ecl_status ecl_wavelet_tree_destroy(ecl_wavelet_tree wt);
/* Query lines (no counterpart in the reference; an opt-in space-for-time
 * structure): a second copy of every level's bits in 128-byte lines of 960
 * bits, each headed by the count of ones before it, so that rank and access
 * read one memory line per level instead of three (bit block, L2 and L1
 * entry); and one 64-byte select hint per select-sample bracket (the
 * sample's position and the counts before the bracket's first 24 blocks),
 * so that a select reads its hint and one block.  Costs 128/120 + 1/64 of
 * the bits in memory and one pass over them; results do not change.  The
 * reference-shaped directory and samples stay as they are.  A second call
 * keeps the existing lines.  The call rewrites the tree's device level views:
 * make it while no query on this tree is in flight. */
ecl_status ecl_wavelet_tree_build_query_lines(ecl_wavelet_tree wt);
ecl_status ecl_rank_select_build_query_lines(ecl_rank_select rs);
/* Select lines (no counterpart in the reference; opt-in, independent of the
 * query lines): for ones and zeros of every array, one 128-byte line every r
 * of them, holding the position of that one (or zero) and the 960 bits from
 * the 64-bit word that holds it, with r = max(1, floor(800 x density)) so
 * that the next r lie in those bits.  A select then reads one memory line
 * per level.  Costs about 0.32 bytes per bit (2.7x the bits); results do not
 * change.  A second call keeps the existing lines.  Like the query lines,
 * build them while no query on this tree is in flight. */
ecl_status ecl_wavelet_tree_build_select_lines(ecl_wavelet_tree wt);
ecl_status ecl_rank_select_build_select_lines(ecl_rank_select rs);
/* Quad lines (no counterpart in the reference; opt-in, for trees over a
 * power-of-two alphabet with an even number of levels, at most 16: sigma =
 * 4, 16, 64, 256, ..., 65536): for every pair of levels (2q, 2q + 1), the
 * codes' 2-bit digits in the order of level 2q, in 128-byte lines of 448
 * digits headed by the counts of digits 0, 1 and 2 before the line.  Access
 * and rank then read one memory line per two levels (the tree's host and
 * device access/rank entry points use them, mixed batches once the quad
 * select lines exist too, and the C++ drop-in's device API for groups of
 * two or more lanes).  Costs
 * 128/112 of the bits (about one
 * byte per symbol at sigma = 256); results do not change.  Other trees get
 * ECL_ERROR_INVALID_ARGUMENT.  A second call keeps the existing lines; build
 * them while no query on the tree is in flight. */
ecl_status ecl_wavelet_tree_build_quad_lines(ecl_wavelet_tree wt);
/* Quad select lines (no counterpart in the reference; opt-in, for the trees
 * quad lines take, which this builds first when absent): for every quad
 * level and digit, a 128-byte line every r occurrences (r from the digit's
 * density, so that 448 digits hold the next r), the position of the first
 * and the 448 digits from there.  The tree's select entry points then read
 * one memory line per two levels (an occurrence past its line's digits is
 * found by a walk over the quad lines).  Costs about 1.5 bytes per symbol and quad
 * level (36 GB for 6e9 symbols at sigma = 256); results do not change.
 * Build them while no query on the tree is in flight. */
ecl_status ecl_wavelet_tree_build_quad_select_lines(ecl_wavelet_tree wt);
/* Hex lines (no counterpart in the reference; opt-in, for trees over a
 * power-of-two alphabet whose level count is a multiple of 4, at most 16:
 * sigma = 16, 256, 4096, 65536; built from the quad lines, which this
 * builds first when absent): for every four levels (4h .. 4h + 3), the
 * codes' 4-bit digits in the order of level 4h, in 128-byte lines of 128
 * digits with the counts of the 16 digits before the line.  Access and rank
 * then read one memory line per four levels (the tree's host and device
 * access/rank entry points use them; select and mixed batches keep the quad
 * layouts).  Costs about one byte per symbol and hex level; results do not
 * change.  Other trees get ECL_ERROR_INVALID_ARGUMENT.  A second call keeps
 * the existing lines; build them while no query on the tree is in flight. */
ecl_status ecl_wavelet_tree_build_hex_lines(ecl_wavelet_tree wt);
/* Device addresses of the tree's quad and hex views (ecl::detail::QuadView and
 * HexView of ecl_wavelet/detail/quad_core.hpp; no counterpart in the
 * reference).  The library keeps them current as the quad lines, quad select
 * lines and hex lines are built, so that copies of a tree passed into
 * kernels see layouts built later; the C++ drop-in's device API reads them.
 * Valid until the tree is destroyed. */
ecl_status ecl_wavelet_tree_layout_views(ecl_wavelet_tree wt, const void** quad_view,
                                         const void** hex_view);
/* Persistence (no counterpart in the reference, which rebuilds every tree
 * from its text; for warm starts of servers).  The file holds a header, the
 * alphabet, the symbol counts and every level's bits; loading checks them
 * for consistency, rebuilds the rank/select directory on the device and
 * recomputes the metadata, so a loaded tree answers exactly as the saved
 * one.  Files are little-endian and not portable across format versions. */
ecl_status ecl_wavelet_tree_save(ecl_wavelet_tree wt, const char* path);
ecl_status ecl_wavelet_tree_load(const char* path, int device, ecl_wavelet_tree* out);
ecl_status ecl_wavelet_tree_info_get(ecl_wavelet_tree wt, ecl_wavelet_tree_info* out);
/* getAlphabet (:2173): alphabet_size elements of type T */
ecl_status ecl_wavelet_tree_alphabet(ecl_wavelet_tree wt, void* out);
/* getCounts for every symbol (:2193-2201): exclusive prefix counts */
ecl_status ecl_wavelet_tree_counts(ecl_wavelet_tree wt, uint64_t* out);
/* createMinimalCodes result (:2100-2159): codes of symbols codes_start.. */
ecl_status ecl_wavelet_tree_codes(ecl_wavelet_tree wt, uint64_t* codes, uint8_t* lens);
/* getPrecomputedRank over all node ranks (:2270-2272) and the node starts
 * (getNodeInfos :2501-2532) */
ecl_status ecl_wavelet_tree_node_ranks(ecl_wavelet_tree wt, uint64_t* ranks,
                                       uint64_t* node_start, uint8_t* node_level);
/* getNumNodesAtLevel for levels 0..L-2 (:2278-2282) */
ecl_status ecl_wavelet_tree_nodes_at_level(ecl_wavelet_tree wt, uint64_t* out);
/* the public member rank_select_ (:561), borrowed */
ecl_status ecl_wavelet_tree_rank_select(ecl_wavelet_tree wt, ecl_rank_select* out);
/* device view of the whole tree (detail::TreeView in
 * include/ecl_wavelet/detail/device_core.hpp) */
ecl_status ecl_wavelet_tree_device_view(ecl_wavelet_tree wt, void* view, size_t view_bytes);
/* WaveletTree::getNeededGPUMemory (:2544-2559) */
size_t ecl_wavelet_tree_needed_gpu_memory(size_t data_size, size_t alphabet_size,
                                          uint8_t num_levels, size_t num_codes);

/* Host batch queries (access :956-1231, rank :1247-1551, select :1567-1797).
 * Inputs and outputs are host arrays owned by the caller.  access writes
 * sizeof(T)-byte symbols (original alphabet unless ECL_RESULTS_MINIMAL);
 * rank/select read ecl_rank_select_query[] (original-alphabet symbols unless
 * ECL_SYMBOLS_MINIMAL) and write u64 results. */
ecl_status ecl_wavelet_tree_access(ecl_wavelet_tree wt, const uint64_t* indices, size_t m,
                                   void* out, uint32_t flags);
ecl_status ecl_wavelet_tree_rank(ecl_wavelet_tree wt, const ecl_rank_select_query* queries,
                                 size_t m, uint64_t* out, uint32_t flags);
ecl_status ecl_wavelet_tree_select(ecl_wavelet_tree wt, const ecl_rank_select_query* queries,
                                   size_t m, uint64_t* out, uint32_t flags);
/* The same over device-resident arrays (accessKernel :373-422, rankKernel
 * :443-494, selectKernel :510-544), launched on `stream` (NULL = default) */
ecl_status ecl_wavelet_tree_access_device(ecl_wavelet_tree wt, const uint64_t* d_indices,
                                          size_t m, void* d_out, uint32_t flags, void* stream);
ecl_status ecl_wavelet_tree_rank_device(ecl_wavelet_tree wt,
                                        const ecl_rank_select_query* d_queries, size_t m,
                                        uint64_t* d_out, uint32_t flags, void* stream);
ecl_status ecl_wavelet_tree_select_device(ecl_wavelet_tree wt,
                                          const ecl_rank_select_query* d_queries, size_t m,
                                          uint64_t* d_out, uint32_t flags, void* stream);
/* new: one fused launch over a batch mixing the three kinds */
ecl_status ecl_wavelet_tree_query_batch_device(ecl_wavelet_tree wt, const ecl_query* d_queries,
                                               size_t m, uint64_t* d_out, void* stream);
/* new: uniform queries drawn on the device (kind 0: u64 indices in [0,n);
 * kind 1: ecl_rank_select_query with c uniform over the alphabet and i in
 * [0,n); kind 2: c uniform, k uniform in [1, count(c)] (1 when count is 0);
 * kind 3: ecl_query mixing the three kinds 1:1:1).  Symbols are minimal. */
ecl_status ecl_wavelet_tree_generate_queries(ecl_wavelet_tree wt, int kind, uint64_t seed,
                                             size_t m, void* d_out, void* stream);
/* new: lanes per access/rank query for this tree's kernels: 0 (default: the
 * library-wide ecl_set_query_lanes setting), 1, or 2.  No effect on results. */
ecl_status ecl_wavelet_tree_set_query_lanes(ecl_wavelet_tree wt, uint32_t lanes);
/* new: the tree's query autotuner.  Times access and rank with one and with
 * two lanes per query over m device-generated queries of each kind (seeded)
 * on the tree as it is now (with or without query lines) and keeps the
 * faster for the tree; *lanes gets 1 or 2, ms[0] and ms[1] (may be NULL)
 * the time of one access and one rank launch with each setting. */
ecl_status ecl_wavelet_tree_autotune(ecl_wavelet_tree wt, uint64_t m, uint64_t seed,
                                     uint32_t* lanes, float* ms);

/* ==========================================================================
 * Forests: many independent trees built and queried together.  No
 * counterpart in the reference (its trees are built and queried one at a
 * time, wavelet_tree.cuh:576-867, 956-2079); for callers with thousands of
 * small texts, where per-tree launches and synchronisations dominate.
 * ========================================================================== */
typedef struct ecl_forest_t* ecl_forest;
typedef struct ecl_graph_t* ecl_graph;

/* Builds one tree per text with one set of launches for all of them.
 * d_texts: device uint8 texts, text i = d_texts[starts[i] .. starts[i] +
 * lengths[i]), every start a multiple of 16 and every text non-empty; the
 * alphabet of text i is 0 .. sigmas[i]-1 (2 <= sigma <= 256).  starts,
 * lengths and sigmas are host arrays of num_texts entries.  Each tree
 * answers exactly as ecl_wavelet_tree_create_device over the same text and
 * alphabet.  tile_rows: 0 = automatic, else 1, 2 or 4 (codes per
 * construction tile = 8192 * tile_rows). */
ecl_status ecl_forest_create(const void* d_texts, const uint64_t* starts, const uint64_t* lengths,
                             const uint64_t* sigmas, size_t num_texts, int device,
                             uint32_t tile_rows, ecl_forest* out);
ecl_status ecl_forest_destroy(ecl_forest f);
/* The forest's construction as one HIP graph: replaying it (ecl_graph_launch)
 * rebuilds every tree from the current contents of the texts passed to
 * ecl_forest_create (same buffer, starts, lengths and alphabets), with no
 * host involvement between its launches: tile histograms and their scan,
 * per-tree counts, C arrays, node ranks and level sizes on the device, the
 * level pass, every directory and every level view.  A rebuild drops the
 * query lines (build them again).  The graph keeps the build's device
 * buffers (about 2.6 bytes per symbol, plus one byte per symbol when an
 * alphabet is not a power of two) until it is destroyed; destroy it before
 * the forest, and launch it while no query on the forest is in flight. */
ecl_status ecl_forest_build_graph(ecl_forest f, ecl_graph* out);
/* Times the construction with tiles of 8192, 16384 and 32768 codes on this
 * forest's texts and keeps the fastest for later build graphs (the trees
 * are the same with any tile); *tile_rows gets 1, 2 or 4 and *ms the
 * fastest build time.  Leaves the forest built. */
ecl_status ecl_forest_autotune_build(ecl_forest f, uint32_t* tile_rows, float* ms);
/* Query lines and select hints (see ecl_wavelet_tree_build_query_lines) for
 * every tree of the forest, in three launches; later forest queries and
 * graphs (also graphs captured before) read them.  Make the call while no
 * forest query or graph replay is in flight.  No counterpart in the
 * reference. */
ecl_status ecl_forest_build_query_lines(ecl_forest f);
/* Select lines (see ecl_wavelet_tree_build_select_lines) for every tree of
 * the forest, in one launch over all levels; later forest queries and graphs
 * (also graphs captured before) read them, with or without the query lines.
 * A build graph's replay drops them from the trees it rebuilds (a second
 * call then builds them again).  Make the call while no forest query or
 * graph replay is in flight.  No counterpart in the reference. */
ecl_status ecl_forest_build_select_lines(ecl_forest f);
/* Quad lines and quad select lines (see ecl_wavelet_tree_build_quad_lines
 * and _quad_select_lines) for every tree of the forest over a power-of-two
 * alphabet with an even number of levels (sigma = 4, 16, 64, 256), in three
 * launches over all their levels, and hex lines (one more launch) for those
 * whose level count is a multiple of 4 (sigma = 16, 256).  Later forest batches (ecl_forest_query_device
 * and graphs captured afterwards) then run a pair of lanes per query: over
 * these layouts for those trees, over the binary levels (with their lines,
 * when built) for the others.  A build or build-graph replay drops the
 * layouts of the trees it rebuilds (they answer over their binary levels
 * until a second call builds them again).  Graphs captured before the call
 * keep the binary kernel.  Make the call while no forest query or graph
 * replay is in flight.  No counterpart in the reference. */
ecl_status ecl_forest_build_quad_layouts(ecl_forest f);
ecl_status ecl_forest_num_trees(ecl_forest f, size_t* out);
/* size, alphabet size and levels of tree i */
ecl_status ecl_forest_tree_info(ecl_forest f, size_t i, uint64_t* n, uint64_t* sigma,
                                uint32_t* num_levels);
/* device view of tree i (detail::TreeView) */
ecl_status ecl_forest_device_view(ecl_forest f, size_t i, void* view, size_t view_bytes);
/* One launch over a batch of queries on any trees of the forest (symbols
 * minimal).  Results as ecl_wavelet_tree_query_batch_device. */
ecl_status ecl_forest_query_device(ecl_forest f, const ecl_forest_query* d_queries, size_t m,
                                   uint64_t* d_out, void* stream);
/* Mixed queries drawn on the device: queries [qoff[i], qoff[i+1]) go to
 * tree i (qoff: host, num_trees + 1 entries), kind q % 3, fields drawn as
 * by ecl_wavelet_tree_generate_queries. */
ecl_status ecl_forest_generate_queries(ecl_forest f, const uint64_t* qoff, uint64_t seed,
                                       ecl_forest_query* d_out, void* stream);
/* Times the batch query launch over workgroup sizes and grid sizes on (a
 * prefix of) the given batch and keeps the fastest for this forest.  Once
 * the quad layouts exist it tunes the pair-lane kernel the batches then run
 * (256-thread workgroups: grid sizes only). */
ecl_status ecl_forest_autotune(ecl_forest f, const ecl_forest_query* d_queries, size_t m,
                               uint64_t* d_out, uint32_t* block, uint32_t* grid);
/* The batch launch captured once into a HIP graph and replayed: the
 * queries and outputs must stay allocated while the graph lives. */
ecl_status ecl_forest_graph_create(ecl_forest f, const ecl_forest_query* d_queries, size_t m,
                                   uint64_t* d_out, ecl_graph* out);
ecl_status ecl_graph_launch(ecl_graph g, void* stream);
ecl_status ecl_graph_destroy(ecl_graph g);

/* ==========================================================================
 * Query server: trees stay resident on the device, admitted against an HBM
 * budget; requests (mixed ecl_query batches against one tree, in host
 * memory) are coalesced into device batches while earlier batches are still
 * running (continuous batching).  No counterpart in the reference, whose
 * host queries run one call at a time (wavelet_tree.cuh:956-1797); for
 * callers that serve many concurrent requests.
 * ========================================================================== */
typedef struct ecl_server_t* ecl_server;
typedef struct ecl_server_config {
  uint64_t hbm_budget;    /* bytes the resident trees may occupy (0: 90% of free HBM) */
  uint64_t max_batch;     /* queries per device batch (0: 2^22) */
  uint32_t max_inflight;  /* device batches in flight at once (0: 3) */
  uint32_t max_trees;     /* resident trees (0: 1024) */
  uint32_t coalesce;      /* 1: coalesce requests into batches; 0: one request per
                           * batch and one batch on the device at a time (the plain
                           * per-call form) */
  uint32_t query_lines;   /* 1: every admitted tree also gets its query lines and
                           * select hints (ecl_wavelet_tree_build_query_lines; about
                           * 1.08x its bits more memory); 2: and its select lines
                           * (ecl_wavelet_tree_build_select_lines; about 2.7x its
                           * bits more); 3: the quad layouts for trees over a
                           * power-of-two alphabet with an even number of levels
                           * (ecl_wavelet_tree_build_quad_select_lines, and
                           * ecl_wavelet_tree_build_hex_lines where the level count
                           * is a multiple of 4; about 9 bytes per symbol at sigma
                           * = 256) and query and select lines for the other trees;
                           * batches then run a pair of lanes per query.  Charged
                           * to the budget. */
} ecl_server_config;
typedef struct ecl_server_stats {
  uint64_t requests;      /* completed requests since the last reset */
  uint64_t queries;
  uint64_t batches;
  double p50_ms, p99_ms, mean_ms, max_ms;  /* submit -> results in host memory */
  double elapsed_s;       /* first submit to last completion */
  double queries_per_s;
  uint64_t hbm_used;      /* bytes held by the resident trees */
  uint64_t trees;
} ecl_server_stats;
ecl_status ecl_server_create(int device, const ecl_server_config* config, ecl_server* out);
/* completes every submitted request, then frees the trees */
ecl_status ecl_server_destroy(ecl_server s);
/* Builds and admits a tree over a uint8 text (host or device) with the
 * alphabet 0 .. sigma-1; ECL_ERROR_OUT_OF_MEMORY_TREE when the tree would
 * exceed the budget. */
ecl_status ecl_server_add_tree(ecl_server s, const void* text, int text_on_device, size_t n,
                               uint64_t sigma, uint32_t* tree);
/* Admits a tree read from a file written by ecl_wavelet_tree_save (a uint8
 * tree over 0 .. sigma-1), without rebuilding it: the warm start of a
 * server.  Same budget rule and error as ecl_server_add_tree. */
ecl_status ecl_server_add_tree_file(ecl_server s, const char* path, uint32_t* tree);
/* Admits a tree over the first num_symbols bytes of a raw uint8 text file,
 * converted to its minimal alphabet first (ecl_read_text_file, then
 * ecl_convert_to_min_alphabet): requests name symbols by their index in
 * that alphabet.  alphabet_out (256 bytes, may be NULL) receives the
 * alphabet, *sigma its size. */
ecl_status ecl_server_add_tree_text_file(ecl_server s, const char* path, size_t num_symbols,
                                         uint8_t* alphabet_out, uint64_t* sigma,
                                         uint32_t* tree);
/* No new requests for the tree; its memory is freed once its requests finish. */
ecl_status ecl_server_remove_tree(ecl_server s, uint32_t tree);
/* Queues a request; `queries` must stay valid and `results` (m entries)
 * writable until ecl_server_wait returns for it. */
ecl_status ecl_server_submit(ecl_server s, uint32_t tree, const ecl_query* queries, size_t m,
                             uint64_t* results, uint64_t* request);
/* Blocks until the request's results are in host memory. */
ecl_status ecl_server_wait(ecl_server s, uint64_t request, double* latency_ms);
/* Switches between coalescing (1) and the plain per-call form (0: one
 * request per batch, one batch on the device at a time) for later batches. */
ecl_status ecl_server_set_mode(ecl_server s, uint32_t coalesce);
ecl_status ecl_server_stats_get(ecl_server s, ecl_server_stats* out);
ecl_status ecl_server_stats_reset(ecl_server s);
/* Open-loop synthetic load: Poisson arrivals at `rate` requests/s for
 * `seconds`; request sizes log-uniform in [min_size, max_size] queries
 * against a uniformly chosen resident tree, taken from per-tree pools of
 * `pool` device-generated mixed queries.  Returns the stats of the run. */
typedef struct ecl_load_spec {
  double rate;
  double seconds;
  uint64_t seed;
  uint64_t min_size, max_size;
  uint64_t pool;
} ecl_load_spec;
ecl_status ecl_server_load(ecl_server s, const ecl_load_spec* spec, ecl_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* ECL_WAVELET_C_H */
