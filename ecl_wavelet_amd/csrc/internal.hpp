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
#include "ecl_wavelet/detail/quad_core.hpp"
#include "ecl_wavelet_c.h"

namespace ecl {
namespace lib {

// the quad and hex layouts (ecl_wavelet/detail/quad_core.hpp)
using detail::HexView;
using detail::QuadView;
using detail::hex_match;
using detail::kHexDigits;
using detail::kHexSbShift;
using detail::kMaxHexLevels;
using detail::kMaxQuadLevels;
using detail::kQuadDigits;
using detail::quad_match;

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
// Device memory owned by the library.  Blocks of kPoolMinBytes or more come
// from and go back to a per-device block cache (see capi.hip): cap is the
// block's size, dev the device it lives on.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  size_t cap = 0;
  int dev = -1;
  DevBuf() = default;
  explicit DevBuf(size_t b);
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes), cap(o.cap), dev(o.dev) {
    o.ptr = nullptr; o.bytes = 0; o.cap = 0; o.dev = -1;
  }
  DevBuf& operator=(DevBuf&& o) noexcept;
  void reset();
  template <typename T> T* as() const { return static_cast<T*>(ptr); }
};
constexpr size_t kPoolMinBytes = size_t(1) << 20;
// bytes of freed blocks the device's block cache holds (available to the
// library's next allocations; freed by ecl_release_cached_memory)
size_t pooled_bytes(int device);
// frees every cached block of the device
void release_block_pool(int device);

// ---- kernel profiling (events around selected launches) -------------------
enum KernelId : int {
  K_HIST = 0, K_PASS, K_RS_BUILD, K_WT_ACCESS, K_WT_RANK, K_WT_SELECT,
  K_WT_MIXED, K_RS_RANK, K_RS_SELECT, K_GEN, K_FILL, K_TILE_HIST, K_TILE_SCAN,
  K_FOREST_QUERY, K_COUNT
};
const char* kernel_name(int id);
struct ProfScope {
  int id; hipStream_t stream; hipEvent_t e0 = nullptr, e1 = nullptr;
  ProfScope(int id, hipStream_t s);
  ~ProfScope();
};

// ---- small host->device copies carried in kernel arguments: the launch
// copies the bytes, so the host neither stages them nor waits for earlier
// work of the stream (kernels_build.hip) ---------------------------------------
constexpr uint32_t kUploadBytes = 2048;
constexpr uint32_t kUploadSegs = 12;
struct UploadBlob {
  uint64_t dst[kUploadSegs];       // destination of segment i
  uint32_t off[kUploadSegs + 1];   // segment i = data[off[i] .. off[i+1])
  uint32_t nseg;
  uint8_t data[kUploadBytes];
};
void launch_upload(const UploadBlob& b, hipStream_t s);
// query lines of one array (detail::LevelView::lines): nlines =
// ceil(size / kLineBits) lines of 128 bytes at `lines`
void launch_query_lines(const LevelView& v, uint4* lines, uint64_t nlines, hipStream_t s);
// select lines of one array: n1 for the ones at rate r1, then n0 for the
// zeros at rate r0 (layout at LevelView::sel_lines)
void launch_select_lines(const LevelView& v, uint4* out, uint64_t n1, uint64_t n0, uint32_t r1,
                         uint32_t r0, hipStream_t s);
// The select-line rate of a Value with `count` of `size` bits: r = max(1,
// floor(800 x density)), so that the next r Values lie within a line's 960
// bits (0 when the Value is absent); and the multiplier that turns i / r into
// a multiply-high (LevelView::sel_magic).
inline uint32_t select_line_rate(uint64_t count, uint64_t size) {
  if (count == 0) return 0;
  const uint64_t r = 800 * count / size;
  return static_cast<uint32_t>(r < 1 ? 1 : r);
}
__host__ __device__ inline uint64_t select_line_magic(uint32_t r) { return r > 1 ? ~0ull / r + 1 : 0; }
// Select lines of many arrays (the levels of a forest) in one launch: job a
// takes the workgroups [wg_start[a], wg_start[a] + 2 nchunks) (its ones'
// chunks, then its zeros'; a chunk is kSelLineChunk 512-bit blocks), and
// writes its n1 + n0 lines from line line_off of `out`.
constexpr uint32_t kSelLineChunk = 256;
struct SelLinesJob {
  LevelView v;
  uint64_t line_off, nl2, nchunks, n1;
  uint32_t r1, r0;
};
void launch_multi_select_lines(const SelLinesJob* jobs, const uint64_t* wg_start, uint32_t njobs,
                               uint64_t total_wgs, uint4* out, hipStream_t s);
// Query lines and select hints of many arrays in three launches: job a's
// lines and hints start at line_start[a] / hint_start[a] of the shared
// buffers, and its device view d_view gets their addresses and hint_values
// (fit: 2 x njobs u32, zeroed by the caller).
struct LinesJob {
  LevelView v;        // the array's view without lines or hints
  LevelView* d_view;  // its view in device memory
  uint64_t nl, nh1, nh0;
};
void launch_multi_lines(const LinesJob* jobs, const uint64_t* line_start,
                        const uint64_t* hint_start, uint32_t njobs, uint4* lines,
                        uint64_t total_lines, uint4* hints, uint64_t total_hints, uint32_t* fit,
                        hipStream_t s);
// the hints of one array and its hint_values, set in its device view d_view
// (block_fit: 2 x select_hint_blocks(nh1, nh0) u32 of scratch)
uint64_t select_hint_blocks(uint64_t nh1, uint64_t nh0);
void launch_select_hints(const LevelView& v, uint4* hints, uint64_t nh1, uint64_t nh0,
                         uint32_t* block_fit, LevelView* d_view, hipStream_t s);

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

// One construction pass over levels l0 .. l0+k-1 (kernels_tree.hip).
struct PassParams {
  // tile_hist input: the text (first pass) or the level-l0 codes
  const void* src; int src_esize;
  void* codes_out;            // first pass: mapped codes (width cw), or nullptr
  int map_kind;               // 0 identity, 1 LUT, 2 binary search
  const uint32_t* lut; const void* alphabet; uint64_t sigma;
  uint64_t codes_start; int pow2; const uint64_t* code_table;
  // pass input: level-l0 codes of width cw (src, or codes_out after mapping)
  const void* codes; int cw;
  uint32_t k, dshift;         // digit = (code >> dshift) & (2^k - 1)
  uint64_t n;                 // uniform tiles over [0, n) when tile_start == nullptr
  uint32_t tile_rows;         // 16-byte chunks per thread: 1, 2 or 4
  uint32_t tile_codes;        // codes per tile = 512 * tile_rows * 16 / cw
  uint64_t num_tiles;
  const uint64_t* tile_start; const uint32_t* tile_len; const uint32_t* tile_node;
  const uint32_t* node_first_tile;   // [nodes + 1], nullptr for uniform tiles
  uint32_t* hist;             // [tiles][2^k] digit counts per tile
  uint64_t* excl;             // [tiles + 1][2^k] exclusive column scan
  uint64_t* partial;          // scan scratch, scan_partial_entries() u64
  const uint64_t* pc;         // [nodes][2^k + 1] global start of (node, digit)
  const uint64_t* obase;      // [nodes][2^k] padded start of (node, digit) in dst
  const uint16_t* thr;        // [nodes][k + 1] first absent digit per level
  void* dst;                  // level-(l0+k) codes, or nullptr on the last pass
  uint64_t dst_cap;           // codes dst can hold
  uint32_t* words;            // bit array words
  uint64_t level_word[8];     // word offset of levels l0 .. l0+k-1
  uint64_t level_nwords[8];   // words of each of those levels
  uint32_t* err;              // set to 1 if a write fell outside its array
  // Batched construction of independent trees (one node per text, single
  // pass): per-node levels, level words and symbol->code tables.  nullptr
  // entries mean "the same for every node" (the fields above).
  const uint8_t* node_k;      // [nodes] levels of the node's tree
  const uint64_t* node_lw;    // [nodes][8] word offsets of its levels
  const uint64_t* node_lnw;   // [nodes][8] words of its levels
  const uint8_t* node_lut;    // [nodes][256] symbol -> code, for remapped nodes
  const uint8_t* node_remap;  // [nodes] 1: codes come from codes_out
  uint32_t row_bins;          // entries per hist/excl row (0: 2^k)
  uint32_t thr_stride;        // entries per node in thr (0: k + 1)
  // construction: consecutive tiles per workgroup (0: chosen by launch_pass)
  uint32_t tiles_per_wg;
};
uint64_t pass_tile_codes(int cw, uint32_t rows);    // codes per tile
uint64_t scan_partial_entries(uint64_t tiles, uint32_t k);
void launch_tile_hist(const PassParams& p, hipStream_t s);
void launch_tile_scan(const PassParams& p, hipStream_t s);
void launch_pass(const PassParams& p, hipStream_t s);
void set_pass_tiles_per_wg(uint32_t span);  // 0: automatic
void set_query_lanes(uint32_t lanes);      // lanes per access/rank query: 0 auto, 1 or 2
uint32_t pass_span(uint64_t num_tiles);     // tiles per workgroup of a pass
// zeroes the words that several workgroups of a single-node pass write
// (p.tiles_per_wg set); the pass then needs no zeroed bit arrays
void launch_span_words(const PassParams& p, hipStream_t s);
void launch_gather_rows(const uint64_t* excl, const uint32_t* rows, uint64_t nrows,
                        uint32_t nb, uint64_t* out, hipStream_t s);
// pc[d] = sum of totals[0 .. d) for d <= nb (the start of every digit's codes
// in a single-node pass), computed on the device from the scan's totals row
void launch_digit_starts(const uint64_t* totals, uint32_t nb, uint64_t* pc, hipStream_t s);

// One bit array of a directory build: its words, its directory arrays and
// the global index of its first L1 block.
struct RsArrayDesc {
  const uint64_t* words64; uint64_t m;
  uint64_t* l1; uint16_t* l2; uint64_t num_l1, num_l2;
  uint64_t* samples; uint64_t cap;
  uint64_t* total_ones;
  uint64_t first_tile;  // first global L1 block of the array
  // L1 blocks reserved for the array in the global numbering (>= num_l1;
  // 0 means num_l1): blocks past num_l1 do nothing, so a launch laid out by
  // capacity serves arrays whose sizes are only known on the device
  uint64_t cap_l1;
  // first chunk of kRsScanChunk L1 blocks of the array in the global
  // numbering of the L1 scan (the arrays' chunks follow each other)
  uint64_t first_chunk;
};
constexpr uint64_t kRsScanChunk = 2048;  // L1 blocks per chunk of the L1 scan
__host__ __device__ inline uint64_t rs_scan_chunks(uint64_t l1_blocks) {
  return (l1_blocks + kRsScanChunk - 1) / kRsScanChunk;
}
// arrays: narrays entries (device) covering nblocks global L1 blocks and
// nchunks scan chunks in order; scratch: 2 * nblocks u32 (the array of every
// block, block totals)
void launch_rs_build(const RsArrayDesc* arrays, uint64_t narrays, uint64_t nblocks,
                     uint64_t nchunks, uint32_t* scratch, hipStream_t s);

void launch_gen_text(void* out, int esize, uint64_t n, uint64_t seed,
                     uint64_t sigma, const uint64_t* thresholds,
                     hipStream_t s);
void launch_gen_bits(uint32_t* words, uint64_t nbits, uint64_t nwords_total,
                     uint64_t seed, uint64_t threshold, hipStream_t s);
void launch_presence(const void* text, int esize, uint64_t n, uint32_t* flags,
                     hipStream_t s);

// ---- launchers (kernels_query.hip) -------------------------------------------
constexpr uint32_t kQuadSelSpan = 340;  // digits the r digits of a select line span on average

// One quad level's select lines of every digit, from its quad lines; digit
// d with r[d] == 0 (absent) has none.
struct QuadSelBuildParams {
  const uint4* qlines;  // the level's quad lines
  uint64_t nql;         // their number (n / 448 + 1)
  uint64_t total[4];    // the level's d's
  uint32_t r[4];
  uint64_t magic[4];
  uint64_t nlines[4];   // total / r + 1
  uint4* out[4];
};
void launch_quad_select_lines(const QuadSelBuildParams& qp, hipStream_t s);
// build workgroups of the quad select lines of a level of nql quad lines
uint64_t quad_select_build_workgroups(uint64_t nql);

// One quad level's build: binary levels a = 2q and b = 2q + 1, the level-2q
// node starts S[0..nodes] (S[nodes] = n) and per node and digit the digits
// before the node, CD[4N + d] for N in [0, nodes] (row nodes: the totals).
struct QuadBuildParams {
  LevelView a, b;
  const uint64_t* S;
  const uint64_t* CD;
  uint64_t nodes;
  uint64_t n;
  uint64_t nlines;  // n / 448 + 1: position n has a line (its header: the totals)
  uint4* out;
  uint64_t* wg_info;  // scratch: kQuadWgInfo u64 per build workgroup
  // 4096 bytes: entry a | z << 4 | o << 8 holds the 4 digits of 4
  // positions whose first bits are a (bit i: position i), whose zeros take
  // their second bits in order from z and whose ones from o
  const uint8_t* lut;
};
constexpr uint32_t kQuadWgInfo = 6;
uint64_t quad_build_workgroups(uint64_t nlines);
void launch_quad_lines(const QuadBuildParams& qp, hipStream_t s);
// The quad lines and quad select lines of many quad levels (a forest's) in
// three launches: job a's build workgroups are [wg_start[a], wg_start[a] +
// quad_build_workgroups(nlines)) of the quad-line launches (its wg_info at
// wg_start[a] * kQuadWgInfo of one scratch array) and [sel_start[a], ... +
// quad_select_build_workgroups(nql)) of the select-line launch.
void launch_multi_quad_layouts(const QuadBuildParams* jobs, const QuadSelBuildParams* sel_jobs,
                               const uint64_t* wg_start, const uint64_t* sel_start,
                               uint32_t njobs, uint64_t total_wg, uint64_t total_sel,
                               hipStream_t s);

// One hex level's build from the quad lines of quad levels 2h (qa) and
// 2h + 1 (qb).  Position p of node N with quad digit d0 goes to position
// e = qbase[4N + d0] + (d0's before p) of quad level 2h + 1, whose digit is
// the hex digit's low half.  The x's before position P (x = 4 d0 + d1) are
// CHD[N][x] + (d1's before e_d0(P) at level 2h + 1) - qcd[4 (4N + d0) + d1].
struct HexBuildParams {
  const uint4* qa;
  const uint4* qb;
  const uint64_t* qbase;  // quad bases of level 2h: qbase[4N + d0]
  const uint64_t* qcd;    // quad level 2h + 1's CD table: qcd[4M + d1]
  const uint64_t* chd;    // CHD[16N + x]: the x's before node N (nodes + 1 rows)
  const uint64_t* S;      // starts of the nodes of hex level h (nodes + 1)
  uint64_t nodes;         // 16^h
  uint64_t n;
  uint64_t nql;           // quad lines per quad level: n / 448 + 1
  uint64_t nlines;        // hex lines: n / 128 + 1
  uint4* out;
  uint64_t* sb;           // superblock counts: sb[16 s + x]
};
void launch_hex_lines(const HexBuildParams& hp, hipStream_t s);
// The hex lines of many hex levels (a forest's) in one launch: job a's build
// workgroups are [wg_start[a], wg_start[a] + hex_build_workgroups(nlines)).
uint64_t hex_build_workgroups(uint64_t nlines);
void launch_multi_hex_lines(const HexBuildParams* jobs, const uint64_t* wg_start, uint32_t njobs,
                            uint64_t total_wg, hipStream_t s);

struct TreeQueryParams {
  TreeView view;
  QuadView quad;             // quad.levels == 0: none
  HexView hex;               // hex.levels == 0: none
  int esize;
  int symbols_minimal;       // 1: query symbols already minimal indices
  int map_kind; const uint32_t* lut; const void* alphabet;  // symbol map
  const void* alphabet_out;  // access: map result back (nullable = identity)
  uint64_t max_lds_counts;   // unused by callers; tuning knob
  int has_lines;             // 1: the tree's levels carry query lines
  int has_sel_lines;         // 1: the tree's levels carry select lines
  uint32_t lanes;            // lanes per access/rank query: 0 (the library-wide
                             // setting), 1 or 2 (the tree's own, from its tuner)
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
void launch_forest_query(const TreeView* views, uint64_t num_trees, const ecl_forest_query* q,
                         uint64_t m, uint64_t* out, uint32_t block, uint32_t grid,
                         hipStream_t s);
void launch_segment_query(const TreeView* views, uint64_t num_views, const ecl_query* q,
                          uint64_t m, const uint64_t* seg_start, const uint32_t* seg_tree,
                          uint32_t nseg, uint64_t* out, uint32_t block, uint32_t grid,
                          hipStream_t s);
// A tree's query parameters for minimal symbols and results (the layouts it
// carries included); valid while the tree lives.
TreeQueryParams tree_query_params(ecl_wavelet_tree wt);
// A served batch over trees with quad select lines (TreeQueryParams per
// tree slot, beside the views): a pair of lanes per query, 256 threads per
// workgroup.
void launch_segment_query_quad(const TreeView* views, const TreeQueryParams* params,
                               uint64_t num_views, const ecl_query* q, uint64_t m,
                               const uint64_t* seg_start, const uint32_t* seg_tree, uint32_t nseg,
                               uint64_t* out, uint32_t grid, hipStream_t s);
// The same pair-lane batch over a forest whose trees carry TreeQueryParams
// (query q names its tree; trees without quad select lines run the binary
// descents).
void launch_forest_query_quad(const TreeView* views, const TreeQueryParams* params,
                              uint64_t num_trees, const ecl_forest_query* q, uint64_t m,
                              uint64_t* out, uint32_t grid, hipStream_t s);
void launch_forest_gen(const TreeView* views, uint64_t num_trees, const uint64_t* qoff,
                       uint64_t seed, uint64_t m, ecl_forest_query* out, hipStream_t s);
void launch_rs_rank(const LevelView* lv, int value, const uint64_t* idx,
                    uint64_t m, uint64_t* out, int generated, uint64_t seed,
                    int has_lines, hipStream_t s);
void launch_rs_select(const LevelView* lv, int value, const uint64_t* k,
                      uint64_t m, uint64_t* out, int generated, uint64_t seed,
                      hipStream_t s);

}  // namespace lib
}  // namespace ecl
