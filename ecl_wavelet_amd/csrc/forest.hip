// forest.hip -- many independent wavelet trees built and queried together.
//
// Construction runs the same three kernels as one tree (tile digit
// histograms, their scan, the level pass of kernels_tree.hip), once for
// all texts: every text is one node with its own levels, tables and bit
// arrays, so thousands of trees cost a few launches.  The rank/select
// directories of every level of every tree are built by one launch.
// Queries of any mix of kinds and trees go through one launch whose shape
// is tuned per forest; the launch can be captured into a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "host.hpp"

using namespace ecl::lib;
namespace detail = ecl::detail;
using detail::LevelView;
using detail::TreeView;

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint32_t kRow = 256;  // digits per text: 8-bit codes, one pass
uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// One tree of the forest: where its arrays live in the forest's arena (byte
// offsets).  Every level gets the capacity of a full level (n bits and the
// directory of n bits): the level sizes of a non-power-of-two alphabet
// depend on the symbol counts, which only the device knows during the build,
// and a layout that does not depend on them lets the whole build run
// without the host waiting in between.
struct TreeSlot {
  uint64_t n = 0, sigma = 0;
  uint32_t L = 0;
  const TreeShape* shape = nullptr;
  uint32_t shape_id = 0, num_ranks = 0;
  std::vector<uint64_t> level_word;                 // word offset of each level
  std::vector<uint64_t> l1_at, l2_at, smp_at;      // directory byte offsets
  uint64_t ones_at = 0, lv_at = 0, counts_at = 0, ranks_at = 0, nal_at = 0, codes_at = 0,
           lens_at = 0;
};

int cu_count(int device) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus;
}

// ---- device-side metadata ---------------------------------------------------
// One distinct alphabet size: its minimal codes (as 8-bit digits) and code
// lengths from code_off, its internal nodes in (level, symbol) order from
// node_off (lo, hi and level each).
struct ShapeDev {
  uint32_t sigma, L, nodes, code_off, node_off;
};
struct MetaJob {
  uint64_t n;
  uint64_t counts_at, ranks_at;  // arena byte offsets of the C array and node ranks
  uint32_t shape, desc0;         // shape index; first directory descriptor (L of them)
};

__device__ __forceinline__ uint64_t rs_num_l2(uint64_t m, uint64_t num_l1) {
  uint64_t last = ((m % detail::kL1Bits) + detail::kL2Bits - 1) / detail::kL2Bits;
  if (last == 0 && m) last = detail::kL2PerL1;
  return m ? (num_l1 - 1) * detail::kL2PerL1 + last : 0;
}

// One thread per tree: from the digit-count rows at the text boundaries,
// the tree's digit starts (pc, for the pass), its C array and node ranks
// (wavelet_tree.cuh:773-774, 332-352: the zeros of level l before each node
// start, in (level, symbol) order) and the sizes of its levels
// (wavelet_tree.cuh:776-790) in its directory descriptors.
__global__ void forest_meta_kernel(const uint64_t* __restrict__ rows, const MetaJob* __restrict__ jobs,
                                   uint32_t N, const ShapeDev* __restrict__ shapes,
                                   const uint8_t* __restrict__ sh_code,
                                   const uint8_t* __restrict__ sh_len,
                                   const uint16_t* __restrict__ sh_node,
                                   const uint8_t* __restrict__ sh_node_level,
                                   uint8_t* __restrict__ arena, uint64_t* __restrict__ pc,
                                   RsArrayDesc* __restrict__ desc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const MetaJob J = jobs[i];
  const ShapeDev S = shapes[J.shape];
  const uint64_t* r0 = rows + static_cast<uint64_t>(i) * kRow;
  const uint64_t* r1 = r0 + kRow;
  uint64_t* P = pc + static_cast<uint64_t>(i) * (kRow + 1);
  uint64_t run = 0;
  for (uint32_t d = 0; d < kRow; ++d) {
    P[d] = run;
    run += r1[d] - r0[d];
  }
  P[kRow] = run;
  uint64_t* C = reinterpret_cast<uint64_t*>(arena + J.counts_at);
  uint64_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t acc = 0;
  for (uint32_t c = 0; c < S.sigma; ++c) {
    const uint8_t code = sh_code[S.code_off + c];
    const uint64_t cnt = r1[code] - r0[code];
    C[c] = acc;
    acc += cnt;
    const uint32_t len = sh_len[S.code_off + c];
    for (uint32_t l = 0; l < 8; ++l)
      if (l < len) m[l] += cnt;
  }
  uint64_t* R = reinterpret_cast<uint64_t*>(arena + J.ranks_at);
  uint32_t k = 0, cur = 0xFFFFFFFFu;
  uint64_t zb = 0;
  for (uint32_t j = 0; j < S.nodes; ++j) {
    const uint32_t lo = sh_node[2 * (S.node_off + j)], hi = sh_node[2 * (S.node_off + j) + 1];
    const uint32_t l = sh_node_level[S.node_off + j];
    if (l != cur) {
      cur = l;
      zb = 0;
    }
    const uint32_t mid = lo + static_cast<uint32_t>(detail::prev_pow2(hi - lo + 1));
    if (l >= 1 && lo >= 1) R[k++] = zb;
    const uint64_t cmid = mid < S.sigma ? C[mid] : J.n;
    zb += cmid - C[lo];
  }
  for (uint32_t l = 0; l < S.L; ++l) {
    RsArrayDesc& d = desc[J.desc0 + l];
    d.m = m[l];
    d.num_l1 = (m[l] + detail::kL1Bits - 1) / detail::kL1Bits;
    d.num_l2 = rs_num_l2(m[l], d.num_l1);
    d.cap = m[l] / detail::kSampleRate;
  }
}

// One thread per level of every tree: its view, from its directory
// descriptor and the total the directory build wrote.  Query lines are not
// part of a build: a rebuilt forest answers over its directory until they
// are built again.
__global__ void forest_views_kernel(const RsArrayDesc* __restrict__ desc, uint32_t ndesc,
                                    LevelView* const* __restrict__ lv_ptr) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ndesc) return;
  const RsArrayDesc d = desc[j];
  LevelView v{};
  v.words = reinterpret_cast<const uint32_t*>(d.words64);
  v.l1 = d.l1;
  v.l2 = d.l2;
  v.samples = d.samples;
  v.size = d.m;
  v.ones = *d.total_ones;
  v.num_l1 = d.num_l1;
  v.sample_cap = d.cap;
  *lv_ptr[j] = v;
}

// One thread per tree: its quad view is cleared (a rebuilt tree's quad
// layouts no longer match its bits; it answers over its binary levels until
// they are built again).
__global__ void forest_clear_quad_kernel(TreeQueryParams* __restrict__ params, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  params[t].quad.levels = 0;
  params[t].quad.has_sel = 0;
  params[t].hex.levels = 0;
}

}  // namespace

// The device buffers and launch parameters of one forest build: everything
// the launch sequence reads besides the texts and the arena.  Kept by the
// forest while a build graph (ecl_forest_build_graph) may replay it.
struct ForestBuild {
  uint32_t rows = 1;
  uint64_t ntiles = 0, blocks = 0, chunks = 0, meta_at = 0;
  uint32_t N = 0, ndesc = 0;
  DevBuf d_x, d_tstart, d_tlen, d_tnode, d_tfirst, d_remap, d_lut, d_hist, d_excl, d_partial,
      d_err, d_rows, d_pc, d_thr, d_nk, d_lw, d_lnw, d_jobs, d_shapes, d_code, d_len, d_node,
      d_node_level, d_desc, d_blk, d_lvptr;
  PassParams pp{};
};

struct ecl_graph_t {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::shared_ptr<ForestBuild> keep;  // a build graph's buffers
  ~ecl_graph_t() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct ecl_forest_t {
  int device = 0;
  const void* d_texts = nullptr;
  std::vector<uint64_t> starts, lengths, sigmas;
  std::map<uint64_t, TreeShape> shapes;
  std::vector<TreeSlot> trees;
  DevBuf arena;                  // bits, directories and metadata of every tree
  DevBuf d_views;                // TreeView per tree
  std::vector<TreeView> views;
  uint32_t q_block = 256, q_grid = 0;  // launch shape of the batch query kernel
  uint32_t tile_rows = 0;              // construction tile (0: by padding)
  std::vector<LevelView*> d_levels;    // where each level's view lives in the arena
  DevBuf lines;                        // query lines and select hints (opt-in)
  DevBuf sel_lines;                    // select lines (opt-in)
  // Quad layouts (opt-in): every tree's query parameters (its view, and its
  // quad view once built; a build clears the quad views on the device, so a
  // rebuilt tree answers over its binary levels), and the layouts' arrays.
  DevBuf d_params;
  std::vector<TreeQueryParams> params;
  DevBuf quad_lines, quad_sel, quad_base, quad_tab, hex_lines, hex_tab;
  bool quad_built = false;   // batches run the pair-lane kernel
  uint32_t q_grid_quad = 0;  // its grid (256-thread workgroups)
  std::shared_ptr<ForestBuild> build;  // buffers of a build graph, while one lives
};

namespace {

// Tiles of `rows` 16-byte chunks per thread (0: the largest whose padding
// stays within 1/4 of the text).
uint32_t choose_rows(const ecl_forest_t* f, uint32_t rows) {
  if (rows != 0) return rows;
  uint64_t total = 0;
  for (uint64_t n : f->lengths) total += n;
  for (uint32_t r : {4u, 2u}) {
    const uint64_t T = pass_tile_codes(1, r);
    uint64_t padded = 0;
    for (uint64_t n : f->lengths) padded += (n + T - 1) / T * T;
    if (padded * 4 <= total * 5) return r;
  }
  return 1;
}

// The device buffers and tables of a build of forest f with construction
// tiles of `rows` chunks per thread (the arena must exist).
std::shared_ptr<ForestBuild> make_build(ecl_forest_t* f, uint32_t rows) {
  auto B = std::make_shared<ForestBuild>();
  const size_t N = f->trees.size();
  B->N = static_cast<uint32_t>(N);
  B->rows = rows;
  const uint64_t T = pass_tile_codes(1, rows);
  std::vector<uint64_t> tstart;
  std::vector<uint32_t> tlen, tnode, tfirst(N + 1);
  uint64_t span = 0;
  for (size_t i = 0; i < N; ++i) {
    tfirst[i] = static_cast<uint32_t>(tstart.size());
    const uint64_t n = f->lengths[i];
    span = std::max(span, f->starts[i] + n);
    for (uint64_t o = 0; o < n; o += T) {
      tstart.push_back(f->starts[i] + o);
      tlen.push_back(static_cast<uint32_t>(std::min<uint64_t>(T, n - o)));
      tnode.push_back(static_cast<uint32_t>(i));
    }
  }
  B->ntiles = tstart.size();
  tfirst[N] = static_cast<uint32_t>(B->ntiles);
  require(B->ntiles < (1ull << 31), "too many construction tiles");
  // non-power-of-2 alphabets read mapped codes (same positions)
  bool any_remap = false;
  std::vector<uint8_t> remap(N, 0), lut(N * 256, 0);
  for (size_t i = 0; i < N; ++i) {
    const TreeShape& sh = *f->trees[i].shape;
    if (sh.pow2) continue;
    remap[i] = 1;
    any_remap = true;
    for (uint32_t x = 0; x < 256; ++x)
      lut[i * 256 + x] = static_cast<uint8_t>(sh.code[std::min<uint64_t>(x, sh.sigma - 1)]);
  }
  B->d_x = DevBuf(any_remap ? span : 0);
  B->d_tstart = DevBuf(B->ntiles * 8);
  B->d_tlen = DevBuf(B->ntiles * 4);
  B->d_tnode = DevBuf(B->ntiles * 4);
  B->d_tfirst = DevBuf((N + 1) * 4);
  B->d_remap = DevBuf(N);
  B->d_lut = DevBuf(N * 256);
  h2d(B->d_tstart.ptr, tstart.data(), B->ntiles);
  h2d(B->d_tlen.ptr, tlen.data(), B->ntiles);
  h2d(B->d_tnode.ptr, tnode.data(), B->ntiles);
  h2d(B->d_tfirst.ptr, tfirst.data(), N + 1);
  h2d(B->d_remap.ptr, remap.data(), N);
  h2d(B->d_lut.ptr, lut.data(), N * 256);
  B->d_hist = DevBuf(B->ntiles * kRow * 4);
  B->d_excl = DevBuf((B->ntiles + 1) * kRow * 8);
  B->d_partial = DevBuf(scan_partial_entries(B->ntiles, 8) * 8);
  B->d_err = DevBuf(4);
  B->d_rows = DevBuf((N + 1) * kRow * 8);
  B->d_pc = DevBuf(N * (kRow + 1) * 8);
  // ---- per-text pass tables that do not depend on the counts
  std::vector<uint64_t> lw(N * 8, 0), lnw(N * 8, 0);
  std::vector<uint16_t> thr(N * 9, 0);
  std::vector<uint8_t> nk(N);
  for (size_t i = 0; i < N; ++i) {
    const TreeSlot& t = f->trees[i];
    const TreeShape& sh = *t.shape;
    const uint32_t L = t.L, nb = 1u << L;
    nk[i] = static_cast<uint8_t>(L);
    for (uint32_t l = 0; l < L; ++l) {
      lw[i * 8 + l] = t.level_word[l];
      lnw[i * 8 + l] = (t.n + 31) / 32;  // the level's capacity (bounds checks only)
    }
    if (sh.pow2) {
      for (uint32_t x = 0; x <= L; ++x) thr[i * 9 + x] = static_cast<uint16_t>(nb);
    } else {
      // first absent code per level, aligned to the level's groups
      int32_t maxd[9];
      for (uint32_t x = 0; x <= L; ++x) maxd[x] = -1;
      for (uint64_t c = 0; c < sh.sigma; ++c)
        for (uint32_t x = 0; x <= L; ++x)
          if (sh.len[c] > x) maxd[x] = std::max<int32_t>(maxd[x], static_cast<int32_t>(sh.code[c]));
      for (uint32_t x = 0; x <= L; ++x)
        thr[i * 9 + x] = maxd[x] < 0 ? 0
                                     : static_cast<uint16_t>(
                                           ((static_cast<uint32_t>(maxd[x]) >> (L - x)) + 1)
                                           << (L - x));
    }
  }
  B->d_thr = DevBuf(thr.size() * 2);
  B->d_nk = DevBuf(N);
  B->d_lw = DevBuf(N * 64);
  B->d_lnw = DevBuf(N * 64);
  h2d(B->d_thr.ptr, thr.data(), thr.size());
  h2d(B->d_nk.ptr, nk.data(), N);
  h2d(B->d_lw.ptr, lw.data(), lw.size());
  h2d(B->d_lnw.ptr, lnw.data(), lnw.size());
  // ---- shapes, metadata jobs, directory descriptors (sizes filled on the
  // device), view slots
  std::vector<ShapeDev> shp;
  std::vector<uint8_t> codes8, lens8, node_level;
  std::vector<uint16_t> nodes;
  std::map<uint64_t, uint32_t> shape_index;
  for (auto& kv : f->shapes) {
    const TreeShape& sh = kv.second;
    ShapeDev d{};
    d.sigma = static_cast<uint32_t>(sh.sigma);
    d.L = sh.L;
    d.code_off = static_cast<uint32_t>(codes8.size());
    d.node_off = static_cast<uint32_t>(node_level.size());
    for (uint64_t c = 0; c < sh.sigma; ++c) {
      codes8.push_back(static_cast<uint8_t>(sh.code[c]));
      lens8.push_back(sh.len[c]);
    }
    for (uint32_t l = 0; l < sh.L; ++l)
      for (const Node& nd : sh.nodes[l]) {
        nodes.push_back(static_cast<uint16_t>(nd.lo));
        nodes.push_back(static_cast<uint16_t>(nd.hi));
        node_level.push_back(static_cast<uint8_t>(l));
      }
    d.nodes = static_cast<uint32_t>(node_level.size()) - d.node_off;
    shape_index[kv.first] = static_cast<uint32_t>(shp.size());
    shp.push_back(d);
  }
  uint8_t* A = f->arena.as<uint8_t>();
  std::vector<MetaJob> jobs(N);
  std::vector<RsArrayDesc> desc;
  std::vector<LevelView*> lvptr;
  uint64_t blocks = 0, chunks = 0;
  for (size_t i = 0; i < N; ++i) {
    const TreeSlot& t = f->trees[i];
    MetaJob& J = jobs[i];
    J.n = t.n;
    J.counts_at = t.counts_at;
    J.ranks_at = t.ranks_at;
    J.shape = shape_index.at(t.sigma);
    J.desc0 = static_cast<uint32_t>(desc.size());
    const uint64_t cap_l1 = (t.n + detail::kL1Bits - 1) / detail::kL1Bits;
    for (uint32_t l = 0; l < t.L; ++l) {
      RsArrayDesc d{};
      d.words64 = reinterpret_cast<const uint64_t*>(A + t.level_word[l] * 4);
      d.l1 = reinterpret_cast<uint64_t*>(A + t.l1_at[l]);
      d.l2 = reinterpret_cast<uint16_t*>(A + t.l2_at[l]);
      d.samples = reinterpret_cast<uint64_t*>(A + t.smp_at[l]);
      d.total_ones = reinterpret_cast<uint64_t*>(A + t.ones_at) + l;
      d.first_tile = blocks;
      d.cap_l1 = cap_l1;
      d.first_chunk = chunks;
      blocks += cap_l1;
      chunks += rs_scan_chunks(cap_l1);
      desc.push_back(d);
      lvptr.push_back(reinterpret_cast<LevelView*>(A + t.lv_at) + l);
    }
  }
  B->blocks = blocks;
  B->chunks = chunks;
  B->ndesc = static_cast<uint32_t>(desc.size());
  B->d_jobs = DevBuf(N * sizeof(MetaJob));
  B->d_shapes = DevBuf(shp.size() * sizeof(ShapeDev));
  B->d_code = DevBuf(codes8.size());
  B->d_len = DevBuf(lens8.size());
  B->d_node = DevBuf(std::max<size_t>(nodes.size(), 1) * 2);
  B->d_node_level = DevBuf(std::max<size_t>(node_level.size(), 1));
  B->d_desc = DevBuf(desc.size() * sizeof(RsArrayDesc));
  B->d_blk = DevBuf(std::max<uint64_t>(blocks, 1) * 8);
  B->d_lvptr = DevBuf(lvptr.size() * sizeof(LevelView*));
  h2d(B->d_jobs.ptr, jobs.data(), N);
  h2d(B->d_shapes.ptr, shp.data(), shp.size());
  h2d(B->d_code.ptr, codes8.data(), codes8.size());
  h2d(B->d_len.ptr, lens8.data(), lens8.size());
  h2d(B->d_node.ptr, nodes.data(), nodes.size());
  h2d(B->d_node_level.ptr, node_level.data(), node_level.size());
  h2d(B->d_desc.ptr, desc.data(), desc.size());
  h2d(B->d_lvptr.ptr, lvptr.data(), lvptr.size());
  // ---- pass parameters
  PassParams& pp = B->pp;
  pp = PassParams{};
  pp.src = f->d_texts;
  pp.src_esize = 1;
  pp.codes_out = any_remap ? B->d_x.ptr : nullptr;
  pp.map_kind = 0;
  pp.sigma = 256;
  pp.pow2 = 1;
  pp.codes = f->d_texts;
  pp.cw = 1;
  pp.k = 8;
  pp.dshift = 0;
  pp.tile_rows = rows;
  pp.tile_codes = static_cast<uint32_t>(T);
  pp.num_tiles = B->ntiles;
  pp.tile_start = B->d_tstart.as<uint64_t>();
  pp.tile_len = B->d_tlen.as<uint32_t>();
  pp.tile_node = B->d_tnode.as<uint32_t>();
  pp.node_first_tile = B->d_tfirst.as<uint32_t>();
  pp.hist = B->d_hist.as<uint32_t>();
  pp.excl = B->d_excl.as<uint64_t>();
  pp.partial = B->d_partial.as<uint64_t>();
  pp.err = B->d_err.as<uint32_t>();
  pp.node_remap = any_remap ? B->d_remap.as<uint8_t>() : nullptr;
  pp.node_lut = B->d_lut.as<uint8_t>();
  pp.row_bins = kRow;
  pp.thr_stride = 9;
  pp.pc = B->d_pc.as<uint64_t>();
  pp.thr = B->d_thr.as<uint16_t>();
  pp.node_k = B->d_nk.as<uint8_t>();
  pp.node_lw = B->d_lw.as<uint64_t>();
  pp.node_lnw = B->d_lnw.as<uint64_t>();
  pp.words = f->arena.as<uint32_t>();
  pp.dst = nullptr;
  B->meta_at = f->trees.empty() ? 0 : f->trees.front().ones_at;
  return B;
}

// The build's launches, queued on s with no host wait in between (and so
// capturable into one graph): clear the bits and directories, tile digit
// histograms and their scan, the digit counts of every text, the metadata
// kernel, the level pass, the directories of every level, the level views.
void enqueue_build(ecl_forest_t* f, ForestBuild& B, hipStream_t s) {
  launch_fill_u32(B.d_err.as<uint32_t>(), 1, 0, s);
  launch_fill_u64(f->arena.as<uint64_t>(), B.meta_at / 8, 0, s);
  launch_tile_hist(B.pp, s);
  launch_tile_scan(B.pp, s);
  launch_gather_rows(B.d_excl.as<uint64_t>(), B.d_tfirst.as<uint32_t>(), B.N + 1, kRow,
                     B.d_rows.as<uint64_t>(), s);
  hipLaunchKernelGGL(forest_meta_kernel, dim3((B.N + 127) / 128), dim3(128), 0, s,
                     B.d_rows.as<uint64_t>(), B.d_jobs.as<MetaJob>(), B.N,
                     B.d_shapes.as<ShapeDev>(), B.d_code.as<uint8_t>(), B.d_len.as<uint8_t>(),
                     B.d_node.as<uint16_t>(), B.d_node_level.as<uint8_t>(),
                     f->arena.as<uint8_t>(), B.d_pc.as<uint64_t>(), B.d_desc.as<RsArrayDesc>());
  ECL_HIP(hipGetLastError());
  launch_pass(B.pp, s);
  launch_rs_build(B.d_desc.as<RsArrayDesc>(), B.ndesc, B.blocks, B.chunks,
                  B.d_blk.as<uint32_t>(), s);
  hipLaunchKernelGGL(forest_views_kernel, dim3((B.ndesc + 255) / 256), dim3(256), 0, s,
                     B.d_desc.as<RsArrayDesc>(), B.ndesc, B.d_lvptr.as<LevelView*>());
  hipLaunchKernelGGL(forest_clear_quad_kernel, dim3((B.N + 255) / 256), dim3(256), 0, s,
                     f->d_params.as<TreeQueryParams>(), B.N);
  ECL_HIP(hipGetLastError());
}

// A batch launch: the pair-lane kernel once the quad layouts were built
// (trees without them run their binary levels there), else the binary
// kernel with the forest's tuned shape.
void launch_batch(const ecl_forest_t* f, const ecl_forest_query* q, uint64_t m, uint64_t* out,
                  hipStream_t s) {
  if (f->quad_built)
    launch_forest_query_quad(f->d_views.as<TreeView>(), f->d_params.as<TreeQueryParams>(),
                             f->trees.size(), q, m, out, f->q_grid_quad, s);
  else
    launch_forest_query(f->d_views.as<TreeView>(), f->trees.size(), q, m, out, f->q_block,
                        f->q_grid, s);
}

// Waits for a build on s and checks the pass's bounds flag.
void finish_build(ForestBuild& B, hipStream_t s) {
  ECL_HIP(hipStreamSynchronize(s));
  uint32_t err = 0;
  d2h(&err, B.d_err.ptr, 1);
  if (err) throw Error(ECL_ERROR_INTERNAL, "construction pass wrote outside its arrays");
}

}  // namespace

extern "C" {

ecl_status ecl_forest_create(const void* d_texts, const uint64_t* starts, const uint64_t* lengths,
                             const uint64_t* sigmas, size_t num_texts, int device,
                             uint32_t tile_rows, ecl_forest* out) {
  return guard([&] {
    require(d_texts && starts && lengths && sigmas && out, "NULL argument");
    require(num_texts > 0 && num_texts < (1ull << 31), "invalid number of texts");
    require(tile_rows == 0 || tile_rows == 1 || tile_rows == 2 || tile_rows == 4,
            "tile_rows must be 0, 1, 2 or 4");
    require((reinterpret_cast<uintptr_t>(d_texts) & 15) == 0,
            "the texts must start on a 16-byte boundary");
    set_device(device);
    hipStream_t s = nullptr;
    const size_t N = num_texts;
    auto f = std::make_unique<ecl_forest_t>();
    f->device = device;
    f->d_texts = d_texts;
    f->starts.assign(starts, starts + N);
    f->lengths.assign(lengths, lengths + N);
    f->sigmas.assign(sigmas, sigmas + N);
    for (size_t i = 0; i < N; ++i) {
      require(starts[i] % 16 == 0, "every text must start on a 16-byte boundary");
      require(lengths[i] > 0, "every text must be non-empty");
      require(sigmas[i] >= 2 && sigmas[i] <= 256, "alphabet sizes must be in [2, 256]");
      if (!f->shapes.count(sigmas[i])) f->shapes.emplace(sigmas[i], tree_shape(sigmas[i]));
    }
    f->tile_rows = tile_rows;
    // ---- arena layout: every level at the capacity of a full level
    f->trees.resize(N);
    uint64_t at = 0;  // bits and directories first, then the metadata block
    for (size_t i = 0; i < N; ++i) {
      TreeSlot& t = f->trees[i];
      const TreeShape& sh = f->shapes.at(sigmas[i]);
      t.n = lengths[i];
      t.sigma = sh.sigma;
      t.L = sh.L;
      t.shape = &sh;
      // bit array: each level padded to 32 words (bit_array.cu:57-72)
      at = align_up(at, kAlign);
      const uint64_t nwords = (t.n + 31) / 32;
      uint64_t w = 0;
      t.level_word.resize(t.L);
      for (uint32_t l = 0; l < t.L; ++l) {
        t.level_word[l] = at / 4 + w;
        w += (nwords + 31) & ~31ull;
      }
      w += w % 2;
      at += std::max<uint64_t>(w, 2) * 4;
      // rank/select directory per level (rank_select.cu:57-143 sizes)
      const uint64_t n1 = (t.n + detail::kL1Bits - 1) / detail::kL1Bits;
      uint64_t last = ((t.n % detail::kL1Bits) + detail::kL2Bits - 1) / detail::kL2Bits;
      if (last == 0) last = detail::kL2PerL1;
      const uint64_t n2 = (n1 - 1) * detail::kL2PerL1 + last, cap = t.n / detail::kSampleRate;
      t.l1_at.resize(t.L); t.l2_at.resize(t.L); t.smp_at.resize(t.L);
      for (uint32_t l = 0; l < t.L; ++l) {
        at = align_up(at, 8);
        t.l1_at[l] = at; at += n1 * 8;
        t.smp_at[l] = at; at += cap * 8;
        t.l2_at[l] = at; at += n2 * 2;
      }
    }
    // metadata block: total ones, counts, ranks, nodes per level, codes,
    // code lengths and level views of every tree
    at = align_up(at, kAlign);
    const uint64_t meta_at = at;
    for (auto& t : f->trees) { t.ones_at = at; at += t.L * 8ull; }
    std::map<uint64_t, TreeMeta> shape_meta;  // node counts per shape (from unit counts)
    for (auto& kv : f->shapes)
      shape_meta.emplace(kv.first, tree_meta(kv.second, std::vector<uint64_t>(kv.first, 1),
                                             kv.first));
    for (auto& t : f->trees) {
      const TreeMeta& sm = shape_meta.at(t.sigma);
      t.num_ranks = static_cast<uint32_t>(sm.ranks.size());
      t.counts_at = at; at += t.sigma * 8;
      t.ranks_at = at; at += t.num_ranks * 8ull;
      t.nal_at = at; at += sm.nodes_at_level.size() * 8;
      const uint64_t ncodes = t.shape->pow2 ? 0 : t.sigma - t.shape->codes_start;
      t.codes_at = at; at += ncodes * 8;
      t.lens_at = at; at += align_up(ncodes, 8);
    }
    at = align_up(at, 16);
    for (auto& t : f->trees) { t.lv_at = at; at += t.L * sizeof(LevelView); }
    const uint64_t arena_bytes = align_up(at, kAlign);
    if (arena_bytes > free_device_memory())
      throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, "Not enough memory available for the wavelet tree.");
    f->arena = DevBuf(arena_bytes);
    uint8_t* A = f->arena.as<uint8_t>();
    // ---- the static part of the metadata block (the build's kernels write
    // the counts, node ranks, totals and level views) and the tree views
    std::vector<uint8_t> img(arena_bytes - meta_at, 0);
    auto put = [&](uint64_t where, const void* src, size_t bytes) {
      if (bytes) std::memcpy(img.data() + (where - meta_at), src, bytes);
    };
    f->views.resize(N);
    for (size_t i = 0; i < N; ++i) {
      TreeSlot& t = f->trees[i];
      const TreeShape& sh = *t.shape;
      const TreeMeta& sm = shape_meta.at(t.sigma);
      put(t.nal_at, sm.nodes_at_level.data(), sm.nodes_at_level.size() * 8);
      if (!sh.pow2) {
        std::vector<uint64_t> codes;
        std::vector<uint8_t> lens;
        for (uint64_t c = sh.codes_start; c < sh.sigma; ++c) {
          codes.push_back(sh.code[c]);
          lens.push_back(sh.len[c]);
        }
        put(t.codes_at, codes.data(), codes.size() * 8);
        put(t.lens_at, lens.data(), lens.size());
      }
      for (uint32_t l = 0; l < t.L; ++l)
        f->d_levels.push_back(reinterpret_cast<LevelView*>(A + t.lv_at) + l);
      TreeView& v = f->views[i];
      v = TreeView{};
      v.levels = reinterpret_cast<const LevelView*>(A + t.lv_at);
      v.counts = reinterpret_cast<const uint64_t*>(A + t.counts_at);
      v.ranks = reinterpret_cast<const uint64_t*>(A + t.ranks_at);
      v.nodes_at_level = reinterpret_cast<const uint64_t*>(A + t.nal_at);
      v.codes = sh.pow2 ? nullptr : reinterpret_cast<const uint64_t*>(A + t.codes_at);
      v.code_lens = sh.pow2 ? nullptr : reinterpret_cast<const uint8_t*>(A + t.lens_at);
      v.n = t.n;
      v.sigma = t.sigma;
      v.codes_start = sh.codes_start;
      v.nodes_until_last_level = sm.nodes_until_last_level;
      v.num_ranks = t.num_ranks;
      v.num_levels = t.L;
      v.pow2 = sh.pow2 ? 1 : 0;
    }
    h2d(A + meta_at, img.data(), img.size());
    f->d_views = DevBuf(N * sizeof(TreeView));
    h2d(f->d_views.ptr, f->views.data(), N);
    f->params.assign(N, TreeQueryParams{});
    for (size_t i = 0; i < N; ++i) {
      f->params[i].view = f->views[i];
      f->params[i].esize = 1;
      f->params[i].symbols_minimal = 1;
    }
    f->d_params = DevBuf(N * sizeof(TreeQueryParams));
    h2d(f->d_params.ptr, f->params.data(), N);
    // ---- the build: one queue of launches, one wait at the end
    {
      std::shared_ptr<ForestBuild> B = make_build(f.get(), choose_rows(f.get(), tile_rows));
      enqueue_build(f.get(), *B, s);
      finish_build(*B, s);
    }
    f->q_grid = static_cast<uint32_t>(cu_count(device)) * 8;
    f->q_grid_quad = f->q_grid;
    *out = f.release();
  });
}

// new: the forest's construction as one graph that rebuilds every tree from
// the current contents of its texts (same buffer, starts, lengths and
// alphabets); the graph keeps the build's device buffers.
ecl_status ecl_forest_build_graph(ecl_forest f, ecl_graph* out) {
  return guard([&] {
    require(f && out, "NULL argument");
    set_device(f->device);
    if (!f->build) f->build = make_build(f, choose_rows(f, f->tile_rows));
    auto g = std::make_unique<ecl_graph_t>();
    g->device = f->device;
    g->keep = f->build;
    hipStream_t s = nullptr;
    ECL_HIP(hipStreamCreate(&s));
    try {
      ECL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      enqueue_build(f, *f->build, s);
      ECL_HIP(hipStreamEndCapture(s, &g->graph));
      ECL_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      hipGraph_t dummy = nullptr;
      (void)hipStreamEndCapture(s, &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      (void)hipStreamDestroy(s);
      throw;
    }
    ECL_HIP(hipStreamDestroy(s));
    *out = g.release();
  });
}

// new: times the build (the queued launches) with tiles of 1, 2 and 4
// chunks per thread on this forest's texts, keeps the fastest for later
// build graphs, and leaves the forest built (every tile size builds the
// same trees).
ecl_status ecl_forest_autotune_build(ecl_forest f, uint32_t* tile_rows, float* ms) {
  return guard([&] {
    require(f, "NULL argument");
    set_device(f->device);
    hipStream_t s = nullptr;
    hipEvent_t e0, e1;
    ECL_HIP(hipEventCreate(&e0));
    ECL_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    uint32_t best_rows = choose_rows(f, 0);
    try {
      for (uint32_t r : {1u, 2u, 4u}) {
        std::shared_ptr<ForestBuild> B = make_build(f, r);
        enqueue_build(f, *B, s);  // warm
        finish_build(*B, s);
        ECL_HIP(hipEventRecord(e0, s));
        enqueue_build(f, *B, s);
        ECL_HIP(hipEventRecord(e1, s));
        finish_build(*B, s);
        float t = 0.f;
        ECL_HIP(hipEventElapsedTime(&t, e0, e1));
        if (t < best) {
          best = t;
          best_rows = r;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    f->tile_rows = best_rows;
    f->build.reset();
    if (tile_rows) *tile_rows = best_rows;
    if (ms) *ms = best;
  });
}

ecl_status ecl_forest_build_query_lines(ecl_forest f) {
  return guard([&] {
    require(f, "NULL argument");
    set_device(f->device);
    // the level views as the last build left them (a rebuild from a build
    // graph drops the lines of the trees it rebuilt)
    const size_t na = f->d_levels.size();
    // (the views of all trees are consecutive in the arena: one copy)
    std::vector<LevelView> views(na);
    if (na)
      d2h(views.data(),
          reinterpret_cast<const LevelView*>(f->arena.as<uint8_t>() + f->trees.front().lv_at), na);
    bool have = na > 0;
    for (const LevelView& v : views) have = have && (v.lines != nullptr || v.size == 0);
    if (have && f->lines.ptr != nullptr) return;
    // the views drop the old lines and hints before they are freed (an out-
    // of-memory error below then leaves no view pointing at freed memory)
    if (na) {
      for (LevelView& v : views) {
        v.lines = nullptr;
        v.sel_hints = nullptr;
        v.hint_values = 0;
      }
      h2d(reinterpret_cast<LevelView*>(f->arena.as<uint8_t>() + f->trees.front().lv_at),
          views.data(), na);
    }
    f->lines.reset();
    require(na < (1ull << 32), "too many bit arrays");
    std::vector<LinesJob> jobs(na);
    std::vector<uint64_t> starts(2 * na);  // line starts | hint starts
    uint64_t tl = 0, th = 0;
    for (size_t a = 0; a < na; ++a) {
      const LevelView& v = views[a];
      LinesJob& J = jobs[a];
      J.v = v;
      J.v.lines = nullptr;
      J.v.sel_hints = nullptr;
      J.v.hint_values = 0;
      J.d_view = f->d_levels[a];
      J.nl = (v.size + detail::kLineBits - 1) / detail::kLineBits;
      J.nh1 = v.size ? v.ones / detail::kSampleRate + 1 : 0;
      J.nh0 = v.size ? (v.size - v.ones) / detail::kSampleRate + 1 : 0;
      starts[a] = tl;
      starts[na + a] = th;
      tl += J.nl;
      th += J.nh1 + J.nh0;
    }
    const uint64_t bytes = tl * 128 + th * 64;
    const uint64_t meta = na * (sizeof(LinesJob) + 16 + 8);
    if (bytes + meta + (64u << 20) > free_device_memory())
      throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, "Not enough memory available for the wavelet tree.");
    DevBuf lines(std::max<uint64_t>(bytes, 1));
    DevBuf d_jobs(std::max<size_t>(na, 1) * sizeof(LinesJob));
    DevBuf d_starts(std::max<size_t>(na, 1) * 16);
    DevBuf d_fit(std::max<size_t>(na, 1) * 8);
    hipStream_t s = nullptr;
    h2d(d_jobs.ptr, jobs.data(), na);
    h2d(d_starts.ptr, starts.data(), 2 * na);
    launch_fill_u32(d_fit.as<uint32_t>(), std::max<size_t>(na, 1) * 2, 0, s);
    uint4* L = lines.as<uint4>();
    uint4* H = L + tl * 8;
    launch_multi_lines(d_jobs.as<LinesJob>(), d_starts.as<uint64_t>(), d_starts.as<uint64_t>() + na,
                       static_cast<uint32_t>(na), L, tl, H, th, d_fit.as<uint32_t>(), s);
    ECL_HIP(hipStreamSynchronize(s));
    f->lines = std::move(lines);
  });
}

// Select lines for every level of every tree (ecl_forest_build_select_lines):
// the rates and line counts from each level's size and ones (as for one
// tree's directory), then one launch over all levels, then the level views
// get the lines.  Old lines are dropped from the views before they are
// freed, so that the views never point at freed memory.
ecl_status ecl_forest_build_select_lines(ecl_forest f) {
  return guard([&] {
    require(f, "NULL argument");
    set_device(f->device);
    const size_t na = f->d_levels.size();
    if (na == 0) return;
    require(na < (1ull << 32), "too many bit arrays");
    // (the views of all trees are consecutive in the arena: one copy)
    LevelView* d_views =
        reinterpret_cast<LevelView*>(f->arena.as<uint8_t>() + f->trees.front().lv_at);
    std::vector<LevelView> views(na);
    d2h(views.data(), d_views, na);
    bool have = f->sel_lines.ptr != nullptr;
    for (const LevelView& v : views) have = have && (v.sel_lines != nullptr || v.size == 0);
    if (have) return;
    for (LevelView& v : views) {
      v.sel_lines = nullptr;
      v.sel_zero_line = 0;
      v.sel_rate[0] = v.sel_rate[1] = 0;
      v.sel_magic[0] = v.sel_magic[1] = 0;
    }
    h2d(d_views, views.data(), na);
    f->sel_lines.reset();
    std::vector<SelLinesJob> jobs(na);
    std::vector<uint64_t> wg_start(na);
    uint64_t total = 0, wgs = 0;
    for (size_t a = 0; a < na; ++a) {
      const LevelView& v = views[a];
      SelLinesJob& J = jobs[a];
      J.v = v;
      const uint64_t zeros = v.size - v.ones;
      J.r1 = v.size ? select_line_rate(v.ones, v.size) : 0;
      J.r0 = v.size ? select_line_rate(zeros, v.size) : 0;
      J.n1 = J.r1 ? v.ones / J.r1 + 1 : 0;
      const uint64_t n0 = J.r0 ? zeros / J.r0 + 1 : 0;
      J.nl2 = (v.size + detail::kL2Bits - 1) / detail::kL2Bits;
      J.nchunks = J.n1 + n0 ? (J.nl2 + kSelLineChunk - 1) / kSelLineChunk : 0;
      J.line_off = total;
      wg_start[a] = wgs;
      total += J.n1 + n0;
      wgs += 2 * J.nchunks;
    }
    const uint64_t meta = na * (sizeof(SelLinesJob) + 8);
    if (total * 128 + meta + (64u << 20) > free_device_memory())
      throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, "Not enough memory available for the wavelet tree.");
    DevBuf lines(std::max<uint64_t>(total, 1) * 128);
    DevBuf d_jobs(na * sizeof(SelLinesJob));
    DevBuf d_start(na * 8);
    h2d(d_jobs.ptr, jobs.data(), na);
    h2d(d_start.ptr, wg_start.data(), na);
    hipStream_t s = nullptr;
    launch_multi_select_lines(d_jobs.as<SelLinesJob>(), d_start.as<uint64_t>(),
                              static_cast<uint32_t>(na), wgs, lines.as<uint4>(), s);
    ECL_HIP(hipStreamSynchronize(s));
    for (size_t a = 0; a < na; ++a) {
      const SelLinesJob& J = jobs[a];
      if (J.nchunks == 0) continue;
      LevelView& v = views[a];
      v.sel_lines = lines.as<uint4>() + J.line_off * 8;
      v.sel_zero_line = J.n1;
      v.sel_rate[1] = J.r1;
      v.sel_rate[0] = J.r0;
      v.sel_magic[1] = select_line_magic(J.r1);
      v.sel_magic[0] = select_line_magic(J.r0);
    }
    h2d(d_views, views.data(), na);
    f->sel_lines = std::move(lines);
  });
}

// Quad lines and quad select lines for every tree over a power-of-two
// alphabet with an even number of levels (ecl_forest_build_quad_layouts).
// Per tree and quad level the host tables follow from the tree's symbol
// counts (one copy of the forest's metadata block), as for a single tree:
// node starts S and digit counts CD for the build and the select ascent,
// bases for the descents, the select lines' rates from the level's digit
// totals.  The lines of all levels of all trees are then built in three
// launches, and the trees' parameters get their quad views.
ecl_status ecl_forest_build_quad_layouts(ecl_forest f) {
  return guard([&] {
    require(f, "NULL argument");
    set_device(f->device);
    const size_t N = f->trees.size();
    auto eligible = [&](const TreeSlot& t) {
      return t.shape->pow2 && t.L % 2 == 0 && t.L >= 2 && t.L <= 2 * kMaxQuadLevels;
    };
    // the parameters as the last build left them (a build clears the quad
    // views of the trees it rebuilt)
    std::vector<TreeQueryParams> cur(N);
    d2h(cur.data(), f->d_params.as<TreeQueryParams>(), N);
    bool have = f->quad_built;
    for (size_t i = 0; i < N; ++i) have = have && (!eligible(f->trees[i]) || cur[i].quad.has_sel);
    if (have) return;
    // every tree answers over its binary levels before the old arrays go
    for (size_t i = 0; i < N; ++i) {
      f->params[i].quad = QuadView{};
      f->params[i].hex = HexView{};
    }
    h2d(f->d_params.ptr, f->params.data(), N);
    f->quad_lines.reset();
    f->quad_sel.reset();
    f->quad_base.reset();
    f->quad_tab.reset();
    f->hex_lines.reset();
    f->hex_tab.reset();
    // the symbol counts of every tree (the metadata block, one copy)
    const uint64_t meta_at = f->trees.front().ones_at;
    const TreeSlot& last = f->trees.back();
    const uint64_t meta_end = last.lv_at + last.L * sizeof(LevelView);
    std::vector<uint8_t> meta(meta_end - meta_at);
    d2h(meta.data(), f->arena.as<uint8_t>() + meta_at, meta.size());
    auto level_view = [&](const TreeSlot& t, uint32_t l) {
      LevelView v;
      std::memcpy(&v, meta.data() + (t.lv_at - meta_at) + l * sizeof(LevelView), sizeof(v));
      return v;
    };
    // host tables and the jobs (one per tree and quad level)
    std::vector<uint64_t> base, tab;
    std::vector<QuadBuildParams> jobs;
    std::vector<QuadSelBuildParams> sel_jobs;
    std::vector<uint64_t> wg_start, sel_start;
    std::vector<uint32_t> job_tree;
    std::vector<uint64_t> tree_base(N, 0), tree_tab(N, 0);
    std::vector<QuadView> qviews(N);
    // per job: the first of its level's lines and select lines in the
    // forest's arrays
    std::vector<uint64_t> job_line, job_sel;
    uint64_t nql_total = 0, sel_total = 0, wg_total = 0, selwg_total = 0;
    for (size_t i = 0; i < N; ++i) {
      const TreeSlot& t = f->trees[i];
      if (!eligible(t)) continue;
      const uint32_t L = t.L, Q = L / 2;
      const uint64_t n = t.n, sigma = t.sigma;
      std::vector<uint64_t> C(sigma + 1);
      std::memcpy(C.data(), meta.data() + (t.counts_at - meta_at), sigma * 8);
      C[sigma] = n;
      QuadView& qv = qviews[i];
      tree_base[i] = base.size();
      tree_tab[i] = tab.size();
      const uint64_t nql = n / kQuadDigits + 1;
      for (uint32_t q = 0; q < Q; ++q) {
        const uint64_t nn = 1ull << (2 * q);
        const uint32_t shift = L - 2 * q - 2;
        qv.off[q] = base.size() - tree_base[i];
        qv.s_off[q] = tab.size() - tree_tab[i];
        for (uint64_t Nd = 0; Nd <= nn; ++Nd)
          tab.push_back(C[std::min<uint64_t>(Nd << (shift + 2), sigma)]);
        qv.cd_off[q] = tab.size() - tree_tab[i];
        uint64_t acc[4] = {0, 0, 0, 0};
        for (uint64_t Nd = 0; Nd <= nn; ++Nd) {
          for (uint32_t d = 0; d < 4; ++d) tab.push_back(acc[d]);
          if (Nd == nn) break;
          for (uint32_t d = 0; d < 4; ++d) {
            const uint64_t lo = (4 * Nd + d) << shift, hi = (4 * Nd + d + 1) << shift;
            base.push_back(C[lo] - acc[d]);
            acc[d] += C[hi] - C[lo];
          }
        }
        // the level's select lines: a line every r of each digit's acc[d]
        uint64_t run = 0;
        for (uint32_t d = 0; d < 4; ++d) {
          const uint64_t tot = acc[d];
          const uint64_t r = tot ? std::max<uint64_t>(1, kQuadSelSpan * tot / n) : 0;
          qv.sel_rate[q][d] = static_cast<uint32_t>(r);
          qv.sel_magic[q][d] = select_line_magic(static_cast<uint32_t>(r));
          qv.sel_first[q][d] = run;
          run += r ? tot / r + 1 : 0;
        }
        QuadBuildParams qp{};
        qp.a = level_view(t, 2 * q);
        qp.b = level_view(t, 2 * q + 1);
        qp.nodes = nn;
        qp.n = n;
        qp.nlines = nql;
        jobs.push_back(qp);  // table, line and scratch addresses below
        QuadSelBuildParams sp{};
        sp.nql = nql;
        for (uint32_t d = 0; d < 4; ++d) {
          sp.total[d] = acc[d];
          sp.r[d] = qv.sel_rate[q][d];
          sp.magic[d] = qv.sel_magic[q][d];
          sp.nlines[d] = sp.r[d] ? acc[d] / sp.r[d] + 1 : 0;
        }
        sel_jobs.push_back(sp);
        job_tree.push_back(static_cast<uint32_t>(i));
        wg_start.push_back(wg_total);
        sel_start.push_back(selwg_total);
        wg_total += quad_build_workgroups(nql);
        selwg_total += quad_select_build_workgroups(nql);
        job_line.push_back(nql_total);
        job_sel.push_back(sel_total);
        nql_total += nql;
        sel_total += run;
      }
      qv.levels = Q;
      qv.nbase = base.size() - tree_base[i];
      qv.ntab = tab.size() - tree_tab[i];
      qv.has_sel = 1;
    }
    // hex lines for the trees whose level count is a multiple of 4, from
    // their quad lines: per hex level the bases and (built on the device)
    // superblock counts, kept; the x's before each node (CHD) and the node
    // starts, for the build only
    std::vector<uint64_t> htab, hbtab;
    std::vector<HexBuildParams> hjobs;
    std::vector<uint32_t> hjob_tree;
    std::vector<uint64_t> hwg_start, hjob_line, hjob_chd, hjob_s, tree_htab(N, 0);
    std::vector<HexView> hviews(N);
    uint64_t nhl_total = 0, hwg_total = 0;
    for (size_t i = 0; i < N; ++i) {
      const TreeSlot& t = f->trees[i];
      if (!eligible(t) || t.L % 4 != 0) continue;
      const uint32_t L = t.L, H = L / 4;
      const uint64_t n = t.n, sigma = t.sigma;
      std::vector<uint64_t> C(sigma + 1);
      std::memcpy(C.data(), meta.data() + (t.counts_at - meta_at), sigma * 8);
      C[sigma] = n;
      const uint64_t nlines = n / kHexDigits + 1;
      const uint64_t nsb = (nlines >> kHexSbShift) + 1;
      HexView& hv = hviews[i];
      tree_htab[i] = htab.size();
      for (uint32_t h = 0; h < H; ++h) {
        const uint64_t nn = 1ull << (4 * h);
        const uint32_t shift = L - 4 * h - 4;
        hv.base_off[h] = htab.size() - tree_htab[i];
        hjob_chd.push_back(hbtab.size());
        uint64_t acc[16] = {};
        for (uint64_t Nd = 0; Nd <= nn; ++Nd) {
          for (uint32_t x = 0; x < 16; ++x) hbtab.push_back(acc[x]);
          if (Nd == nn) break;
          for (uint32_t x = 0; x < 16; ++x) {
            const uint64_t lo = (16 * Nd + x) << shift, hi = (16 * Nd + x + 1) << shift;
            htab.push_back(C[lo] - acc[x]);
            acc[x] += C[hi] - C[lo];
          }
        }
        hjob_s.push_back(hbtab.size());
        for (uint64_t Nd = 0; Nd <= nn; ++Nd)
          hbtab.push_back(C[std::min<uint64_t>(Nd << (shift + 4), sigma)]);
        HexBuildParams hp{};
        hp.nodes = nn;
        hp.n = n;
        hp.nql = n / kQuadDigits + 1;
        hp.nlines = nlines;
        hjobs.push_back(hp);  // addresses below
        hjob_tree.push_back(static_cast<uint32_t>(i));
        hjob_line.push_back(nhl_total);
        hwg_start.push_back(hwg_total);
        nhl_total += nlines;
        hwg_total += hex_build_workgroups(nlines);
      }
      for (uint32_t h = 0; h < H; ++h) {
        hv.sb_off[h] = htab.size() - tree_htab[i];
        htab.resize(htab.size() + 16 * nsb, 0);  // written by the build kernel
      }
      hv.levels = H;
      hv.ntab = htab.size() - tree_htab[i];
    }
    const size_t nj = jobs.size(), nh = hjobs.size();
    if (nj == 0) return;  // no tree takes them: batches keep the binary kernel
    require(nj < (1ull << 32) && nh < (1ull << 32), "too many quad levels");
    const uint64_t bytes = (nql_total + sel_total + nhl_total) * 128 +
                           (base.size() + tab.size() + htab.size() + hbtab.size()) * 8 +
                           wg_total * kQuadWgInfo * 8 +
                           nj * (sizeof(QuadBuildParams) + sizeof(QuadSelBuildParams) + 16) +
                           nh * (sizeof(HexBuildParams) + 8);
    if (bytes + (64u << 20) > free_device_memory())
      throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, "Not enough memory available for the wavelet tree.");
    DevBuf lines(nql_total * 128), sel(std::max<uint64_t>(sel_total, 1) * 128);
    DevBuf dbase(std::max<size_t>(base.size(), 1) * 8), dtab(tab.size() * 8);
    DevBuf wg_info(wg_total * kQuadWgInfo * 8);
    DevBuf dlut(4096);
    {
      // the 4-position deposit table (QuadBuildParams::lut)
      std::vector<uint8_t> lut(4096);
      for (uint32_t x = 0; x < 4096; ++x) {
        const uint32_t a = x & 15, z = (x >> 4) & 15, o = x >> 8;
        uint32_t iz = 0, io = 0, v = 0;
        for (uint32_t k = 0; k < 4; ++k) {
          const uint32_t ak = (a >> k) & 1, bk = ak ? (o >> io++) & 1 : (z >> iz++) & 1;
          v |= ((ak << 1) | bk) << (2 * k);
        }
        lut[x] = static_cast<uint8_t>(v);
      }
      h2d(dlut.ptr, lut.data(), 4096);
    }
    if (!base.empty()) h2d(dbase.ptr, base.data(), base.size());
    h2d(dtab.ptr, tab.data(), tab.size());
    // the device addresses of every view and job (a tree's jobs are its
    // quad levels in order)
    for (size_t i = 0; i < N; ++i) {
      if (!qviews[i].levels) continue;
      qviews[i].base = dbase.as<uint64_t>() + tree_base[i];
      qviews[i].tab = dtab.as<uint64_t>() + tree_tab[i];
    }
    for (size_t j = 0, q = 0; j < nj; ++j) {
      const uint32_t i = job_tree[j];
      q = (j > 0 && job_tree[j - 1] == i) ? q + 1 : 0;
      QuadView& qv = qviews[i];
      qv.lines[q] = lines.as<uint4>() + job_line[j] * 8;
      qv.sel_lines[q] = sel.as<uint4>() + job_sel[j] * 8;
      QuadBuildParams& qp = jobs[j];
      qp.S = qv.tab + qv.s_off[q];
      qp.CD = qv.tab + qv.cd_off[q];
      qp.out = const_cast<uint4*>(qv.lines[q]);
      qp.wg_info = wg_info.as<uint64_t>() + wg_start[j] * kQuadWgInfo;
      qp.lut = dlut.as<uint8_t>();
      QuadSelBuildParams& sp = sel_jobs[j];
      sp.qlines = qv.lines[q];
      for (uint32_t d = 0; d < 4; ++d)
        sp.out[d] = const_cast<uint4*>(qv.sel_lines[q]) + qv.sel_first[q][d] * 8;
    }
    DevBuf d_jobs(nj * sizeof(QuadBuildParams)), d_sel_jobs(nj * sizeof(QuadSelBuildParams));
    DevBuf d_wg(nj * 8), d_selwg(nj * 8);
    h2d(d_jobs.ptr, jobs.data(), nj);
    h2d(d_sel_jobs.ptr, sel_jobs.data(), nj);
    h2d(d_wg.ptr, wg_start.data(), nj);
    h2d(d_selwg.ptr, sel_start.data(), nj);
    hipStream_t s = nullptr;
    launch_multi_quad_layouts(d_jobs.as<QuadBuildParams>(), d_sel_jobs.as<QuadSelBuildParams>(),
                              d_wg.as<uint64_t>(), d_selwg.as<uint64_t>(),
                              static_cast<uint32_t>(nj), wg_total, selwg_total, s);
    DevBuf hlines(std::max<uint64_t>(nhl_total, 1) * 128);
    DevBuf dhtab(std::max<size_t>(htab.size(), 1) * 8), dhbtab(std::max<size_t>(hbtab.size(), 1) * 8);
    DevBuf d_hjobs(std::max<size_t>(nh, 1) * sizeof(HexBuildParams)), d_hwg(std::max<size_t>(nh, 1) * 8);
    if (nh) {
      h2d(dhtab.ptr, htab.data(), htab.size());
      h2d(dhbtab.ptr, hbtab.data(), hbtab.size());
      for (size_t j = 0, h = 0; j < nh; ++j) {
        const uint32_t i = hjob_tree[j];
        h = (j > 0 && hjob_tree[j - 1] == i) ? h + 1 : 0;
        const QuadView& qv = qviews[i];
        HexView& hv = hviews[i];
        HexBuildParams& hp = hjobs[j];
        hp.qa = qv.lines[2 * h];
        hp.qb = qv.lines[2 * h + 1];
        hp.qbase = qv.base + qv.off[2 * h];
        hp.qcd = qv.tab + qv.cd_off[2 * h + 1];
        hp.chd = dhbtab.as<uint64_t>() + hjob_chd[j];
        hp.S = dhbtab.as<uint64_t>() + hjob_s[j];
        hp.out = hlines.as<uint4>() + hjob_line[j] * 8;
        hp.sb = dhtab.as<uint64_t>() + tree_htab[i] + hv.sb_off[h];
        hv.lines[h] = hp.out;
        hv.tab = dhtab.as<uint64_t>() + tree_htab[i];
      }
      h2d(d_hjobs.ptr, hjobs.data(), nh);
      h2d(d_hwg.ptr, hwg_start.data(), nh);
      launch_multi_hex_lines(d_hjobs.as<HexBuildParams>(), d_hwg.as<uint64_t>(),
                             static_cast<uint32_t>(nh), hwg_total, s);
    }
    ECL_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < N; ++i) {
      if (qviews[i].levels) f->params[i].quad = qviews[i];
      if (hviews[i].levels) f->params[i].hex = hviews[i];
    }
    h2d(f->d_params.ptr, f->params.data(), N);
    f->quad_lines = std::move(lines);
    f->quad_sel = std::move(sel);
    f->quad_base = std::move(dbase);
    f->quad_tab = std::move(dtab);
    f->hex_lines = std::move(hlines);
    f->hex_tab = std::move(dhtab);
    f->quad_built = true;
  });
}

ecl_status ecl_forest_destroy(ecl_forest f) {
  return guard([&] {
    if (!f) return;
    set_device(f->device);
    delete f;
  });
}

ecl_status ecl_forest_num_trees(ecl_forest f, size_t* out) {
  return guard([&] {
    require(f && out, "NULL argument");
    *out = f->trees.size();
  });
}

ecl_status ecl_forest_tree_info(ecl_forest f, size_t i, uint64_t* n, uint64_t* sigma,
                                uint32_t* num_levels) {
  return guard([&] {
    require(f, "NULL argument");
    require(i < f->trees.size(), "tree index out of range");
    if (n) *n = f->trees[i].n;
    if (sigma) *sigma = f->trees[i].sigma;
    if (num_levels) *num_levels = f->trees[i].L;
  });
}

ecl_status ecl_forest_device_view(ecl_forest f, size_t i, void* view, size_t view_bytes) {
  return guard([&] {
    require(f && view && view_bytes == sizeof(TreeView), "view size mismatch");
    require(i < f->trees.size(), "tree index out of range");
    std::memcpy(view, &f->views[i], sizeof(TreeView));
  });
}

ecl_status ecl_forest_query_device(ecl_forest f, const ecl_forest_query* d_queries, size_t m,
                                   uint64_t* d_out, void* stream) {
  return guard([&] {
    require(f, "NULL forest");
    if (m == 0) return;
    require(d_queries && d_out, "NULL argument");
    set_device(f->device);
    launch_batch(f, d_queries, m, d_out, static_cast<hipStream_t>(stream));
  });
}

ecl_status ecl_forest_generate_queries(ecl_forest f, const uint64_t* qoff, uint64_t seed,
                                       ecl_forest_query* d_out, void* stream) {
  return guard([&] {
    require(f && qoff, "NULL argument");
    const size_t N = f->trees.size();
    for (size_t i = 0; i < N; ++i) require(qoff[i + 1] >= qoff[i], "query offsets must not decrease");
    const uint64_t m = qoff[N] - qoff[0];
    require(qoff[0] == 0, "query offsets must start at 0");
    if (m == 0) return;
    require(d_out, "NULL output");
    set_device(f->device);
    DevBuf d_qoff((N + 1) * 8);
    h2d(d_qoff.ptr, qoff, N + 1);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    launch_forest_gen(f->d_views.as<TreeView>(), N, d_qoff.as<uint64_t>(), seed, m, d_out, s);
    ECL_HIP(hipStreamSynchronize(s));  // d_qoff is freed on return
  });
}

ecl_status ecl_forest_autotune(ecl_forest f, const ecl_forest_query* d_queries, size_t m,
                               uint64_t* d_out, uint32_t* block, uint32_t* grid) {
  return guard([&] {
    require(f && d_queries && d_out, "NULL argument");
    require(m > 0, "empty batch");
    set_device(f->device);
    const uint64_t sample = std::min<uint64_t>(m, 1ull << 22);
    const uint32_t cus = static_cast<uint32_t>(cu_count(f->device));
    const bool quad = f->quad_built;  // the kernel the batches launch
    hipEvent_t e0, e1;
    ECL_HIP(hipEventCreate(&e0));
    ECL_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    uint32_t bb = quad ? 256u : f->q_block, bg = quad ? f->q_grid_quad : f->q_grid;
    auto launch = [&](uint32_t b, uint32_t g) {
      if (quad)
        launch_forest_query_quad(f->d_views.as<TreeView>(), f->d_params.as<TreeQueryParams>(),
                                 f->trees.size(), d_queries, sample, d_out, g, nullptr);
      else
        launch_forest_query(f->d_views.as<TreeView>(), f->trees.size(), d_queries, sample, d_out,
                            b, g, nullptr);
    };
    try {
      // the pair-lane kernel runs 256-thread workgroups
      const std::vector<uint32_t> blocks =
          quad ? std::vector<uint32_t>{256u} : std::vector<uint32_t>{64u, 128u, 256u, 512u};
      for (uint32_t b : blocks) {
        for (uint32_t per_cu : {1u, 2u, 4u, 8u, 16u, 32u}) {
          const uint32_t g = cus * per_cu * 256 / b;  // same waves per CU across block sizes
          launch(b, g);
          ECL_HIP(hipEventRecord(e0, nullptr));
          for (int r = 0; r < 3; ++r) launch(b, g);
          ECL_HIP(hipEventRecord(e1, nullptr));
          ECL_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          ECL_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) {
            best = ms;
            bb = b;
            bg = g;
          }
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (quad) {
      f->q_grid_quad = bg;
    } else {
      f->q_block = bb;
      f->q_grid = bg;
    }
    if (block) *block = bb;
    if (grid) *grid = bg;
  });
}

ecl_status ecl_forest_graph_create(ecl_forest f, const ecl_forest_query* d_queries, size_t m,
                                   uint64_t* d_out, ecl_graph* out) {
  return guard([&] {
    require(f && d_queries && d_out && out, "NULL argument");
    require(m > 0, "empty batch");
    set_device(f->device);
    auto g = std::make_unique<ecl_graph_t>();
    g->device = f->device;
    hipStream_t s = nullptr;
    ECL_HIP(hipStreamCreate(&s));
    try {
      ECL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      launch_batch(f, d_queries, m, d_out, s);
      ECL_HIP(hipStreamEndCapture(s, &g->graph));
      ECL_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      hipGraph_t dummy = nullptr;
      (void)hipStreamEndCapture(s, &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      (void)hipStreamDestroy(s);
      throw;
    }
    ECL_HIP(hipStreamDestroy(s));
    *out = g.release();
  });
}

ecl_status ecl_graph_launch(ecl_graph g, void* stream) {
  return guard([&] {
    require(g, "NULL graph");
    set_device(g->device);
    ECL_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  });
}

ecl_status ecl_graph_destroy(ecl_graph g) {
  return guard([&] {
    if (!g) return;
    set_device(g->device);
    delete g;
  });
}

}  // extern "C"
