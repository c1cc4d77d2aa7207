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
uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// One tree of the forest: its shape, metadata and where its arrays live in
// the forest's arena (byte offsets).
struct TreeSlot {
  uint64_t n = 0, sigma = 0;
  uint32_t L = 0;
  const TreeShape* shape = nullptr;
  TreeMeta meta;
  std::vector<uint64_t> level_size, level_word, level_nwords;  // bits, word offset, words
  std::vector<uint64_t> num_l1, num_l2, cap, l1_at, l2_at, smp_at;
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

}  // namespace

struct ecl_graph_t {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~ecl_graph_t() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct ecl_forest_t {
  int device = 0;
  std::map<uint64_t, TreeShape> shapes;
  std::vector<TreeSlot> trees;
  DevBuf arena;                  // bits, directories and metadata of every tree
  DevBuf d_views;                // TreeView per tree
  std::vector<TreeView> views;
  uint32_t q_block = 256, q_grid = 0;  // launch shape of the batch query kernel
  std::vector<LevelView> levels;       // every tree's level views, tree by tree
  std::vector<LevelView*> d_levels;    // where each of them lives in the arena
  DevBuf lines;                        // query lines and select hints (opt-in)
};

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
    uint64_t total = 0;
    uint64_t span = 0;  // bytes of d_texts the texts reach
    for (size_t i = 0; i < N; ++i) {
      require(starts[i] % 16 == 0, "every text must start on a 16-byte boundary");
      require(lengths[i] > 0, "every text must be non-empty");
      span = std::max(span, starts[i] + lengths[i]);
      require(sigmas[i] >= 2 && sigmas[i] <= 256, "alphabet sizes must be in [2, 256]");
      total += lengths[i];
      if (!f->shapes.count(sigmas[i])) f->shapes.emplace(sigmas[i], tree_shape(sigmas[i]));
    }
    // ---- tiles: the largest tile whose padding stays within 1/4 of the text
    uint32_t rows = tile_rows;
    if (rows == 0) {
      rows = 1;
      for (uint32_t r : {4u, 2u}) {
        const uint64_t T = pass_tile_codes(1, r);
        uint64_t padded = 0;
        for (size_t i = 0; i < N; ++i) padded += (lengths[i] + T - 1) / T * T;
        if (padded * 4 <= total * 5) {
          rows = r;
          break;
        }
      }
    }
    const uint64_t T = pass_tile_codes(1, rows);
    std::vector<uint64_t> tstart;
    std::vector<uint32_t> tlen, tnode, tfirst(N + 1);
    for (size_t i = 0; i < N; ++i) {
      tfirst[i] = static_cast<uint32_t>(tstart.size());
      const uint64_t n = lengths[i];
      for (uint64_t o = 0; o < n; o += T) {
        tstart.push_back(starts[i] + o);
        tlen.push_back(static_cast<uint32_t>(std::min<uint64_t>(T, n - o)));
        tnode.push_back(static_cast<uint32_t>(i));
      }
    }
    const uint64_t ntiles = tstart.size();
    tfirst[N] = static_cast<uint32_t>(ntiles);
    require(ntiles < (1ull << 31), "too many construction tiles");
    // ---- non-power-of-2 alphabets read mapped codes (same positions)
    bool any_remap = false;
    std::vector<uint8_t> remap(N, 0), lut(N * 256, 0);
    for (size_t i = 0; i < N; ++i) {
      const TreeShape& sh = f->shapes.at(sigmas[i]);
      if (sh.pow2) continue;
      remap[i] = 1;
      any_remap = true;
      for (uint32_t x = 0; x < 256; ++x)
        lut[i * 256 + x] = static_cast<uint8_t>(sh.code[std::min<uint64_t>(x, sh.sigma - 1)]);
    }
    DevBuf d_x(any_remap ? span : 0);
    DevBuf d_tstart(ntiles * 8), d_tlen(ntiles * 4), d_tnode(ntiles * 4), d_tfirst((N + 1) * 4);
    DevBuf d_remap(N), d_lut(N * 256);
    h2d(d_tstart.ptr, tstart.data(), ntiles);
    h2d(d_tlen.ptr, tlen.data(), ntiles);
    h2d(d_tnode.ptr, tnode.data(), ntiles);
    h2d(d_tfirst.ptr, tfirst.data(), N + 1);
    h2d(d_remap.ptr, remap.data(), N);
    h2d(d_lut.ptr, lut.data(), N * 256);
    constexpr uint32_t kRow = 256;
    DevBuf d_hist(ntiles * kRow * 4), d_excl((ntiles + 1) * kRow * 8);
    DevBuf d_partial(scan_partial_entries(ntiles, 8) * 8);
    DevBuf d_err(4);
    launch_fill_u32(d_err.as<uint32_t>(), 1, 0, s);

    PassParams pp{};
    pp.src = d_texts;
    pp.src_esize = 1;
    pp.codes_out = any_remap ? d_x.ptr : nullptr;
    pp.map_kind = 0;
    pp.sigma = 256;
    pp.pow2 = 1;
    pp.codes = d_texts;
    pp.cw = 1;
    pp.k = 8;
    pp.dshift = 0;
    pp.tile_rows = rows;
    pp.tile_codes = static_cast<uint32_t>(T);
    pp.num_tiles = ntiles;
    pp.tile_start = d_tstart.as<uint64_t>();
    pp.tile_len = d_tlen.as<uint32_t>();
    pp.tile_node = d_tnode.as<uint32_t>();
    pp.node_first_tile = d_tfirst.as<uint32_t>();
    pp.hist = d_hist.as<uint32_t>();
    pp.excl = d_excl.as<uint64_t>();
    pp.partial = d_partial.as<uint64_t>();
    pp.err = d_err.as<uint32_t>();
    pp.node_remap = any_remap ? d_remap.as<uint8_t>() : nullptr;
    pp.node_lut = d_lut.as<uint8_t>();
    pp.row_bins = kRow;
    pp.thr_stride = 9;
    launch_tile_hist(pp, s);
    launch_tile_scan(pp, s);
    // digit (= code) counts of every text: scan rows at the text boundaries
    std::vector<uint64_t> rows_h((N + 1) * kRow);
    {
      DevBuf d_rows((N + 1) * kRow * 8);
      launch_gather_rows(d_excl.as<uint64_t>(), d_tfirst.as<uint32_t>(), N + 1, kRow,
                         d_rows.as<uint64_t>(), s);
      ECL_HIP(hipStreamSynchronize(s));
      d2h(rows_h.data(), d_rows.ptr, rows_h.size());
    }

    // ---- per tree: counts, metadata, level sizes, arena layout
    f->trees.resize(N);
    uint64_t at = 0;  // bits and directories first, then the metadata block
    for (size_t i = 0; i < N; ++i) {
      TreeSlot& t = f->trees[i];
      const TreeShape& sh = f->shapes.at(sigmas[i]);
      t.n = lengths[i];
      t.sigma = sh.sigma;
      t.L = sh.L;
      t.shape = &sh;
      std::vector<uint64_t> sym(sh.sigma);
      for (uint64_t c = 0; c < sh.sigma; ++c)
        sym[c] = rows_h[(i + 1) * kRow + sh.code[c]] - rows_h[i * kRow + sh.code[c]];
      t.meta = tree_meta(sh, sym, t.n);
      t.level_size = level_sizes(sh, sym);
      // bit array: each level padded to 32 words (bit_array.cu:57-72)
      at = align_up(at, kAlign);
      uint64_t w = 0;
      t.level_word.resize(t.L);
      t.level_nwords.resize(t.L);
      for (uint32_t l = 0; l < t.L; ++l) {
        t.level_word[l] = at / 4 + w;
        t.level_nwords[l] = (t.level_size[l] + 31) / 32;
        w += (t.level_nwords[l] + 31) & ~31ull;
      }
      w += w % 2;
      at += std::max<uint64_t>(w, 2) * 4;
      // rank/select directory per level (rank_select.cu:57-143 sizes)
      t.num_l1.resize(t.L); t.num_l2.resize(t.L); t.cap.resize(t.L);
      t.l1_at.resize(t.L); t.l2_at.resize(t.L); t.smp_at.resize(t.L);
      for (uint32_t l = 0; l < t.L; ++l) {
        const uint64_t m = t.level_size[l];
        t.num_l1[l] = (m + detail::kL1Bits - 1) / detail::kL1Bits;
        uint64_t last = ((m % detail::kL1Bits) + detail::kL2Bits - 1) / detail::kL2Bits;
        if (last == 0 && m) last = detail::kL2PerL1;
        t.num_l2[l] = m ? (t.num_l1[l] - 1) * detail::kL2PerL1 + last : 0;
        t.cap[l] = m / detail::kSampleRate;
        at = align_up(at, 8);
        t.l1_at[l] = at; at += t.num_l1[l] * 8;
        t.smp_at[l] = at; at += t.cap[l] * 8;
        t.l2_at[l] = at; at += t.num_l2[l] * 2;
      }
    }
    // metadata block: total ones, counts, ranks, nodes per level, codes,
    // code lengths and level views of every tree, uploaded in one copy
    at = align_up(at, kAlign);
    const uint64_t meta_at = at;
    for (auto& t : f->trees) { t.ones_at = at; at += t.L * 8ull; }
    const uint64_t ones_end = at;
    for (auto& t : f->trees) {
      t.counts_at = at; at += t.sigma * 8;
      t.ranks_at = at; at += t.meta.ranks.size() * 8;
      t.nal_at = at; at += t.meta.nodes_at_level.size() * 8;
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
    // bits are OR-ed in at run edges and the totals accumulate: zero all
    launch_fill_u64(f->arena.as<uint64_t>(), arena_bytes / 8, 0, s);

    // ---- the level pass of every tree
    {
      std::vector<uint64_t> pc(N * (kRow + 1)), lw(N * 8, 0), lnw(N * 8, 0);
      std::vector<uint16_t> thr(N * 9, 0);
      std::vector<uint8_t> nk(N);
      for (size_t i = 0; i < N; ++i) {
        const TreeSlot& t = f->trees[i];
        const TreeShape& sh = *t.shape;
        const uint32_t L = t.L, nb = 1u << L;
        nk[i] = static_cast<uint8_t>(L);
        uint64_t run = 0;
        for (uint32_t d = 0; d <= kRow; ++d) {
          pc[i * (kRow + 1) + d] = run;
          if (d < kRow) run += rows_h[(i + 1) * kRow + d] - rows_h[i * kRow + d];
        }
        for (uint32_t l = 0; l < L; ++l) {
          lw[i * 8 + l] = t.level_word[l];
          lnw[i * 8 + l] = t.level_nwords[l];
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
      DevBuf d_pc(pc.size() * 8), d_thr(thr.size() * 2), d_nk(N), d_lw(N * 64), d_lnw(N * 64);
      h2d(d_pc.ptr, pc.data(), pc.size());
      h2d(d_thr.ptr, thr.data(), thr.size());
      h2d(d_nk.ptr, nk.data(), N);
      h2d(d_lw.ptr, lw.data(), lw.size());
      h2d(d_lnw.ptr, lnw.data(), lnw.size());
      pp.pc = d_pc.as<uint64_t>();
      pp.thr = d_thr.as<uint16_t>();
      pp.node_k = d_nk.as<uint8_t>();
      pp.node_lw = d_lw.as<uint64_t>();
      pp.node_lnw = d_lnw.as<uint64_t>();
      pp.words = f->arena.as<uint32_t>();
      pp.dst = nullptr;
      launch_pass(pp, s);
      uint32_t err = 0;
      ECL_HIP(hipStreamSynchronize(s));
      d2h(&err, d_err.ptr, 1);
      if (err) throw Error(ECL_ERROR_INTERNAL, "construction pass wrote outside its arrays");
    }
    d_x.reset();
    d_hist.reset();
    d_excl.reset();

    // ---- directories of every level of every tree, one launch
    {
      std::vector<RsArrayDesc> desc;
      uint64_t blocks = 0;
      for (auto& t : f->trees) {
        for (uint32_t l = 0; l < t.L; ++l) {
          if (t.level_size[l] == 0) continue;
          RsArrayDesc d{};
          d.words64 = reinterpret_cast<const uint64_t*>(A + t.level_word[l] * 4);
          d.m = t.level_size[l];
          d.l1 = reinterpret_cast<uint64_t*>(A + t.l1_at[l]);
          d.l2 = reinterpret_cast<uint16_t*>(A + t.l2_at[l]);
          d.num_l1 = t.num_l1[l];
          d.num_l2 = t.num_l2[l];
          d.samples = reinterpret_cast<uint64_t*>(A + t.smp_at[l]);
          d.cap = t.cap[l];
          d.total_ones = reinterpret_cast<uint64_t*>(A + t.ones_at) + l;
          d.first_tile = blocks;
          blocks += d.num_l1;
          desc.push_back(d);
        }
      }
      DevBuf d_desc(std::max<size_t>(desc.size(), 1) * sizeof(RsArrayDesc));
      DevBuf d_blk(std::max<uint64_t>(blocks, 1) * 8);
      h2d(d_desc.ptr, desc.data(), desc.size());
      launch_rs_build(d_desc.as<RsArrayDesc>(), desc.size(), blocks, d_blk.as<uint32_t>(), s);
      ECL_HIP(hipStreamSynchronize(s));
    }

    // ---- metadata block and views
    std::vector<uint8_t> img(arena_bytes - meta_at, 0);
    d2h(img.data(), A + meta_at, ones_end - meta_at);  // total ones per level
    auto put = [&](uint64_t where, const void* src, size_t bytes) {
      if (bytes) std::memcpy(img.data() + (where - meta_at), src, bytes);
    };
    f->views.resize(N);
    for (size_t i = 0; i < N; ++i) {
      TreeSlot& t = f->trees[i];
      const TreeShape& sh = *t.shape;
      put(t.counts_at, t.meta.counts.data(), t.sigma * 8);
      put(t.ranks_at, t.meta.ranks.data(), t.meta.ranks.size() * 8);
      put(t.nal_at, t.meta.nodes_at_level.data(), t.meta.nodes_at_level.size() * 8);
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
      std::vector<LevelView> lv(t.L);
      const uint64_t* ones = reinterpret_cast<const uint64_t*>(img.data() + (t.ones_at - meta_at));
      for (uint32_t l = 0; l < t.L; ++l) {
        LevelView& v = lv[l];
        v.words = reinterpret_cast<const uint32_t*>(A + t.level_word[l] * 4);
        v.l1 = reinterpret_cast<const uint64_t*>(A + t.l1_at[l]);
        v.l2 = reinterpret_cast<const uint16_t*>(A + t.l2_at[l]);
        v.samples = reinterpret_cast<const uint64_t*>(A + t.smp_at[l]);
        v.size = t.level_size[l];
        v.ones = ones[l];
        v.num_l1 = t.num_l1[l];
        v.sample_cap = t.cap[l];
      }
      put(t.lv_at, lv.data(), lv.size() * sizeof(LevelView));
      for (uint32_t l = 0; l < t.L; ++l) {
        f->levels.push_back(lv[l]);
        f->d_levels.push_back(reinterpret_cast<LevelView*>(A + t.lv_at) + l);
      }
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
      v.nodes_until_last_level = t.meta.nodes_until_last_level;
      v.num_ranks = t.meta.ranks.size();
      v.num_levels = t.L;
      v.pow2 = sh.pow2 ? 1 : 0;
    }
    h2d(A + meta_at, img.data(), img.size());
    f->d_views = DevBuf(N * sizeof(TreeView));
    h2d(f->d_views.ptr, f->views.data(), N);
    f->q_grid = static_cast<uint32_t>(cu_count(device)) * 8;
    *out = f.release();
  });
}

ecl_status ecl_forest_build_query_lines(ecl_forest f) {
  return guard([&] {
    require(f, "NULL argument");
    if (f->lines.ptr != nullptr) return;
    set_device(f->device);
    const size_t na = f->levels.size();
    require(na < (1ull << 32), "too many bit arrays");
    std::vector<LinesJob> jobs(na);
    std::vector<uint64_t> starts(2 * na);  // line starts | hint starts
    uint64_t tl = 0, th = 0;
    for (size_t a = 0; a < na; ++a) {
      const LevelView& v = f->levels[a];
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
    ECL_HIP(hipMemsetAsync(d_fit.ptr, 0, std::max<size_t>(na, 1) * 8, s));
    uint4* L = lines.as<uint4>();
    uint4* H = L + tl * 8;
    launch_multi_lines(d_jobs.as<LinesJob>(), d_starts.as<uint64_t>(), d_starts.as<uint64_t>() + na,
                       static_cast<uint32_t>(na), L, tl, H, th, d_fit.as<uint32_t>(), s);
    std::vector<uint32_t> fit(2 * na);
    if (na) d2h(fit.data(), d_fit.ptr, 2 * na);  // waits for the launches
    for (size_t a = 0; a < na; ++a) {
      LevelView& v = f->levels[a];
      v.lines = jobs[a].nl ? L + starts[a] * 8 : nullptr;
      v.sel_hints = jobs[a].nh1 + jobs[a].nh0 ? H + starts[na + a] * 4 : nullptr;
      v.hint_values = (2ull * fit[2 * a + 1] >= jobs[a].nh1 ? 2u : 0u) |
                      (2ull * fit[2 * a] >= jobs[a].nh0 ? 1u : 0u);
    }
    f->lines = std::move(lines);
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
    launch_forest_query(f->d_views.as<TreeView>(), f->trees.size(), d_queries, m, d_out,
                        f->q_block, f->q_grid, static_cast<hipStream_t>(stream));
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
    hipEvent_t e0, e1;
    ECL_HIP(hipEventCreate(&e0));
    ECL_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    uint32_t bb = f->q_block, bg = f->q_grid;
    try {
      for (uint32_t b : {64u, 128u, 256u, 512u}) {
        for (uint32_t per_cu : {1u, 2u, 4u, 8u, 16u, 32u}) {
          const uint32_t g = cus * per_cu * 256 / b;  // same waves per CU across block sizes
          launch_forest_query(f->d_views.as<TreeView>(), f->trees.size(), d_queries, sample, d_out,
                              b, g, nullptr);
          ECL_HIP(hipEventRecord(e0, nullptr));
          for (int r = 0; r < 3; ++r)
            launch_forest_query(f->d_views.as<TreeView>(), f->trees.size(), d_queries, sample,
                                d_out, b, g, nullptr);
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
    f->q_block = bb;
    f->q_grid = bg;
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
      launch_forest_query(f->d_views.as<TreeView>(), f->trees.size(), d_queries, m, d_out,
                          f->q_block, f->q_grid, s);
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
