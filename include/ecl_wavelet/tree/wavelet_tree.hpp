// ecl_wavelet/tree/wavelet_tree.hpp -- drop-in for tree/wavelet_tree.cuh
// (mfranzreb/ECL_wavelet) on HIP / gfx950.
//
// Construction and host batch queries go through the C ABI of
// libecl_wavelet_hip.so.  The object is the reference's shallow view: copies
// (e.g. kernel arguments) share the device tree and never free it, and the
// device member functions deviceAccess / deviceRank / deviceSelect run the
// library's descents inline in user kernels.
//
// Differences kept deliberately small:
//  * errors other than out-of-memory throw std::runtime_error instead of
//    printing and calling exit();
//  * ThreadsPerQuery may be any power of two dividing the wave width (64);
//    a group is that many consecutive lanes of one wave, all active and
//    calling with the same query; its lanes split each level's rank or
//    select and all return the answer;
//  * the `offsets` LDS argument of deviceAccess / deviceRank is accepted and
//    ignored (level descriptors carry pointers, not word offsets);
//  * when the tree carries hex or quad lines (built through the C ABI on
//    handle(), e.g. ecl_wavelet_tree_build_hex_lines), deviceAccess and
//    deviceRank with ThreadsPerQuery >= 2 read them, and deviceSelect reads
//    the quad select lines: one line per four or two levels, each pair of
//    adjacent lanes of a group running the descent.  Answers do not change;
//    with one thread per query the binary levels are read.
#pragma once

#include <algorithm>
#include <limits>
#include <numeric>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "ecl_wavelet/bitarray/rank_select.hpp"
#include "ecl_wavelet/detail/quad_core.hpp"

namespace ecl {

// pinned host allocator behind PinnedVector (wavelet_tree.cuh:59-62)
template <typename T>
struct PinnedAllocator {
  using value_type = T;
  PinnedAllocator() = default;
  template <typename U>
  PinnedAllocator(const PinnedAllocator<U>&) {}
  T* allocate(size_t n) {
    void* p = nullptr;
    utils::hipThrow(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocPortable),
                    "hipHostMalloc");
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t) { (void)hipHostFree(p); }
  template <typename U>
  bool operator==(const PinnedAllocator<U>&) const { return true; }
  template <typename U>
  bool operator!=(const PinnedAllocator<U>&) const { return false; }
};
template <typename T>
using PinnedVector = std::vector<T, PinnedAllocator<T>>;

template <typename T>
class WaveletTree {
  static_assert(std::is_integral<T>::value && std::is_unsigned<T>::value,
                "T must be an unsigned integral type");
  static constexpr ecl_type kType = static_cast<ecl_type>(sizeof(T));

 public:
  struct Code {
    uint8_t len_;
    T code_;
  };
  RankSelect rank_select_;

  WaveletTree() = default;

  // WaveletTree(data, data_size, alphabet&&, GPU_index) (wavelet_tree.cuh:
  // 576-867).  An empty alphabet is discovered from the data.
  __host__ WaveletTree(T const* data, size_t data_size, std::vector<T>&& alphabet,
                       uint32_t const GPU_index) {
    utils::hipThrow(hipSetDevice(static_cast<int>(GPU_index)), "hipSetDevice");
    utils::throwOnError(ecl_wavelet_tree_create(kType, data, data_size, alphabet.data(),
                                                alphabet.size(), static_cast<int>(GPU_index),
                                                &handle_));
    is_copy_ = false;
    host_ = new Host();
    host_->gpu = GPU_index;
    bind();
  }
  // A tree read from a file written by save() (no counterpart in the
  // reference); its element size must be sizeof(T).
  struct FromFile {};
  __host__ WaveletTree(FromFile, std::string const& path, uint32_t const GPU_index) {
    utils::hipThrow(hipSetDevice(static_cast<int>(GPU_index)), "hipSetDevice");
    utils::throwOnError(ecl_wavelet_tree_load(path.c_str(), static_cast<int>(GPU_index),
                                              &handle_));
    is_copy_ = false;
    host_ = new Host();
    host_->gpu = GPU_index;
    bind();
    if (info_.element_size != sizeof(T)) {
      release();
      throw std::runtime_error("the tree file holds another element type");
    }
  }
  // writes the tree (alphabet, symbol counts, level bits) to a file
  __host__ void save(std::string const& path) const {
    utils::throwOnError(ecl_wavelet_tree_save(handle_, path.c_str()));
  }
  // copy = shallow view (wavelet_tree.cuh:870-887)
  __host__ WaveletTree(WaveletTree const& o)
      : rank_select_(o.rank_select_), handle_(o.handle_), host_(o.host_), view_(o.view_),
        quad_(o.quad_), hex_(o.hex_), info_(o.info_), is_copy_(true) {}
  WaveletTree& operator=(const WaveletTree&) = delete;
  WaveletTree(WaveletTree&&) = delete;
  __host__ WaveletTree& operator=(WaveletTree&& o) {
    release();
    rank_select_ = std::move(o.rank_select_);
    handle_ = o.handle_; host_ = o.host_; view_ = o.view_; info_ = o.info_;
    quad_ = o.quad_; hex_ = o.hex_;
    is_copy_ = o.is_copy_;
    o.is_copy_ = true;
    return *this;
  }
  __host__ ~WaveletTree() { release(); }

  // ---- host batch queries ---------------------------------------------------
  // access (wavelet_tree.cuh:956-1231): a span into a pinned pool that the
  // next access call overwrites.
  __host__ std::span<T> access(size_t* indices, size_t const num_indices) {
    auto& pool = host_->access_pool;
    if (pool.size() < num_indices) pool.resize(num_indices);
    utils::throwOnError(ecl_wavelet_tree_access(handle_, reinterpret_cast<uint64_t*>(indices),
                                                num_indices, pool.data(), 0));
    return std::span<T>(pool.data(), num_indices);
  }
  // rank (wavelet_tree.cuh:1247-1551): symbols are rewritten in place to
  // their minimal-alphabet index, as the reference does.
  __host__ std::span<size_t> rank(RankSelectQuery<T>* queries, size_t const num_queries) {
    to_minimal(queries, num_queries);
    auto& pool = host_->rank_pool;
    if (pool.size() < num_queries) pool.resize(num_queries);
    utils::throwOnError(ecl_wavelet_tree_rank(
        handle_, reinterpret_cast<const ecl_rank_select_query*>(queries), num_queries,
        reinterpret_cast<uint64_t*>(pool.data()), ECL_SYMBOLS_MINIMAL));
    return std::span<size_t>(pool.data(), num_queries);
  }
  // select (wavelet_tree.cuh:1567-1797)
  __host__ std::span<size_t> select(RankSelectQuery<T>* queries, size_t const num_queries) {
    to_minimal(queries, num_queries);
    auto& pool = host_->select_pool;
    if (pool.size() < num_queries) pool.resize(num_queries);
    utils::throwOnError(ecl_wavelet_tree_select(
        handle_, reinterpret_cast<const ecl_rank_select_query*>(queries), num_queries,
        reinterpret_cast<uint64_t*>(pool.data()), ECL_SYMBOLS_MINIMAL));
    return std::span<size_t>(pool.data(), num_queries);
  }

  // ---- device API (wavelet_tree.cuh:1815-2079) --------------------------------
  template <int ThreadsPerQuery>
  __device__ T deviceAccess(size_t index, bool const /*is_pow_two*/, size_t* counts = nullptr,
                            size_t* /*offsets*/ = nullptr, size_t* ranks = nullptr) {
    check_tpq<ThreadsPerQuery>();
    if constexpr (ThreadsPerQuery >= 2) {
      if (hex_->levels) {
        const detail::LayoutRef p{view_, *quad_, *hex_};
        return static_cast<T>(detail::hex_access(p, detail::QuadTable<false>{hex_->tab}, index));
      }
      if (quad_->levels) {
        const detail::LayoutRef p{view_, *quad_, *hex_};
        return static_cast<T>(detail::quad_access(p, detail::QuadTable<false>{quad_->base}, index));
      }
    }
    return static_cast<T>(detail::tree_access<false, ThreadsPerQuery>(
        view_, view_.levels, counts_or(counts), ranks_or(ranks), view_.nodes_at_level, index));
  }
  template <int ThreadsPerQuery>
  __device__ size_t deviceRank(RankSelectQuery<T> const& query, bool const /*is_pow_two*/,
                               size_t* counts = nullptr, size_t* /*offsets*/ = nullptr,
                               size_t* ranks = nullptr) {
    check_tpq<ThreadsPerQuery>();
    if constexpr (ThreadsPerQuery >= 2) {
      if (hex_->levels) {
        const detail::LayoutRef p{view_, *quad_, *hex_};
        return detail::hex_rank(p, detail::QuadTable<false>{hex_->tab}, query.symbol_,
                                query.index_);
      }
      if (quad_->levels) {
        const detail::LayoutRef p{view_, *quad_, *hex_};
        return detail::quad_rank(p, detail::QuadTable<false>{quad_->base}, query.symbol_,
                                 query.index_);
      }
    }
    return detail::tree_rank<false, ThreadsPerQuery>(view_, view_.levels, counts_or(counts),
                                                     ranks_or(ranks), view_.nodes_at_level,
                                                     query.symbol_, query.index_);
  }
  template <int ThreadsPerQuery>
  __device__ size_t deviceSelect(RankSelectQuery<T> const& query, bool const /*is_pow_two*/,
                                 size_t* ranks = nullptr) {
    check_tpq<ThreadsPerQuery>();
    if constexpr (ThreadsPerQuery >= 2) {
      if (quad_->has_sel) {
        const detail::LayoutRef p{view_, *quad_, *hex_};
        return detail::quad_select(p, detail::QuadTable<false>{quad_->tab}, query.symbol_,
                                   query.index_);
      }
    }
    return detail::tree_select<true, true, ThreadsPerQuery>(view_, view_.levels, view_.counts,
                                                            ranks_or(ranks), view_.nodes_at_level,
                                                            query.symbol_, query.index_);
  }

  // encode (wavelet_tree.cuh:2087-2091)
  __device__ Code encode(T const c) const {
    return Code{view_.code_lens[c - view_.codes_start],
                static_cast<T>(view_.codes[c - view_.codes_start])};
  }

  // createMinimalCodes (wavelet_tree.cuh:2100-2159): codes of the symbols
  // from the codes start on, as leaves of the left-complete binary tree
  // (a node of s symbols splits after the largest power of two below s).
  __host__ static std::vector<Code> createMinimalCodes(std::vector<T> const& alphabet) {
    const uint64_t sigma = alphabet.size();
    if (utils::isPowTwo<uint64_t>(sigma)) return {};
    const uint32_t L = detail::ceil_log2(sigma);
    uint64_t num = sigma;  // symbols past the first complete subtrees
    for (uint32_t i = L - 1; i > 0; --i) {
      const uint64_t p = 1ull << i;
      if (num <= p) break;
      num -= p;
    }
    std::vector<Code> out;
    out.reserve(num);
    const uint64_t first = sigma - num;
    // walk the left-complete tree down to every leaf >= first
    struct Frame { uint64_t lo, hi, prefix; uint32_t depth; };
    std::vector<Frame> stack{{0, sigma - 1, 0, 0}};
    std::vector<Code> codes(sigma);
    while (!stack.empty()) {
      Frame f = stack.back();
      stack.pop_back();
      if (f.hi < first) continue;
      if (f.lo == f.hi) {
        codes[f.lo] = Code{static_cast<uint8_t>(f.depth),
                           static_cast<T>(f.prefix << (L - f.depth))};
        continue;
      }
      const uint64_t mid = f.lo + detail::prev_pow2(f.hi - f.lo + 1);
      stack.push_back({mid, f.hi, (f.prefix << 1) | 1, f.depth + 1});
      stack.push_back({f.lo, mid - 1, f.prefix << 1, f.depth + 1});
    }
    for (uint64_t c = first; c < sigma; ++c) out.push_back(codes[c]);
    return out;
  }

  // ---- accessors (wavelet_tree.cuh:2165-2282) ----------------------------------
  __host__ __device__ size_t getAlphabetSize() const { return view_.sigma; }
  __host__ std::vector<T> const& getAlphabet() const { return host_->alphabet; }
  __device__ uint8_t getNumLevels() const { return static_cast<uint8_t>(view_.num_levels); }
  __host__ __device__ size_t getCounts(size_t i) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return view_.counts[i];
#else
    return host_->counts[i];
#endif
  }
  __host__ __device__ size_t getTotalAppearances(size_t i) const {
    if (i == getAlphabetSize() - 1) return view_.n - getCounts(i);
    return getCounts(i + 1) - getCounts(i);
  }
  __host__ __device__ bool isMinAlphabet() const { return info_.is_min_alphabet != 0; }
  __device__ T getCodesStart() const { return static_cast<T>(view_.codes_start); }
  template <bool IsPowTwo>
  __device__ T getNodePosAtLevel(T symbol, uint8_t const level) const {
    return static_cast<T>(detail::tree_node_pos(view_, symbol, level));
  }
  __device__ size_t getPrecomputedRank(T const i) const { return view_.ranks[i]; }
  __device__ void setPrecomputedRank(T const i, size_t const r) {
    const_cast<uint64_t*>(view_.ranks)[i] = r;
  }
  __device__ T getNumNodesAtLevel(uint8_t const level) const {
    return static_cast<T>(detail::tree_nodes_at(view_, view_.nodes_at_level, level));
  }
  // getNeededGPUMemory (wavelet_tree.cuh:2544-2559)
  __host__ static size_t getNeededGPUMemory(size_t const data_size, size_t const alphabet_size,
                                            uint8_t const num_levels, T const num_codes) {
    return ecl_wavelet_tree_needed_gpu_memory(data_size, alphabet_size, num_levels, num_codes);
  }

  __host__ ecl_wavelet_tree handle() const { return handle_; }

 private:
  struct Host {
    uint32_t gpu = 0;
    std::vector<T> alphabet;
    std::vector<size_t> counts;
    PinnedVector<T> access_pool;
    PinnedVector<size_t> rank_pool, select_pool;
  };

  template <int TPQ>
  __device__ static constexpr void check_tpq() {
    static_assert(TPQ > 0 && TPQ <= WS && (WS % TPQ) == 0,
                  "ThreadsPerQuery must divide the wave width");
  }
  __device__ const uint64_t* counts_or(size_t* c) const {
    return c ? reinterpret_cast<const uint64_t*>(c) : view_.counts;
  }
  __device__ const uint64_t* ranks_or(size_t* r) const {
    return r ? reinterpret_cast<const uint64_t*>(r) : view_.ranks;
  }

  __host__ void bind() {
    utils::throwOnError(ecl_wavelet_tree_info_get(handle_, &info_));
    utils::throwOnError(ecl_wavelet_tree_device_view(handle_, &view_, sizeof(view_)));
    const void *qv = nullptr, *hv = nullptr;
    utils::throwOnError(ecl_wavelet_tree_layout_views(handle_, &qv, &hv));
    quad_ = static_cast<const detail::QuadView*>(qv);
    hex_ = static_cast<const detail::HexView*>(hv);
    host_->alphabet.resize(info_.alphabet_size);
    utils::throwOnError(ecl_wavelet_tree_alphabet(handle_, host_->alphabet.data()));
    host_->counts.resize(info_.alphabet_size);
    utils::throwOnError(ecl_wavelet_tree_counts(
        handle_, reinterpret_cast<uint64_t*>(host_->counts.data())));
    ecl_rank_select rs = nullptr;
    utils::throwOnError(ecl_wavelet_tree_rank_select(handle_, &rs));
    rank_select_.adopt(rs);
  }
  __host__ void release() {
    if (!is_copy_) {
      if (handle_ != nullptr) ecl_wavelet_tree_destroy(handle_);
      delete host_;
    }
    handle_ = nullptr;
    host_ = nullptr;
    is_copy_ = true;
  }
  // in-place symbol -> minimal index (lower_bound, wavelet_tree.cuh:1426-1436)
  __host__ void to_minimal(RankSelectQuery<T>* q, size_t m) {
    if (info_.is_min_alphabet) return;
    const auto& a = host_->alphabet;
    for (size_t i = 0; i < m; ++i)
      q[i].symbol_ = static_cast<T>(std::lower_bound(a.begin(), a.end(), q[i].symbol_) - a.begin());
  }

  ecl_wavelet_tree handle_ = nullptr;
  Host* host_ = nullptr;
  detail::TreeView view_{};
  const detail::QuadView* quad_ = nullptr;  // device copies, kept current by the library
  const detail::HexView* hex_ = nullptr;
  ecl_wavelet_tree_info info_{};
  bool is_copy_ = true;
};

}  // namespace ecl
