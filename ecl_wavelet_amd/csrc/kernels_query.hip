// kernels_query.hip -- batched query kernels for gfx950.
//
// One lane per query, grid-stride.  Per-block LDS caches the small, hot
// metadata every descent touches (level descriptors, node counts per level,
// and, when they fit, the C array and the node ranks), like the reference's
// shmem variants of accessKernel / rankKernel / selectKernel
// (tree/wavelet_tree.cuh:373-544).  The random bit-array lines are the only
// HBM traffic left per level.
#include <hip/hip_runtime.h>

#include <atomic>

#include "internal.hpp"

namespace ecl {
namespace lib {
namespace {

using detail::PairLine;
using detail::QuadStep;
using detail::QuadTable;
using detail::hex_access;
using detail::hex_rank;
using detail::hex_step;
using detail::pair_line;
using detail::pair_select_in_line;
using detail::pair_swap;
using detail::pair_swap64;
using detail::quad_access;
using detail::quad_rank;
using detail::quad_select;
using detail::quad_select_line;
using detail::quad_select_walk;
using detail::quad_step;

using detail::LevelView;
using detail::TreeView;

constexpr uint64_t kMaxLdsTable = 2048;  // u64 entries for C and for ranks
constexpr uint32_t kMaxLevels = 64;

struct Meta {
  const LevelView* lv;
  const uint64_t* counts;
  const uint64_t* ranks;
  const uint64_t* nal;
};

__host__ __device__ __forceinline__ bool lds_counts(const TreeView& t) {
  return t.sigma <= kMaxLdsTable;
}
__host__ __device__ __forceinline__ bool lds_ranks(const TreeView& t) {
  return t.num_ranks <= kMaxLdsTable;
}
// LDS layout: levels[L] | nal[L] | counts[sigma]? | ranks[R]?
__host__ __device__ __forceinline__ size_t meta_bytes_of(const TreeView& t) {
  size_t b = t.num_levels * (sizeof(LevelView) + sizeof(uint64_t));
  if (lds_counts(t)) b += t.sigma * sizeof(uint64_t);
  if (lds_ranks(t)) b += t.num_ranks * sizeof(uint64_t);
  return b;
}

__device__ Meta stage_meta(const TreeView& t, uint8_t* lds) {
  const uint32_t L = t.num_levels;
  LevelView* s_lv = reinterpret_cast<LevelView*>(lds);
  uint64_t* s_nal = reinterpret_cast<uint64_t*>(lds + L * sizeof(LevelView));
  uint64_t* s_next = s_nal + L;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < L; i += blockDim.x) s_lv[i] = t.levels[i];
  for (uint32_t i = tid; i < L; i += blockDim.x)
    s_nal[i] = (t.pow2 || i + 1 >= L) ? 0 : t.nodes_at_level[i];
  Meta m{s_lv, t.counts, t.ranks, s_nal};
  if (lds_counts(t)) {
    for (uint64_t i = tid; i < t.sigma; i += blockDim.x) s_next[i] = t.counts[i];
    m.counts = s_next;
    s_next += t.sigma;
  }
  if (lds_ranks(t)) {
    for (uint64_t i = tid; i < t.num_ranks; i += blockDim.x) s_next[i] = t.ranks[i];
    m.ranks = s_next;
  }
  __syncthreads();
  return m;
}

__device__ __forceinline__ uint64_t load_sym(const void* q16, int esize) {
  // symbol of a 16-byte RankSelectQuery<T>: first sizeof(T) bytes at +8
  const uint8_t* p = static_cast<const uint8_t*>(q16) + 8;
  switch (esize) {
    case 1: return *p;
    case 2: return *reinterpret_cast<const uint16_t*>(p);
    case 4: return *reinterpret_cast<const uint32_t*>(p);
    default: return *reinterpret_cast<const uint64_t*>(p);
  }
}

template <typename T>
__device__ __forceinline__ uint64_t map_symbol(const TreeQueryParams& p,
                                               uint64_t x) {
  if (p.symbols_minimal || p.map_kind == 0) return x;
  if (p.map_kind == 1) {
    const uint64_t lim = sizeof(T) == 1 ? 256 : 65536;
    return x < lim ? p.lut[x] : p.view.sigma;
  }
  const T* a = static_cast<const T*>(p.alphabet);
  uint64_t lo = 0, hi = p.view.sigma;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (static_cast<uint64_t>(a[mid]) < x) lo = mid + 1; else hi = mid;
  }
  return lo;  // == sigma when x is larger than every symbol
}

// G lanes per query (1, or 2: the adjacent lanes 2k, 2k+1 descend for the
// same query, each fetching half of every level's bits; lane 2k writes).
template <typename T, int G>
__global__ __launch_bounds__(256) void tree_access_kernel(
    TreeQueryParams p, const uint64_t* __restrict__ idx, uint64_t m,
    T* __restrict__ out) {
  extern __shared__ uint8_t s_meta[];
  const Meta mt = stage_meta(p.view, s_meta);
  const T* amap = static_cast<const T*>(p.alphabet_out);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / G;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / G;
       q < m; q += stride) {
    const uint64_t i = idx[q];
    if (i >= p.view.n) continue;  // accessKernel skips such indices
    const uint64_t r =
        detail::tree_access<G == 2>(p.view, mt.lv, mt.counts, mt.ranks, mt.nal, i);
    if (G == 1 || (threadIdx.x & 1) == 0) out[q] = amap ? amap[r] : static_cast<T>(r);
  }
}

// Occupancy target (waves per SIMD): left to itself the compiler gives the
// rank kernel 78 VGPRs (6 waves); at 8 waves it fits in 61 without spilling
// and the pair kernel over query lines runs 7% faster (2.03 vs 2.17 ms per
// 10^7 at n = 6e9, sigma = 256).
template <typename T, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void tree_rank_kernel(
    TreeQueryParams p, const uint8_t* __restrict__ qs, uint64_t m,
    uint64_t* __restrict__ out) {
  extern __shared__ uint8_t s_meta[];
  const Meta mt = stage_meta(p.view, s_meta);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / G;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / G;
       q < m; q += stride) {
    const uint8_t* e = qs + q * 16;
    const uint64_t i = *reinterpret_cast<const uint64_t*>(e);
    const uint64_t c = map_symbol<T>(p, load_sym(e, sizeof(T)));
    const uint64_t r = i > p.view.n ? 0
                                    : detail::tree_rank<G == 2>(p.view, mt.lv, mt.counts,
                                                                mt.ranks, mt.nal, c, i);
    if (G == 1 || (threadIdx.x & 1) == 0) out[q] = r;
  }
}

// ---------------------------------------------------------------------------
// Access and rank over quad lines (see QuadView): one 128-byte line per two
// levels.  Two adjacent lanes share a query: lane 0 loads the line's header
// and digit words 0..5, lane 1 words 6..13; each counts the digit in its
// words before the offset, and one xor-shuffle adds the counts.  The bases
// sit in LDS when they fit.
constexpr uint64_t kQuadLdsBase = 4096;  // u64 entries

// Lds only when nbase <= kQuadLdsBase (the launch's LDS holds the table)
template <bool Lds>
__device__ __forceinline__ QuadTable<Lds> stage_quad_base(const QuadView& q, uint64_t* lds) {
  if constexpr (!Lds) {
    return QuadTable<Lds>{q.base};
  } else {
    for (uint64_t i = threadIdx.x; i < q.nbase; i += blockDim.x) lds[i] = q.base[i];
    __syncthreads();
    return QuadTable<Lds>{lds};
  }
}




template <typename T, bool Lds>
__global__ __launch_bounds__(256) void tree_access_quad_kernel(
    TreeQueryParams p, const uint64_t* __restrict__ idx, uint64_t m, T* __restrict__ out) {
  extern __shared__ uint64_t s_qbase[];
  const QuadTable<Lds> base = stage_quad_base<Lds>(p.quad, s_qbase);
  const T* amap = static_cast<const T*>(p.alphabet_out);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / 2;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / 2;
       q < m; q += stride) {
    const uint64_t i = idx[q];
    if (i >= p.view.n) continue;  // accessKernel skips such indices (both lanes)
    const uint64_t N = quad_access(p, base, i);
    if ((threadIdx.x & 1) == 0) out[q] = amap ? amap[N] : static_cast<T>(N);
  }
}

template <typename T, bool Lds>
__global__ __launch_bounds__(256) void tree_rank_quad_kernel(
    TreeQueryParams p, const uint8_t* __restrict__ qs, uint64_t m, uint64_t* __restrict__ out) {
  extern __shared__ uint64_t s_qbase[];
  const QuadTable<Lds> base = stage_quad_base<Lds>(p.quad, s_qbase);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / 2;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / 2;
       q < m; q += stride) {
    const uint8_t* e = qs + q * 16;
    const uint64_t i = *reinterpret_cast<const uint64_t*>(e);
    const uint64_t c = map_symbol<T>(p, load_sym(e, sizeof(T)));
    const uint64_t r = i <= p.view.n && c < p.view.sigma ? quad_rank(p, base, c, i) : 0;
    if ((threadIdx.x & 1) == 0) out[q] = r;
  }
}

// ---------------------------------------------------------------------------
// Access and rank over hex lines (see HexView): one 128-byte line per four
// levels.  Two adjacent lanes share a query: lane h loads the line's half h
// (the counts of digits 8h .. 8h + 7 and digits 64h .. 64h + 63); each
// counts the digit in its digits before the offset, the lane holding the
// digit's stored count adds it, and one pair swap adds the two parts.  The
// bases and superblock counts sit in LDS when they fit.
constexpr uint64_t kHexLdsTab = 4096;  // u64 entries

template <bool Lds>
__device__ __forceinline__ QuadTable<Lds> stage_hex_tab(const HexView& x, uint64_t* lds) {
  if constexpr (!Lds) {
    return QuadTable<Lds>{x.tab};
  } else {
    for (uint64_t i = threadIdx.x; i < x.ntab; i += blockDim.x) lds[i] = x.tab[i];
    __syncthreads();
    return QuadTable<Lds>{lds};
  }
}


template <typename T, bool Lds>
__global__ __launch_bounds__(256) void tree_access_hex_kernel(
    TreeQueryParams p, const uint64_t* __restrict__ idx, uint64_t m, T* __restrict__ out) {
  extern __shared__ uint64_t s_htab[];
  const QuadTable<Lds> tab = stage_hex_tab<Lds>(p.hex, s_htab);
  const T* amap = static_cast<const T*>(p.alphabet_out);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / 2;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / 2;
       q < m; q += stride) {
    const uint64_t i = idx[q];
    if (i >= p.view.n) continue;  // accessKernel skips such indices (both lanes)
    const uint64_t N = hex_access(p, tab, i);
    if ((threadIdx.x & 1) == 0) out[q] = amap ? amap[N] : static_cast<T>(N);
  }
}

template <typename T, bool Lds>
__global__ __launch_bounds__(256) void tree_rank_hex_kernel(
    TreeQueryParams p, const uint8_t* __restrict__ qs, uint64_t m, uint64_t* __restrict__ out) {
  extern __shared__ uint64_t s_htab[];
  const QuadTable<Lds> tab = stage_hex_tab<Lds>(p.hex, s_htab);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / 2;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / 2;
       q < m; q += stride) {
    const uint8_t* e = qs + q * 16;
    const uint64_t i = *reinterpret_cast<const uint64_t*>(e);
    const uint64_t c = map_symbol<T>(p, load_sym(e, sizeof(T)));
    const uint64_t r = i <= p.view.n && c < p.view.sigma ? hex_rank(p, tab, c, i) : 0;
    if ((threadIdx.x & 1) == 0) out[q] = r;
  }
}

// ---------------------------------------------------------------------------
// Select over quad select lines: the ascent takes two levels a step.  At
// quad level q, the kk-th code of child 4N + d of node N is the
// (CD[N][d] + kk)-th d of the level, found in one select line read by the
// pair (lane 0: position and digit words 0..5, lane 1: words 6..13); its
// index in node N is its position minus the node's start, plus one.  When
// the d lies past the line's 448 digits, the step walks on over the quad
// lines from the next one, whose headers count the d's before each line.
constexpr uint64_t kQuadLdsTab = 4096;  // u64 entries


// Lds only when ntab <= kQuadLdsTab (the launch's LDS holds the table)
template <bool Lds>
__device__ __forceinline__ QuadTable<Lds> stage_quad_tab(const QuadView& q, uint64_t* lds) {
  if constexpr (!Lds) {
    return QuadTable<Lds>{q.tab};
  } else {
    for (uint64_t i = threadIdx.x; i < q.ntab; i += blockDim.x) lds[i] = q.tab[i];
    __syncthreads();
    return QuadTable<Lds>{lds};
  }
}

// Occupancy: the kernel takes 66 VGPRs (7 waves per SIMD) and the target
// keeps it at 6 or more.  With the directory selects inlined for lines that
// overflow, it had taken 99 (4 waves), and 6 waves were worth 1.32 vs 1.50 ms
// per 10^7 at n = 6e9, sigma = 256; forcing 8 now spills 2 VGPRs and is no
// faster (1.26 vs 1.20-1.24 ms).
template <typename T, bool Lds>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void tree_select_quad_kernel(
    TreeQueryParams p, const uint8_t* __restrict__ qs, uint64_t m, uint64_t* __restrict__ out) {
  extern __shared__ uint64_t s_qtab[];
  const QuadTable<Lds> tab = stage_quad_tab<Lds>(p.quad, s_qtab);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / 2;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / 2;
       q < m; q += stride) {
    const uint8_t* e = qs + q * 16;
    const uint64_t k = *reinterpret_cast<const uint64_t*>(e);
    const uint64_t c = map_symbol<T>(p, load_sym(e, sizeof(T)));
    const uint64_t res = quad_select(p, tab, c, k);
    if ((threadIdx.x & 1) == 0) out[q] = res;
  }
}

// Occupancy target: with select lines (and without the hint path, below),
// 7 waves per SIMD (68 VGPRs, no spills, instead of the compiler's 123 and 4
// waves): 2.7 vs 3.0 ms per 10^7 at n = 6e9, sigma = 256.  Without them the
// compiler's choice is kept: more waves made the hint and directory searches
// slower (4.71 vs 3.94 ms).
template <typename T, bool SelLines>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SelLines ? 7 : 1))) void tree_select_kernel(
    TreeQueryParams p, const uint8_t* __restrict__ qs, uint64_t m,
    uint64_t* __restrict__ out) {
  extern __shared__ uint8_t s_meta[];
  const Meta mt = stage_meta(p.view, s_meta);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       q < m; q += stride) {
    const uint8_t* e = qs + q * 16;
    const uint64_t k = *reinterpret_cast<const uint64_t*>(e);
    const uint64_t c = map_symbol<T>(p, load_sym(e, sizeof(T)));
    // with select lines the hints would only serve the rare overflow of a
    // line: the directory search does that, with fewer registers
    out[q] = detail::tree_select<SelLines, !SelLines>(p.view, mt.lv, mt.counts, mt.ranks,
                                                      mt.nal, c, k);
  }
}

// ---------------------------------------------------------------------------
// Mixed-kind batches: a workgroup takes the batch in chunks of up to
// kKindChunk queries (a multiple of its size, sized so that every workgroup
// of the grid has a chunk) and first lists each chunk's queries by kind in LDS (per wave,
// three ballots give each lane its place; one LDS atomic per kind and wave
// reserves the places), then walks the three lists one after the other.  A
// wave then runs one kind's descent for all its lanes, except the waves at
// the two list boundaries, instead of running the access, rank and select
// descents one after the other for the lanes of each kind.  Results are
// written by query index, so the order of the output is unchanged.
constexpr uint32_t kKindChunk = 1024;

struct KindLists {
  uint16_t idx[3][kKindChunk];
  uint32_t cnt[3];
};

// kind_of(j) for j < len (<= kKindChunk): the kind of the chunk's query j
// (0 access, 1 rank, anything else select, as the per-query dispatch reads
// it).  Every thread of the workgroup calls this; blockDim.x is a multiple
// of 64.
template <typename KindOf>
__device__ __forceinline__ void group_by_kind(KindLists& L, uint32_t len, KindOf kind_of) {
  if (threadIdx.x < 3) L.cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t rounded = (len + blockDim.x - 1) / blockDim.x * blockDim.x;
  for (uint32_t j = threadIdx.x; j < rounded; j += blockDim.x) {
    uint32_t k = 3;  // past the chunk: no kind
    if (j < len) {
      const uint32_t kk = kind_of(j);
      k = kk > 2 ? 2u : kk;
    }
#pragma unroll
    for (uint32_t c = 0; c < 3; ++c) {
      const uint64_t mask = __ballot(k == c);
      if (mask != 0) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&L.cnt[c], static_cast<uint32_t>(__popcll(mask)));
        base = __shfl(base, 0);
        if (k == c) L.idx[c][base + static_cast<uint32_t>(__popcll(mask & below))] = static_cast<uint16_t>(j);
      }
    }
  }
  __syncthreads();
}

// Calls run(j, kind) for every query of the chunk, in kind order, spread
// over the workgroup; then waits for the workgroup (the lists are reused).
template <typename Run>
__device__ __forceinline__ void run_by_kind(const KindLists& L, Run run) {
  const uint32_t c0 = L.cnt[0], c1 = L.cnt[1], tot = c0 + c1 + L.cnt[2];
  for (uint32_t e = threadIdx.x; e < tot; e += blockDim.x) {
    uint32_t j, k;
    if (e < c0) { j = L.idx[0][e]; k = 0; }
    else if (e < c0 + c1) { j = L.idx[1][e - c0]; k = 1; }
    else { j = L.idx[2][e - c0 - c1]; k = 2; }
    run(j, k);
  }
  __syncthreads();
}

// Occupancy target: 5 waves per SIMD (96 VGPRs, no spills; the compiler
// alone gives the kernel, with all three descents inlined, 138 VGPRs and 3
// waves).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void tree_mixed_kernel(
    TreeQueryParams p, const ecl_query* __restrict__ qs, uint64_t m,
    uint64_t* __restrict__ out, uint32_t chunk) {
  extern __shared__ uint8_t s_meta[];
  __shared__ KindLists lists;
  const Meta mt = stage_meta(p.view, s_meta);
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * chunk; base < m;
       base += static_cast<uint64_t>(gridDim.x) * chunk) {
    const uint32_t len = static_cast<uint32_t>(m - base < chunk ? m - base : chunk);
    const ecl_query* cq = qs + base;
    group_by_kind(lists, len, [&](uint32_t j) { return cq[j].kind; });
    run_by_kind(lists, [&](uint32_t j, uint32_t k) {
      const ecl_query e = cq[j];
      uint64_t r;
      if (k == 0) {
        r = e.arg < p.view.n ? detail::tree_access(p.view, mt.lv, mt.counts,
                                                   mt.ranks, mt.nal, e.arg)
                             : 0;
      } else if (k == 1) {
        r = e.arg > p.view.n ? 0
                             : detail::tree_rank(p.view, mt.lv, mt.counts,
                                                 mt.ranks, mt.nal, e.symbol, e.arg);
      } else {
        r = detail::tree_select(p.view, mt.lv, mt.counts, mt.ranks, mt.nal,
                                e.symbol, e.arg);
      }
      out[base + j] = r;
    });
  }
}

// A mixed batch over the quad lines and quad select lines: the chunk's
// queries grouped by kind as above, then the three lists walked with a pair
// of lanes per query.  Symbols are minimal indices.  Occupancy target: 5
// waves per SIMD (96 VGPRs, no spills), which the kernel now reaches by
// itself; with the directory selects inlined it had taken 121 (4 waves), and
// 5 waves were worth 3.52 vs 3.66 ms per 3e7 at n = 6e9, sigma = 256.  6
// waves spill 10 VGPRs and are no faster (3.53 ms).  Hex: access and rank
// go over the hex lines instead of the quad lines.
template <bool Lds, bool Hex>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void tree_mixed_quad_kernel(
    TreeQueryParams p, const ecl_query* __restrict__ qs, uint64_t m,
    uint64_t* __restrict__ out, uint32_t chunk) {
  extern __shared__ uint64_t s_qmix[];
  __shared__ KindLists lists;
  // Lds: the tables used fit (the quad node table, and the quad bases or,
  // with Hex, the hex table, each of at most 4096 entries)
  const QuadTable<Lds> tab = stage_quad_tab<Lds>(p.quad, s_qmix);
  const QuadTable<Lds> base = Hex ? stage_hex_tab<Lds>(p.hex, s_qmix + (Lds ? p.quad.ntab : 0))
                                  : stage_quad_base<Lds>(p.quad, s_qmix + (Lds ? p.quad.ntab : 0));
  const uint32_t pair = threadIdx.x >> 1, npairs = blockDim.x >> 1;
  for (uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * chunk; b0 < m;
       b0 += static_cast<uint64_t>(gridDim.x) * chunk) {
    const uint32_t len = static_cast<uint32_t>(m - b0 < chunk ? m - b0 : chunk);
    const ecl_query* cq = qs + b0;
    group_by_kind(lists, len, [&](uint32_t j) { return cq[j].kind; });
    const uint32_t c0 = lists.cnt[0], c1 = lists.cnt[1], tot = c0 + c1 + lists.cnt[2];
    for (uint32_t e = pair; e < tot; e += npairs) {
      uint32_t j, k;
      if (e < c0) { j = lists.idx[0][e]; k = 0; }
      else if (e < c0 + c1) { j = lists.idx[1][e - c0]; k = 1; }
      else { j = lists.idx[2][e - c0 - c1]; k = 2; }
      const ecl_query q = cq[j];
      uint64_t r;
      if (k == 0) {
        r = q.arg < p.view.n ? (Hex ? hex_access(p, base, q.arg) : quad_access(p, base, q.arg)) : 0;
      } else if (k == 1) {
        r = q.arg <= p.view.n && q.symbol < p.view.sigma
                ? (Hex ? hex_rank(p, base, q.symbol, q.arg) : quad_rank(p, base, q.symbol, q.arg))
                : 0;
      } else {
        r = quad_select(p, tab, q.symbol, q.arg);
      }
      if ((threadIdx.x & 1) == 0) out[b0 + j] = r;
    }
    __syncthreads();
  }
}

// Query generators: field f of query q uses hash(seed + f * golden2, q).
constexpr uint64_t kField = 0xD1B54A32D192ED03ull;

__device__ __forceinline__ uint64_t symbol_count(const TreeView& t, uint64_t c) {
  return c == t.sigma - 1 ? t.n - t.counts[c] : t.counts[c + 1] - t.counts[c];
}

__global__ void tree_gen_kernel(TreeView t, int kind, uint64_t seed,
                                uint64_t m, void* out, int esize) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       q < m; q += stride) {
    const uint64_t h0 = detail::hash64(seed, q);
    const uint64_t h1 = detail::hash64(seed + kField, q);
    const uint64_t h2 = detail::hash64(seed + 2 * kField, q);
    int k = kind == 3 ? static_cast<int>(q % 3) : kind;
    const uint64_t c = detail::bounded64(h1, t.sigma);
    uint64_t arg;
    if (k == 0) {
      arg = detail::bounded64(h0, t.n);
    } else if (k == 1) {
      arg = detail::bounded64(h2, t.n);
    } else {
      const uint64_t cnt = symbol_count(t, c);
      arg = cnt ? 1 + detail::bounded64(h2, cnt) : 1;
    }
    if (kind == 0) {
      static_cast<uint64_t*>(out)[q] = arg;
    } else if (kind == 3) {
      ecl_query e;
      e.arg = arg;
      e.symbol = static_cast<uint32_t>(k == 0 ? 0 : c);
      e.kind = static_cast<uint32_t>(k);
      static_cast<ecl_query*>(out)[q] = e;
    } else {
      uint8_t* e = static_cast<uint8_t*>(out) + q * 16;
      *reinterpret_cast<uint64_t*>(e) = arg;
      *reinterpret_cast<uint64_t*>(e + 8) = 0;
      switch (esize) {
        case 1: e[8] = static_cast<uint8_t>(c); break;
        case 2: *reinterpret_cast<uint16_t*>(e + 8) = static_cast<uint16_t>(c); break;
        case 4: *reinterpret_cast<uint32_t*>(e + 8) = static_cast<uint32_t>(c); break;
        default: *reinterpret_cast<uint64_t*>(e + 8) = c; break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Batched queries over many trees: query q names its tree; the tree's view
// (levels, counts, node ranks) is read from global memory, where the
// metadata of all trees stays in the L2 / MALL.  A grid of a few
// workgroups per CU strides over the batch.
// One query of kind k against tree t with its metadata at mt (LDS or the
// tree's own arrays in global memory).
__device__ __forceinline__ uint64_t tree_one(const TreeView& t, const Meta& mt, uint32_t k,
                                             uint64_t arg, uint32_t symbol) {
  if (k == 0)
    return arg < t.n ? detail::tree_access(t, mt.lv, mt.counts, mt.ranks, mt.nal, arg) : 0;
  if (k == 1)
    return arg > t.n ? 0 : detail::tree_rank(t, mt.lv, mt.counts, mt.ranks, mt.nal, symbol, arg);
  return detail::tree_select(t, mt.lv, mt.counts, mt.ranks, mt.nal, symbol, arg);
}

__device__ __forceinline__ uint64_t forest_one(const TreeView* __restrict__ views,
                                               uint64_t num_views, uint64_t tree, uint32_t k,
                                               uint64_t arg, uint32_t symbol) {
  if (tree >= num_views) return 0;
  const TreeView t = views[tree];
  return tree_one(t, Meta{t.levels, t.counts, t.ranks, t.nodes_at_level}, k, arg, symbol);
}

// LDS for one tree's staged metadata in the forest and served-batch kernels
// (level views, nodes per level, C array, node ranks; under 5 KB for the
// uint8 trees of 2..256 symbols they hold).
constexpr uint32_t kTreeMetaLds = 8192;

// A chunk whose queries all name one tree stages that tree's metadata in
// LDS, as the single-tree kernels do, so the descents read it there instead
// of from global memory at every level; other chunks read each query's tree
// from global memory.
__global__ void forest_query_kernel(const TreeView* __restrict__ views, uint64_t num_trees,
                                    const ecl_forest_query* __restrict__ qs, uint64_t m,
                                    uint64_t* __restrict__ out, uint32_t chunk) {
  __shared__ KindLists lists;
  __shared__ __align__(16) uint8_t s_meta[kTreeMetaLds];
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * chunk; base < m;
       base += static_cast<uint64_t>(gridDim.x) * chunk) {
    const uint32_t len = static_cast<uint32_t>(m - base < chunk ? m - base : chunk);
    const ecl_forest_query* cq = qs + base;
    group_by_kind(lists, len, [&](uint32_t j) { return static_cast<uint32_t>(cq[j].kind); });
    const uint32_t t0 = cq[0].tree;
    bool other = false;
    for (uint32_t j = threadIdx.x; j < len; j += blockDim.x) other |= cq[j].tree != t0;
    const bool one_tree = !__syncthreads_or(other) && t0 < num_trees &&
                          meta_bytes_of(views[t0]) <= kTreeMetaLds;
    if (one_tree) {
      const TreeView t = views[t0];
      const Meta mt = stage_meta(t, s_meta);
      run_by_kind(lists, [&](uint32_t j, uint32_t k) {
        const ecl_forest_query e = cq[j];
        out[base + j] = tree_one(t, mt, k, e.arg, e.symbol);
      });
    } else {
      run_by_kind(lists, [&](uint32_t j, uint32_t k) {
        const ecl_forest_query e = cq[j];
        out[base + j] = forest_one(views, num_trees, e.tree, k, e.arg, e.symbol);
      });
    }
  }
}

// A served batch: requests' queries back to back, segment j (queries from
// seg_start[j] on) belonging to the tree in view slot seg_tree[j].
__device__ __forceinline__ uint32_t segment_of(const uint64_t* __restrict__ seg_start,
                                               uint32_t nseg, uint64_t q) {
  uint32_t lo = 0, hi = nseg;  // last segment with seg_start <= q
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (seg_start[mid] <= q) lo = mid; else hi = mid;
  }
  return lo;
}

// Chunks whose segments all name one tree stage its metadata in LDS (as in
// forest_query_kernel) and skip the per-query segment search.
__global__ void segment_query_kernel(const TreeView* __restrict__ views, uint64_t num_views,
                                     const ecl_query* __restrict__ qs, uint64_t m,
                                     const uint64_t* __restrict__ seg_start,
                                     const uint32_t* __restrict__ seg_tree, uint32_t nseg,
                                     uint64_t* __restrict__ out, uint32_t chunk) {
  __shared__ KindLists lists;
  __shared__ __align__(16) uint8_t s_meta[kTreeMetaLds];
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * chunk; base < m;
       base += static_cast<uint64_t>(gridDim.x) * chunk) {
    const uint32_t len = static_cast<uint32_t>(m - base < chunk ? m - base : chunk);
    const ecl_query* cq = qs + base;
    group_by_kind(lists, len, [&](uint32_t j) { return cq[j].kind; });
    const uint32_t s0 = segment_of(seg_start, nseg, base);
    const uint32_t s1 = segment_of(seg_start, nseg, base + len - 1);
    const uint32_t t0 = seg_tree[s0];
    bool other = false;
    for (uint32_t x = s0 + 1 + threadIdx.x; x <= s1; x += blockDim.x) other |= seg_tree[x] != t0;
    const bool one_tree = !__syncthreads_or(other) && t0 < num_views &&
                          meta_bytes_of(views[t0]) <= kTreeMetaLds;
    if (one_tree) {
      const TreeView t = views[t0];
      const Meta mt = stage_meta(t, s_meta);
      run_by_kind(lists, [&](uint32_t j, uint32_t k) {
        const ecl_query e = cq[j];
        out[base + j] = tree_one(t, mt, k, e.arg, e.symbol);
      });
    } else {
      run_by_kind(lists, [&](uint32_t j, uint32_t k) {
        const uint64_t q = base + j;
        const ecl_query e = cq[j];
        out[q] = forest_one(views, num_views, seg_tree[segment_of(seg_start, nseg, q)], k, e.arg,
                            e.symbol);
      });
    }
  }
}

// A served batch over trees that carry quad lines and quad select lines
// (and hex lines where built): as segment_query_kernel, but a pair of lanes
// per query, each query's tree parameters read from global memory, and the
// descents of the single-tree quad and hex kernels.  A tree without quad
// select lines takes its binary descents (both lanes of the pair run them).
// One chunk of a batch over many trees (its queries grouped by kind in
// lists), a pair of lanes per query: trees with quad select lines run the
// single-tree quad and hex descents over their parameters (params[t], from
// global memory), the others the binary ones (access and rank fetching
// half of each level's line or block per lane, select in both lanes).
// tree_of(q, query) names the query's tree.
template <typename Query, typename TreeOf>
__device__ __forceinline__ void pair_chunk(const TreeQueryParams* __restrict__ params,
                                           uint64_t num_views, const Query* __restrict__ cq,
                                           uint64_t b0, const KindLists& lists,
                                           uint64_t* __restrict__ out, TreeOf tree_of) {
  const uint32_t pair = threadIdx.x >> 1, npairs = blockDim.x >> 1;
  const uint32_t c0 = lists.cnt[0], c1 = lists.cnt[1], tot = c0 + c1 + lists.cnt[2];
  for (uint32_t e = pair; e < tot; e += npairs) {
    uint32_t j, k;
    if (e < c0) { j = lists.idx[0][e]; k = 0; }
    else if (e < c0 + c1) { j = lists.idx[1][e - c0]; k = 1; }
    else { j = lists.idx[2][e - c0 - c1]; k = 2; }
    const uint64_t q = b0 + j;
    const Query qq = cq[j];
    const uint64_t t = tree_of(q, qq);
    uint64_t r = 0;
    if (t < num_views) {
      const TreeQueryParams& p = params[t];
      if (p.quad.has_sel) {
        if (k == 0) {
          r = qq.arg < p.view.n
                  ? (p.hex.levels ? hex_access(p, QuadTable<false>{p.hex.tab}, qq.arg)
                                  : quad_access(p, QuadTable<false>{p.quad.base}, qq.arg))
                  : 0;
        } else if (k == 1) {
          r = qq.arg <= p.view.n && qq.symbol < p.view.sigma
                  ? (p.hex.levels ? hex_rank(p, QuadTable<false>{p.hex.tab}, qq.symbol, qq.arg)
                                  : quad_rank(p, QuadTable<false>{p.quad.base}, qq.symbol, qq.arg))
                  : 0;
        } else {
          r = quad_select(p, QuadTable<false>{p.quad.tab}, qq.symbol, qq.arg);
        }
      } else {
        const TreeView& v = p.view;
        if (k == 0)
          r = qq.arg < v.n ? detail::tree_access<true>(v, v.levels, v.counts, v.ranks,
                                                       v.nodes_at_level, qq.arg)
                           : 0;
        else if (k == 1)
          r = qq.arg > v.n ? 0
                           : detail::tree_rank<true>(v, v.levels, v.counts, v.ranks,
                                                     v.nodes_at_level, qq.symbol, qq.arg);
        else
          r = detail::tree_select(v, v.levels, v.counts, v.ranks, v.nodes_at_level, qq.symbol,
                                  qq.arg);
      }
    }
    if ((threadIdx.x & 1) == 0) out[q] = r;
  }
  __syncthreads();
}

// (4 waves per SIMD, as forest_query_quad_kernel below)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void segment_query_quad_kernel(
    const TreeView* __restrict__ views, const TreeQueryParams* __restrict__ params,
    uint64_t num_views, const ecl_query* __restrict__ qs, uint64_t m,
    const uint64_t* __restrict__ seg_start, const uint32_t* __restrict__ seg_tree, uint32_t nseg,
    uint64_t* __restrict__ out, uint32_t chunk) {
  __shared__ KindLists lists;
  for (uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * chunk; b0 < m;
       b0 += static_cast<uint64_t>(gridDim.x) * chunk) {
    const uint32_t len = static_cast<uint32_t>(m - b0 < chunk ? m - b0 : chunk);
    const ecl_query* cq = qs + b0;
    group_by_kind(lists, len, [&](uint32_t j) { return cq[j].kind; });
    pair_chunk(params, num_views, cq, b0, lists, out,
               [&](uint64_t q, const ecl_query&) -> uint64_t {
                 return seg_tree[segment_of(seg_start, nseg, q)];
               });
  }
}

// A forest batch over the trees' quad layouts: query q names its tree.  A
// chunk whose queries all name one tree without quad layouts runs as in
// forest_query_kernel (the tree's metadata staged in LDS, a lane per
// query); other chunks run pair_chunk.  Occupancy target: the compiler's
// choice is about 170 VGPRs (2 waves per SIMD); 4 waves fit in 128 without
// spills (a forest batch of 10^8 queries: 22.1 -> 14.2 ms).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void forest_query_quad_kernel(
    const TreeView* __restrict__ views, const TreeQueryParams* __restrict__ params,
    uint64_t num_trees, const ecl_forest_query* __restrict__ qs, uint64_t m,
    uint64_t* __restrict__ out, uint32_t chunk) {
  __shared__ KindLists lists;
  __shared__ __align__(16) uint8_t s_meta[kTreeMetaLds];
  for (uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * chunk; b0 < m;
       b0 += static_cast<uint64_t>(gridDim.x) * chunk) {
    const uint32_t len = static_cast<uint32_t>(m - b0 < chunk ? m - b0 : chunk);
    const ecl_forest_query* cq = qs + b0;
    group_by_kind(lists, len, [&](uint32_t j) { return static_cast<uint32_t>(cq[j].kind); });
    const uint32_t t0 = cq[0].tree;
    bool other = false;
    for (uint32_t j = threadIdx.x; j < len; j += blockDim.x) other |= cq[j].tree != t0;
    const bool binary_tree = !__syncthreads_or(other) && t0 < num_trees &&
                             !params[t0].quad.has_sel && meta_bytes_of(views[t0]) <= kTreeMetaLds;
    if (binary_tree) {
      const TreeView t = views[t0];
      const Meta mt = stage_meta(t, s_meta);
      run_by_kind(lists, [&](uint32_t j, uint32_t k) {
        const ecl_forest_query e = cq[j];
        out[b0 + j] = tree_one(t, mt, k, e.arg, e.symbol);
      });
    } else {
      pair_chunk(params, num_trees, cq, b0, lists, out,
                 [](uint64_t, const ecl_forest_query& e) -> uint64_t { return e.tree; });
    }
  }
}

// Mixed queries for a forest: queries [qoff[i], qoff[i+1]) go to tree i;
// kind q % 3, fields drawn like tree_gen_kernel's.
__global__ void forest_gen_kernel(const TreeView* __restrict__ views, uint64_t num_trees,
                                  const uint64_t* __restrict__ qoff, uint64_t seed, uint64_t m,
                                  ecl_forest_query* __restrict__ out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       q < m; q += stride) {
    uint64_t lo = 0, hi = num_trees;  // last tree with qoff[tree] <= q
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (qoff[mid] <= q) lo = mid; else hi = mid;
    }
    const TreeView t = views[lo];
    const uint64_t h0 = detail::hash64(seed, q);
    const uint64_t h1 = detail::hash64(seed + kField, q);
    const uint64_t h2 = detail::hash64(seed + 2 * kField, q);
    const uint32_t k = static_cast<uint32_t>(q % 3);
    const uint64_t c = detail::bounded64(h1, t.sigma);
    uint64_t arg;
    if (k == 0) {
      arg = detail::bounded64(h0, t.n);
    } else if (k == 1) {
      arg = detail::bounded64(h2, t.n);
    } else {
      const uint64_t cnt = symbol_count(t, c);
      arg = cnt ? 1 + detail::bounded64(h2, cnt) : 1;
    }
    ecl_forest_query e;
    e.arg = arg;
    e.symbol = static_cast<uint32_t>(k == 0 ? 0 : c);
    e.kind = static_cast<uint16_t>(k);
    e.reserved = 0;
    e.tree = static_cast<uint32_t>(lo);
    e.reserved2 = 0;
    out[q] = e;
  }
}

// G lanes per query as in tree_access_kernel (2: the pair fetch over query
// lines).
template <int Value, int G>
__global__ __launch_bounds__(256) void rs_rank_kernel(
    const LevelView* lvp, const uint64_t* __restrict__ idx, uint64_t m,
    uint64_t* __restrict__ out, int generated, uint64_t seed) {
  const LevelView v = *lvp;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / G;
  for (uint64_t q = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / G;
       q < m; q += stride) {
    const uint64_t i =
        generated ? detail::bounded64(detail::hash64(seed, q), v.size) : idx[q];
    const uint64_t r = G == 2 ? detail::level_rank_pair<Value>(v, i)
                              : detail::level_rank<Value>(v, i);
    if (G == 1 || (threadIdx.x & 1) == 0) out[q] = r;
  }
}

template <int Value>
__global__ __launch_bounds__(256) void rs_select_kernel(
    const LevelView* lvp, const uint64_t* __restrict__ ks, uint64_t m,
    uint64_t* __restrict__ out, int generated, uint64_t seed) {
  const LevelView v = *lvp;
  const uint64_t total = Value ? v.ones : v.size - v.ones;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       q < m; q += stride) {
    const uint64_t k =
        generated ? (total ? 1 + detail::bounded64(detail::hash64(seed, q), total) : 1)
                  : ks[q];
    out[q] = detail::level_select<Value, true>(v, k);
  }
}

// The kind grouping works on whole waves of one workgroup.
void check_kind_launch(uint32_t block, uint32_t grid) {
  if (block == 0 || block % 64 != 0 || block > 1024 || grid == 0)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "query launch: block must be a multiple of 64 up to 1024");
}

// Chunk length for a grid of `grid` workgroups of `block` lanes: the
// smallest multiple of the block that gives every workgroup a chunk, between
// one block and kKindChunk.
uint32_t kind_chunk(uint64_t m, uint32_t grid, uint32_t block) {
  uint64_t c = (m + grid - 1) / grid;
  c = (c + block - 1) / block * block;
  if (c < block) c = block;
  if (c > kKindChunk) c = kKindChunk / block * block;
  return static_cast<uint32_t>(c);
}

uint32_t query_grid(uint64_t m) {
  uint64_t g = (m + 255) / 256;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return static_cast<uint32_t>(g);
}

size_t meta_bytes(const TreeView& t) {
  if (t.num_levels > kMaxLevels)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "too many levels");
  return meta_bytes_of(t);
}

}  // namespace

// 0: automatic -- two lanes per access/rank query over query lines (a 128 B
// line split over a pair moves more lines per second), one over the plain
// directory (its 64 B block and two directory entries gain nothing from the
// split).  Select keeps one lane: over the select hints a pair was slower
// (5.7 vs 3.8 ms per 10^7 at n = 6e9, sigma = 256).
std::atomic<uint32_t> g_query_lanes{0};

bool pair_lanes_for(int has_lines) {
  const uint32_t l = g_query_lanes.load();
  return l == 0 ? has_lines != 0 : l == 2;
}

bool pair_lanes(const TreeQueryParams& p) {
  return p.lanes ? p.lanes == 2 : pair_lanes_for(p.has_lines);
}

template <int G>
void access_g(const TreeQueryParams& p, const uint64_t* idx, uint64_t m, void* out, hipStream_t s) {
  const dim3 g(query_grid(m * G)), b(256);
  const size_t sh = meta_bytes(p.view);
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_access_kernel<uint8_t, G>), g, b, sh, s, p, idx, m, static_cast<uint8_t*>(out)); break;
    case 2: hipLaunchKernelGGL((tree_access_kernel<uint16_t, G>), g, b, sh, s, p, idx, m, static_cast<uint16_t*>(out)); break;
    case 4: hipLaunchKernelGGL((tree_access_kernel<uint32_t, G>), g, b, sh, s, p, idx, m, static_cast<uint32_t*>(out)); break;
    default: hipLaunchKernelGGL((tree_access_kernel<uint64_t, G>), g, b, sh, s, p, idx, m, static_cast<uint64_t*>(out)); break;
  }
}

template <int G>
void rank_g(const TreeQueryParams& p, const uint8_t* qq, uint64_t m, uint64_t* out, hipStream_t s) {
  const dim3 g(query_grid(m * G)), b(256);
  const size_t sh = meta_bytes(p.view);
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_rank_kernel<uint8_t, G>), g, b, sh, s, p, qq, m, out); break;
    case 2: hipLaunchKernelGGL((tree_rank_kernel<uint16_t, G>), g, b, sh, s, p, qq, m, out); break;
    case 4: hipLaunchKernelGGL((tree_rank_kernel<uint32_t, G>), g, b, sh, s, p, qq, m, out); break;
    default: hipLaunchKernelGGL((tree_rank_kernel<uint64_t, G>), g, b, sh, s, p, qq, m, out); break;
  }
}

void set_query_lanes(uint32_t lanes) { g_query_lanes.store(lanes); }

template <bool Lds>
void access_quad(const TreeQueryParams& p, const uint64_t* idx, uint64_t m, void* out,
                 hipStream_t s) {
  const dim3 g(query_grid(2 * m)), b(256);
  const size_t sh = Lds ? p.quad.nbase * sizeof(uint64_t) : 0;
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_access_quad_kernel<uint8_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint8_t*>(out)); break;
    case 2: hipLaunchKernelGGL((tree_access_quad_kernel<uint16_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint16_t*>(out)); break;
    case 4: hipLaunchKernelGGL((tree_access_quad_kernel<uint32_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint32_t*>(out)); break;
    default: hipLaunchKernelGGL((tree_access_quad_kernel<uint64_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint64_t*>(out)); break;
  }
}

template <bool Lds>
void access_hex(const TreeQueryParams& p, const uint64_t* idx, uint64_t m, void* out,
                hipStream_t s) {
  const dim3 g(query_grid(2 * m)), b(256);
  const size_t sh = Lds ? p.hex.ntab * sizeof(uint64_t) : 0;
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_access_hex_kernel<uint8_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint8_t*>(out)); break;
    case 2: hipLaunchKernelGGL((tree_access_hex_kernel<uint16_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint16_t*>(out)); break;
    case 4: hipLaunchKernelGGL((tree_access_hex_kernel<uint32_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint32_t*>(out)); break;
    default: hipLaunchKernelGGL((tree_access_hex_kernel<uint64_t, Lds>), g, b, sh, s, p, idx, m, static_cast<uint64_t*>(out)); break;
  }
}

template <bool Lds>
void rank_hex(const TreeQueryParams& p, const uint8_t* qq, uint64_t m, uint64_t* out,
              hipStream_t s) {
  const dim3 g(query_grid(2 * m)), b(256);
  const size_t sh = Lds ? p.hex.ntab * sizeof(uint64_t) : 0;
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_rank_hex_kernel<uint8_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    case 2: hipLaunchKernelGGL((tree_rank_hex_kernel<uint16_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    case 4: hipLaunchKernelGGL((tree_rank_hex_kernel<uint32_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    default: hipLaunchKernelGGL((tree_rank_hex_kernel<uint64_t, Lds>), g, b, sh, s, p, qq, m, out); break;
  }
}

void launch_tree_access(const TreeQueryParams& p, const uint64_t* idx,
                        uint64_t m, void* out, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_WT_ACCESS, s);
  if (p.hex.levels) {
    if (p.hex.ntab <= kHexLdsTab) access_hex<true>(p, idx, m, out, s);
    else access_hex<false>(p, idx, m, out, s);
  } else if (p.quad.levels) {
    if (p.quad.nbase <= kQuadLdsBase) access_quad<true>(p, idx, m, out, s);
    else access_quad<false>(p, idx, m, out, s);
  } else if (pair_lanes(p)) access_g<2>(p, idx, m, out, s);
  else access_g<1>(p, idx, m, out, s);
  ECL_HIP(hipGetLastError());
}

template <bool Lds>
void rank_quad(const TreeQueryParams& p, const uint8_t* qq, uint64_t m, uint64_t* out,
               hipStream_t s) {
  const dim3 g(query_grid(2 * m)), b(256);
  const size_t sh = Lds ? p.quad.nbase * sizeof(uint64_t) : 0;
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_rank_quad_kernel<uint8_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    case 2: hipLaunchKernelGGL((tree_rank_quad_kernel<uint16_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    case 4: hipLaunchKernelGGL((tree_rank_quad_kernel<uint32_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    default: hipLaunchKernelGGL((tree_rank_quad_kernel<uint64_t, Lds>), g, b, sh, s, p, qq, m, out); break;
  }
}

void launch_tree_rank(const TreeQueryParams& p, const void* q, uint64_t m,
                      uint64_t* out, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_WT_RANK, s);
  const uint8_t* qq = static_cast<const uint8_t*>(q);
  if (p.hex.levels) {
    if (p.hex.ntab <= kHexLdsTab) rank_hex<true>(p, qq, m, out, s);
    else rank_hex<false>(p, qq, m, out, s);
  } else if (p.quad.levels) {
    if (p.quad.nbase <= kQuadLdsBase) rank_quad<true>(p, qq, m, out, s);
    else rank_quad<false>(p, qq, m, out, s);
  } else if (pair_lanes(p)) rank_g<2>(p, qq, m, out, s);
  else rank_g<1>(p, qq, m, out, s);
  ECL_HIP(hipGetLastError());
}

template <bool SL>
void select_sl(const TreeQueryParams& p, const uint8_t* qq, uint64_t m, uint64_t* out,
               hipStream_t s) {
  const dim3 g(query_grid(m)), b(256);
  const size_t sh = meta_bytes(p.view);
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_select_kernel<uint8_t, SL>), g, b, sh, s, p, qq, m, out); break;
    case 2: hipLaunchKernelGGL((tree_select_kernel<uint16_t, SL>), g, b, sh, s, p, qq, m, out); break;
    case 4: hipLaunchKernelGGL((tree_select_kernel<uint32_t, SL>), g, b, sh, s, p, qq, m, out); break;
    default: hipLaunchKernelGGL((tree_select_kernel<uint64_t, SL>), g, b, sh, s, p, qq, m, out); break;
  }
}

template <bool Lds>
void select_quad(const TreeQueryParams& p, const uint8_t* qq, uint64_t m, uint64_t* out,
                 hipStream_t s) {
  const dim3 g(query_grid(2 * m)), b(256);
  const size_t sh = Lds ? p.quad.ntab * sizeof(uint64_t) : 0;
  switch (p.esize) {
    case 1: hipLaunchKernelGGL((tree_select_quad_kernel<uint8_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    case 2: hipLaunchKernelGGL((tree_select_quad_kernel<uint16_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    case 4: hipLaunchKernelGGL((tree_select_quad_kernel<uint32_t, Lds>), g, b, sh, s, p, qq, m, out); break;
    default: hipLaunchKernelGGL((tree_select_quad_kernel<uint64_t, Lds>), g, b, sh, s, p, qq, m, out); break;
  }
}

void launch_tree_select(const TreeQueryParams& p, const void* q, uint64_t m,
                        uint64_t* out, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_WT_SELECT, s);
  const uint8_t* qq = static_cast<const uint8_t*>(q);
  if (p.quad.has_sel) {
    if (p.quad.ntab <= kQuadLdsTab) select_quad<true>(p, qq, m, out, s);
    else select_quad<false>(p, qq, m, out, s);
  } else if (p.has_sel_lines) select_sl<true>(p, qq, m, out, s);
  else select_sl<false>(p, qq, m, out, s);
  ECL_HIP(hipGetLastError());
}

void launch_tree_mixed(const TreeQueryParams& p, const ecl_query* q,
                       uint64_t m, uint64_t* out, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_WT_MIXED, s);
  if (p.quad.levels && p.quad.has_sel) {
    const uint32_t g = query_grid(2 * m);
    const uint32_t chunk = kind_chunk(m, g, 256);
    if (p.hex.levels) {
      if (p.hex.ntab <= kHexLdsTab && p.quad.ntab <= kQuadLdsTab) {
        const size_t sh = (p.hex.ntab + p.quad.ntab) * sizeof(uint64_t);
        hipLaunchKernelGGL((tree_mixed_quad_kernel<true, true>), dim3(g), dim3(256), sh, s, p, q, m,
                           out, chunk);
      } else {
        hipLaunchKernelGGL((tree_mixed_quad_kernel<false, true>), dim3(g), dim3(256), 0, s, p, q, m,
                           out, chunk);
      }
    } else if (p.quad.nbase <= kQuadLdsBase && p.quad.ntab <= kQuadLdsTab) {
      const size_t sh = (p.quad.nbase + p.quad.ntab) * sizeof(uint64_t);
      hipLaunchKernelGGL((tree_mixed_quad_kernel<true, false>), dim3(g), dim3(256), sh, s, p, q, m,
                         out, chunk);
    } else {
      hipLaunchKernelGGL((tree_mixed_quad_kernel<false, false>), dim3(g), dim3(256), 0, s, p, q, m,
                         out, chunk);
    }
  } else {
    const uint32_t g = query_grid(m);
    hipLaunchKernelGGL(tree_mixed_kernel, dim3(g), dim3(256), meta_bytes(p.view), s, p, q, m, out,
                       kind_chunk(m, g, 256));
  }
  ECL_HIP(hipGetLastError());
}

void launch_forest_query(const TreeView* views, uint64_t num_trees, const ecl_forest_query* q,
                         uint64_t m, uint64_t* out, uint32_t block, uint32_t grid,
                         hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_FOREST_QUERY, s);
  check_kind_launch(block, grid);
  hipLaunchKernelGGL(forest_query_kernel, dim3(grid), dim3(block), 0, s, views, num_trees, q,
                     m, out, kind_chunk(m, grid, block));
  ECL_HIP(hipGetLastError());
}

void launch_segment_query(const TreeView* views, uint64_t num_views, const ecl_query* q,
                          uint64_t m, const uint64_t* seg_start, const uint32_t* seg_tree,
                          uint32_t nseg, uint64_t* out, uint32_t block, uint32_t grid,
                          hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_FOREST_QUERY, s);
  check_kind_launch(block, grid);
  hipLaunchKernelGGL(segment_query_kernel, dim3(grid), dim3(block), 0, s, views, num_views, q, m,
                     seg_start, seg_tree, nseg, out, kind_chunk(m, grid, block));
  ECL_HIP(hipGetLastError());
}

void launch_segment_query_quad(const TreeView* views, const TreeQueryParams* params,
                               uint64_t num_views, const ecl_query* q, uint64_t m,
                               const uint64_t* seg_start, const uint32_t* seg_tree, uint32_t nseg,
                               uint64_t* out, uint32_t grid, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_FOREST_QUERY, s);
  check_kind_launch(256, grid);
  hipLaunchKernelGGL(segment_query_quad_kernel, dim3(grid), dim3(256), 0, s, views, params,
                     num_views, q, m, seg_start, seg_tree, nseg, out, kind_chunk(m, grid, 256));
  ECL_HIP(hipGetLastError());
}

void launch_forest_query_quad(const TreeView* views, const TreeQueryParams* params,
                              uint64_t num_trees, const ecl_forest_query* q, uint64_t m,
                              uint64_t* out, uint32_t grid, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_FOREST_QUERY, s);
  check_kind_launch(256, grid);
  hipLaunchKernelGGL(forest_query_quad_kernel, dim3(grid), dim3(256), 0, s, views, params,
                     num_trees, q, m, out, kind_chunk(m, grid, 256));
  ECL_HIP(hipGetLastError());
}

void launch_forest_gen(const TreeView* views, uint64_t num_trees, const uint64_t* qoff,
                       uint64_t seed, uint64_t m, ecl_forest_query* out, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_GEN, s);
  hipLaunchKernelGGL(forest_gen_kernel, dim3(query_grid(m)), dim3(256), 0, s, views, num_trees,
                     qoff, seed, m, out);
  ECL_HIP(hipGetLastError());
}

void launch_tree_gen_queries(const TreeView& v, const uint64_t*, int kind,
                             uint64_t seed, uint64_t m, void* out, int esize,
                             hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_GEN, s);
  hipLaunchKernelGGL(tree_gen_kernel, dim3(query_grid(m)), dim3(256), 0, s, v,
                     kind, seed, m, out, esize);
  ECL_HIP(hipGetLastError());
}

void launch_rs_rank(const LevelView* lv, int value, const uint64_t* idx,
                    uint64_t m, uint64_t* out, int generated, uint64_t seed,
                    int has_lines, hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_RS_RANK, s);
  if (pair_lanes_for(has_lines)) {
    const dim3 g(query_grid(2 * m)), b(256);
    if (value)
      hipLaunchKernelGGL((rs_rank_kernel<1, 2>), g, b, 0, s, lv, idx, m, out, generated, seed);
    else
      hipLaunchKernelGGL((rs_rank_kernel<0, 2>), g, b, 0, s, lv, idx, m, out, generated, seed);
  } else {
    const dim3 g(query_grid(m)), b(256);
    if (value)
      hipLaunchKernelGGL((rs_rank_kernel<1, 1>), g, b, 0, s, lv, idx, m, out, generated, seed);
    else
      hipLaunchKernelGGL((rs_rank_kernel<0, 1>), g, b, 0, s, lv, idx, m, out, generated, seed);
  }
  ECL_HIP(hipGetLastError());
}

void launch_rs_select(const LevelView* lv, int value, const uint64_t* k,
                      uint64_t m, uint64_t* out, int generated, uint64_t seed,
                      hipStream_t s) {
  if (m == 0) return;
  ProfScope ps(K_RS_SELECT, s);
  const dim3 g(query_grid(m)), b(256);
  if (value)
    hipLaunchKernelGGL(rs_select_kernel<1>, g, b, 0, s, lv, k, m, out, generated, seed);
  else
    hipLaunchKernelGGL(rs_select_kernel<0>, g, b, 0, s, lv, k, m, out, generated, seed);
  ECL_HIP(hipGetLastError());
}

}  // namespace lib
}  // namespace ecl
