// kernels_tree.hip -- wavelet-tree construction for gfx950: several levels
// per pass over the text.
//
// Level l of a wavelet tree lists the symbols whose code is longer than l,
// stably ordered by their top l code bits, and stores bit L-1-l of each code
// (the per-level radix sort and fill of tree/wavelet_tree.cuh:806-818 and
// fillLevelKernel, :241-323).  The reference sorts the whole text once per
// level.  Here a pass covers k <= 8 consecutive levels l0 .. l0+k-1 and
// reads the level-l0 order once:
//
//  * tile_hist_kernel: per tile of the level-l0 order, the histogram of the
//    k-bit digits (the code bits of the pass's levels).  The first pass also
//    maps the text to codes here when the text is not already its own code.
//  * scan_*_kernel: for every digit, the number of codes with that digit in
//    all earlier tiles (exclusive column scan of the tile histograms).
//  * pass_kernel: the tile is staged in LDS and partitioned level by level
//    inside LDS (stable one-bit split within each node).  At level l0+i the
//    tile's codes of one node form a single run of that level's bit array,
//    starting at the node's global start plus the node's codes in earlier
//    tiles, both known from the digit counts.  Runs are written as 32-bit
//    words; the two edge words of a run are shared with neighbouring runs and
//    use atomicOr on the zeroed bit array.  A pass that is not the last also
//    writes its codes in the order of level l0+k for the next pass.
//
// Tiles of later passes never straddle a node of level l0: every node of the
// pass input starts on a 16-byte boundary (the previous pass pads between
// nodes) and is cut into tiles of its own.
#include <hip/hip_runtime.h>

#include <atomic>

#include "internal.hpp"

namespace ecl {
namespace lib {
namespace {

constexpr uint32_t kPassThreads = 512;
constexpr uint32_t kMaxBins = 256;

template <int CW> struct CodeOf;
template <> struct CodeOf<1> { using type = uint8_t; };
template <> struct CodeOf<2> { using type = uint16_t; };
template <> struct CodeOf<4> { using type = uint32_t; };
template <> struct CodeOf<8> { using type = uint64_t; };

// bit `bit` of each of the 16/sizeof(T) codes of a 16-byte chunk, code k at bit k
template <typename T>
__device__ __forceinline__ uint32_t chunk_mask(const uint4& v, uint32_t bit) {
  const uint32_t d[4] = {v.x, v.y, v.z, v.w};
  uint32_t m = 0;
  if constexpr (sizeof(T) == 1) {
    // isolate the bit of each byte of a dword and gather the four bits into
    // a nibble with one multiply (the partial products do not overlap)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t x = (d[k] >> bit) & 0x01010101u;
      m |= ((x * 0x01020408u) >> 24 & 0xFu) << (4 * k);
    }
  } else if constexpr (sizeof(T) == 2) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      m |= ((d[k] >> bit) & 1u) << (2 * k);
      m |= ((d[k] >> (16 + bit)) & 1u) << (2 * k + 1);
    }
  } else if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int k = 0; k < 4; ++k) m |= ((d[k] >> bit) & 1u) << k;
  } else {
    const uint64_t a = (static_cast<uint64_t>(v.y) << 32) | v.x;
    const uint64_t b = (static_cast<uint64_t>(v.w) << 32) | v.z;
    m = static_cast<uint32_t>((a >> bit) & 1u) | (static_cast<uint32_t>((b >> bit) & 1u) << 1);
  }
  return m;
}

// code k of a 16-byte chunk, by shifts (k is a constant after unrolling)
template <typename T>
__device__ __forceinline__ T chunk_elem(const uint4& q, uint32_t k) {
  const uint32_t wi = (k * static_cast<uint32_t>(sizeof(T))) >> 2;
  const uint32_t d = wi == 0 ? q.x : wi == 1 ? q.y : wi == 2 ? q.z : q.w;
  if constexpr (sizeof(T) == 1) return static_cast<T>(d >> (8 * (k & 3)));
  else if constexpr (sizeof(T) == 2) return static_cast<T>(d >> (16 * (k & 1)));
  else if constexpr (sizeof(T) == 4) return static_cast<T>(d);
  else return static_cast<T>(k == 0 ? ((static_cast<uint64_t>(q.y) << 32) | q.x)
                                    : ((static_cast<uint64_t>(q.w) << 32) | q.z));
}

// `width` bits of code k of the chunk from bit `shift` on: one bit-field
// extract from the chunk's dword below 64-bit codes
template <typename T>
__device__ __forceinline__ uint32_t child_of(const uint4& q, uint32_t k, uint32_t shift,
                                             uint32_t width) {
  if constexpr (sizeof(T) == 8) {
    return static_cast<uint32_t>(chunk_elem<T>(q, k) >> (shift & 63)) & ((1u << width) - 1u);
  } else {
    const uint32_t wi = (k * static_cast<uint32_t>(sizeof(T))) >> 2;
    const uint32_t d = wi == 0 ? q.x : wi == 1 ? q.y : wi == 2 ? q.z : q.w;
    const uint32_t bitpos = (k * 8 * static_cast<uint32_t>(sizeof(T))) & 31;
    return __builtin_amdgcn_ubfe(d, bitpos + shift, width);
  }
}

// the chunk at src[s0 ..] with only the first `valid` codes loaded (rest 0)
template <typename T>
__device__ __forceinline__ uint4 load_partial_chunk(const T* src, uint64_t s0, uint32_t valid) {
  uint32_t d[4] = {0, 0, 0, 0};
  constexpr uint32_t E = 16 / sizeof(T);
#pragma unroll
  for (uint32_t k = 0; k < E; ++k) {
    if (k < valid) {
      const uint64_t x = static_cast<uint64_t>(src[s0 + k]);
      if constexpr (sizeof(T) == 8) {
        d[2 * k] = static_cast<uint32_t>(x);
        d[2 * k + 1] = static_cast<uint32_t>(x >> 32);
      } else {
        const uint32_t bitpos = k * 8 * static_cast<uint32_t>(sizeof(T));
        d[bitpos >> 5] |= static_cast<uint32_t>(x) << (bitpos & 31);
      }
    }
  }
  return make_uint4(d[0], d[1], d[2], d[3]);
}

// Tile geometry: tile t covers codes [start, start + len) of the pass input,
// inside node `node` of level l0.
struct TileGeo {
  uint64_t start;
  uint32_t len, node;
};
__device__ __forceinline__ TileGeo tile_geo(const PassParams& p, uint64_t t, uint32_t T) {
  TileGeo g;
  if (p.tile_start != nullptr) {
    g.start = p.tile_start[t];
    g.len = p.tile_len[t];
    g.node = p.tile_node[t];
  } else {
    g.start = t * T;
    const uint64_t rest = p.n - g.start;
    g.len = rest < T ? static_cast<uint32_t>(rest) : T;
    g.node = 0;
  }
  return g;
}

// Symbol -> code on the first pass (tree/wavelet_tree.cuh:2295-2413 remap):
// map_kind 0 identity, 1 LUT (8/16-bit texts), 2 lower_bound over the sorted
// alphabet; values outside the alphabet clamp to the last symbol.
template <typename TIn, int MapKind>
__device__ __forceinline__ uint64_t map_code(const PassParams& p, uint64_t x) {
  uint64_t c;
  if constexpr (MapKind == 0) {
    c = x;
  } else if constexpr (MapKind == 1) {
    c = p.lut[x];
  } else {
    const TIn* alpha = static_cast<const TIn*>(p.alphabet);
    uint64_t lo = 0, hi = p.sigma;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (static_cast<uint64_t>(alpha[mid]) < x) lo = mid + 1; else hi = mid;
    }
    c = lo;
  }
  if (c >= p.sigma) c = p.sigma - 1;
  return (p.pow2 || c < p.codes_start) ? c : p.code_table[c - p.codes_start];
}

// ---------------------------------------------------------------------------
// Tile digit histograms.  The workgroup's 8192 LDS counters hold
// C = 8192 / 2^k copies of the 2^k bins, bin d of copy c at d * C + c;
// thread tid counts into copy tid % C, so the lanes of one atomic add hit
// distinct banks whatever their digits (C >= 32).  The copies are summed per
// bin at the end.
constexpr uint32_t kHistSlots = 8192;

template <typename TIn, int CW, int MapKind, bool Remap>
__global__ __launch_bounds__(kPassThreads) void tile_hist_kernel(PassParams p) {
  using Code = typename CodeOf<CW>::type;
  const uint32_t T = p.tile_codes;                   // codes per tile
  constexpr uint32_t EI = 16 / sizeof(TIn);          // input symbols per chunk
  __shared__ uint32_t s_bins[kHistSlots];
  __shared__ uint32_t s_tot[kMaxBins];
  const uint32_t tid = threadIdx.x;
  // C copies of the nb bins, lane-banked (thread tid counts into copy
  // tid % C); never more copies than threads, so small alphabets zero and
  // reduce only the copies that are used
  const uint32_t nb = 1u << p.k;
  const uint32_t C = (kHistSlots >> p.k) < kPassThreads ? (kHistSlots >> p.k) : kPassThreads;
  const uint32_t dmask = nb - 1;
  const uint32_t row = p.row_bins ? p.row_bins : nb;
  for (uint32_t i = tid; i < nb * C / 4; i += kPassThreads)
    reinterpret_cast<uint4*>(s_bins)[i] = make_uint4(0, 0, 0, 0);
  if (tid < kMaxBins) s_tot[tid] = 0;
  const uint64_t t = blockIdx.x;
  __syncthreads();
  uint32_t* bins = s_bins + tid % C;
  const TileGeo g = tile_geo(p, t, T);
  const TIn* in = static_cast<const TIn*>(p.src);
  Code* out = static_cast<Code*>(p.codes_out);
  // batched trees: only remapped nodes are written, through their own table
  const uint8_t* lut = nullptr;
  if constexpr (Remap) {
    if (p.node_remap != nullptr) {
      if (p.node_remap[g.node]) lut = p.node_lut + static_cast<uint64_t>(g.node) * 256;
      else out = nullptr;  // read in place: its symbols are its codes
    }
  }
  const uint32_t nchunks = (g.len + EI - 1) / EI;
  uint32_t c0 = tid;
  if constexpr (!Remap) {
    // four chunks per thread at a time, all loads issued before the counts
    // (one load per thread in flight would leave HBM idle between them)
    constexpr uint32_t U = 4;
    for (; c0 + (U - 1) * kPassThreads < nchunks; c0 += U * kPassThreads) {
      uint4 q[U];
      uint32_t valid[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t c = c0 + u * kPassThreads;
        const uint64_t e0 = g.start + static_cast<uint64_t>(c) * EI;
        valid[u] = (g.len - c * EI < EI) ? g.len - c * EI : EI;
        q[u] = valid[u] == EI ? *reinterpret_cast<const uint4*>(in + e0)
                              : load_partial_chunk<TIn>(in, e0, valid[u]);
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u)
#pragma unroll
        for (uint32_t k = 0; k < EI; ++k)
          if (k < valid[u])
            atomicAdd(&bins[((static_cast<uint64_t>(chunk_elem<TIn>(q[u], k)) >> p.dshift) & dmask) * C],
                      1u);
    }
  }
  for (uint32_t c = c0; c < nchunks; c += kPassThreads) {
    const uint64_t e0 = g.start + static_cast<uint64_t>(c) * EI;
    const uint32_t valid = (g.len - c * EI < EI) ? g.len - c * EI : EI;
    uint4 q;
    if (valid == EI) q = *reinterpret_cast<const uint4*>(in + e0);
    else q = load_partial_chunk<TIn>(in, e0, valid);
    if constexpr (Remap) {
      // codes of this chunk, assembled for one store when the chunk is full
      constexpr uint32_t OB = EI * CW;  // output bytes per chunk (<= 16)
      uint32_t o[4] = {0, 0, 0, 0};
#pragma unroll
      for (uint32_t k = 0; k < EI; ++k) {
        if (k < valid) {
          const uint64_t x = static_cast<uint64_t>(chunk_elem<TIn>(q, k));
          const uint64_t code = lut != nullptr ? lut[x & 255] : map_code<TIn, MapKind>(p, x);
          atomicAdd(&bins[((code >> p.dshift) & dmask) * C], 1u);
          if constexpr (CW == 8) {
            o[2 * k] = static_cast<uint32_t>(code);
            o[2 * k + 1] = static_cast<uint32_t>(code >> 32);
          } else {
            const uint32_t bp = k * 8 * CW;
            o[bp >> 5] |= static_cast<uint32_t>(code) << (bp & 31);
          }
        }
      }
      if (out == nullptr) {
        // counted only (a batched node that is read in place)
      } else if (valid == EI) {
        if constexpr (OB == 16) *reinterpret_cast<uint4*>(out + e0) = make_uint4(o[0], o[1], o[2], o[3]);
        else if constexpr (OB == 8) *reinterpret_cast<uint2*>(out + e0) = make_uint2(o[0], o[1]);
        else if constexpr (OB == 4) *reinterpret_cast<uint32_t*>(out + e0) = o[0];
        else if constexpr (OB == 2) *reinterpret_cast<uint16_t*>(out + e0) = static_cast<uint16_t>(o[0]);
        else *reinterpret_cast<uint8_t*>(out + e0) = static_cast<uint8_t>(o[0]);
      } else {
#pragma unroll
        for (uint32_t k = 0; k < EI; ++k) {
          if (k < valid) {
            if constexpr (CW == 8) {
              out[e0 + k] = (static_cast<uint64_t>(o[2 * k + 1]) << 32) | o[2 * k];
            } else {
              const uint32_t bp = k * 8 * CW;
              out[e0 + k] = static_cast<Code>(o[bp >> 5] >> (bp & 31));
            }
          }
        }
      }
    } else {
#pragma unroll
      for (uint32_t k = 0; k < EI; ++k)
        if (k < valid)
          atomicAdd(&bins[((static_cast<uint64_t>(chunk_elem<TIn>(q, k)) >> p.dshift) & dmask) * C],
                    1u);
    }
  }
  __syncthreads();
  {
    // bin d summed over its copies by kPassThreads / nb threads; the copy
    // order is rotated by d so the lanes of a wave read distinct banks
    const uint32_t per = kPassThreads >> p.k;  // >= 2 (nb <= 256)
    const uint32_t d = tid & dmask, part = tid >> p.k;
    uint32_t sum = 0;
    for (uint32_t j = part; j < C; j += per) sum += s_bins[d * C + ((j + d) & (C - 1))];
    atomicAdd(&s_tot[d], sum);
  }
  __syncthreads();
  for (uint32_t d = tid; d < row; d += kPassThreads) p.hist[t * row + d] = d < nb ? s_tot[d] : 0u;
}

// Tile histogram of 1- or 2-bit digits of a uint8 text that is its own code
// (identity map, digit = symbol & (2^k - 1)): per 32-bit word the digit's
// low and high bits are bit 0 and bit 1 of every byte, so the counts of
// digits 1, 2 and 3 are popcounts of (lo & ~hi), (hi & ~lo) and (hi & lo);
// digit 0 takes the rest of the tile.  Counts stay in registers (no LDS
// atomics) until one wave reduction per workgroup.
__global__ __launch_bounds__(kPassThreads) void tile_hist_small_kernel(PassParams p) {
  __shared__ uint32_t s_tot[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  if (tid < 4) s_tot[tid] = 0;
  const uint64_t t = blockIdx.x;
  const TileGeo g = tile_geo(p, t, p.tile_codes);
  const uint8_t* in = static_cast<const uint8_t*>(p.src);
  const uint32_t hi_mask = p.k >= 2 ? 0x02020202u : 0u;
  uint32_t c1 = 0, c2 = 0, c3 = 0;
  auto count = [&](const uint4 q) {
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t lo = w[j] & 0x01010101u, hi = (w[j] & hi_mask) >> 1;
      c1 += __popc(lo & ~hi);
      c2 += __popc(hi & ~lo);
      c3 += __popc(hi & lo);
    }
  };
  const uint32_t nchunks = (g.len + 15) / 16;
  constexpr uint32_t U = 4;
  uint32_t c0 = tid;
  for (; c0 + (U - 1) * kPassThreads < nchunks; c0 += U * kPassThreads) {
    uint4 q[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t c = c0 + u * kPassThreads;
      const uint64_t e0 = g.start + static_cast<uint64_t>(c) * 16;
      const uint32_t valid = (g.len - c * 16 < 16) ? g.len - c * 16 : 16;
      q[u] = valid == 16 ? *reinterpret_cast<const uint4*>(in + e0)
                         : load_partial_chunk<uint8_t>(in, e0, valid);
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) count(q[u]);
  }
  for (uint32_t c = c0; c < nchunks; c += kPassThreads) {
    const uint64_t e0 = g.start + static_cast<uint64_t>(c) * 16;
    const uint32_t valid = (g.len - c * 16 < 16) ? g.len - c * 16 : 16;
    count(valid == 16 ? *reinterpret_cast<const uint4*>(in + e0)
                      : load_partial_chunk<uint8_t>(in, e0, valid));
  }
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) {
    c1 += __shfl_xor(c1, off, 64);
    c2 += __shfl_xor(c2, off, 64);
    c3 += __shfl_xor(c3, off, 64);
  }
  __syncthreads();
  if (lane == 0) {
    atomicAdd(&s_tot[1], c1);
    atomicAdd(&s_tot[2], c2);
    atomicAdd(&s_tot[3], c3);
  }
  __syncthreads();
  const uint32_t nb = 1u << p.k;
  const uint32_t row = p.row_bins ? p.row_bins : nb;
  for (uint32_t d = tid; d < row; d += kPassThreads) {
    uint32_t v = 0;
    if (d == 0) v = g.len - s_tot[1] - s_tot[2] - s_tot[3];
    else if (d < nb) v = s_tot[d];
    p.hist[t * row + d] = v;
  }
}

// ---------------------------------------------------------------------------
// Exclusive column scan of the tile histograms, in blocks of 256 tiles:
//  partial[b][d] = sum of block b  ->  exclusive over blocks  ->  excl rows.
constexpr uint32_t kScanTiles = 256;

__global__ __launch_bounds__(256) void scan_partial_kernel(const uint32_t* __restrict__ hist,
                                                           uint64_t tiles, uint32_t nb,
                                                           uint64_t* __restrict__ partial) {
  __shared__ uint64_t s[256];
  const uint32_t tid = threadIdx.x;
  const uint32_t rows = 256 / nb, col = tid % nb, r0 = tid / nb;
  const uint64_t t0 = blockIdx.x * static_cast<uint64_t>(kScanTiles);
  const uint64_t t1 = (t0 + kScanTiles < tiles) ? t0 + kScanTiles : tiles;
  uint64_t sum = 0;
  for (uint64_t t = t0 + r0; t < t1; t += rows) sum += hist[t * nb + col];
  s[tid] = sum;
  __syncthreads();
  if (tid < nb) {
    uint64_t tot = 0;
    for (uint32_t r = 0; r < rows; ++r) tot += s[r * nb + tid];
    partial[blockIdx.x * static_cast<uint64_t>(nb) + tid] = tot;
  }
}

// one workgroup per digit column: exclusive scan of the block sums in place
__global__ __launch_bounds__(256) void scan_blocks_kernel(uint64_t* __restrict__ partial,
                                                          uint64_t nblk, uint32_t nb) {
  __shared__ uint64_t s_wave[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t col = blockIdx.x;
  const uint64_t per = (nblk + 255) / 256;
  const uint64_t b0 = tid * per;
  const uint64_t b1 = (b0 + per < nblk) ? b0 + per : nblk;
  uint64_t sum = 0;
  for (uint64_t b = b0; b < b1; ++b) sum += partial[b * nb + col];
  uint64_t incl = sum;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint64_t run = incl - sum;
  for (uint32_t w = 0; w < wave; ++w) run += s_wave[w];
  for (uint64_t b = b0; b < b1; ++b) {
    const uint64_t x = partial[b * nb + col];
    partial[b * nb + col] = run;
    run += x;
  }
}

// excl[t][d] = codes with digit d in tiles < t; row `tiles` holds the totals
__global__ __launch_bounds__(256) void scan_final_kernel(const uint32_t* __restrict__ hist,
                                                         const uint64_t* __restrict__ partial,
                                                         uint64_t tiles, uint32_t nb,
                                                         uint64_t* __restrict__ excl) {
  __shared__ uint64_t s[256];
  const uint32_t tid = threadIdx.x;
  const uint32_t rows = 256 / nb, col = tid % nb, r0 = tid / nb;
  const uint32_t seg = kScanTiles / rows;  // consecutive tiles per thread
  const uint64_t blk = blockIdx.x;
  const uint64_t t0 = blk * kScanTiles + static_cast<uint64_t>(r0) * seg;
  uint64_t sum = 0;
  for (uint32_t j = 0; j < seg; ++j)
    if (t0 + j < tiles) sum += hist[(t0 + j) * nb + col];
  s[tid] = sum;
  __syncthreads();
  uint64_t run = partial[blk * nb + col];
  for (uint32_t r = 0; r < r0; ++r) run += s[r * nb + col];
  for (uint32_t j = 0; j < seg; ++j) {
    const uint64_t t = t0 + j;
    if (t >= tiles) break;
    excl[t * nb + col] = run;
    run += hist[t * nb + col];
  }
  const uint64_t last_blk = (tiles - 1) / kScanTiles;
  if (blk == last_blk && r0 == rows - 1) excl[tiles * nb + col] = run;
}

// rows[j] of excl -> out[j] (digit counts at node boundaries)
__global__ void gather_rows_kernel(const uint64_t* __restrict__ excl,
                                   const uint32_t* __restrict__ rows, uint32_t nb,
                                   uint64_t* __restrict__ out) {
  const uint64_t j = blockIdx.x;
  const uint64_t r = rows[j];
  for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) out[j * nb + d] = excl[r * nb + d];
}

// ---------------------------------------------------------------------------
// Exclusive scan of 256 values held 4 per lane by one wave (lane l holds
// entries 4l .. 4l+3); returns the exclusive prefixes in place.
template <typename V>
__device__ __forceinline__ void wave_excl_scan4(V (&v)[4], V& total) {
  const uint32_t lane = threadIdx.x & 63;
  V s = v[0] + v[1] + v[2] + v[3];
  V incl = s;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const V y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  total = __shfl(incl, 63, 64);
  V run = incl - s;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const V x = v[j];
    v[j] = run;
    run += x;
  }
}

// one wave: pc[d] = totals[0] + .. + totals[d-1] for d <= nb
__global__ __launch_bounds__(64) void digit_starts_kernel(const uint64_t* __restrict__ totals,
                                                          uint32_t nb, uint64_t* __restrict__ pc) {
  const uint32_t lane = threadIdx.x;
  uint64_t v[4], tot;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t d = 4 * lane + j;
    v[j] = d < nb ? totals[d] : 0ull;
  }
  wave_excl_scan4<uint64_t>(v, tot);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (4 * lane + j < nb) pc[4 * lane + j] = v[j];
  if (lane == 0) pc[nb] = tot;
}

// Words written by more than one workgroup of a single-node pass whose
// workgroups take spans of `span` tiles (everything else gets one plain
// store): for every span start and every group of every level, the word
// holding the group's first bit in that span (a run carried in from the
// previous span completes it), and the last word of every level.  One
// workgroup per span; the bit arrays need no other zeroing.
__global__ __launch_bounds__(256) void span_words_kernel(PassParams p) {
  __shared__ uint64_t s_e[kMaxBins + 1];  // codes of digit < x in tiles before the span
  __shared__ uint64_t s_p[kMaxBins + 1];  // global start of digit x
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const uint64_t tbeg = static_cast<uint64_t>(blockIdx.x) * p.tiles_per_wg;
  const uint32_t k = p.k, nb = 1u << k;
  if (tid < 64) {
    uint64_t v[4], tot;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t d = 4 * lane + j;
      v[j] = d < nb ? p.excl[tbeg * nb + d] : 0ull;
    }
    wave_excl_scan4<uint64_t>(v, tot);
#pragma unroll
    for (int j = 0; j < 4; ++j) s_e[4 * lane + j] = v[j];
    if (lane == 0) s_e[kMaxBins] = tot;
  }
  for (uint32_t x = tid; x <= nb; x += 256) s_p[x] = p.pc[x];
  __syncthreads();
  // level-groups x = 2^i - 1 + q (level i, group q)
  for (uint32_t x = tid; x + 1 < nb; x += 256) {
    const uint32_t i = 31u - __builtin_clz(x + 1), q = x + 1 - (1u << i);
    const uint32_t gshift = k - i, a = q << gshift, b = (q + 1) << gshift;
    const uint64_t D = (s_p[a] + s_e[b] - s_e[a]) >> 5;
    if (D < p.level_nwords[i]) p.words[p.level_word[i] + D] = 0u;
  }
  if (blockIdx.x == 0 && tid < k && p.level_nwords[tid] != 0)
    p.words[p.level_word[tid] + p.level_nwords[tid] - 1] = 0u;
}

// ---------------------------------------------------------------------------
// One pass over k levels of one tile (see the file comment).  Thread tid
// owns chunks c = r * 512 + tid (r < ROWS) of 16 bytes; chunk c holds the
// codes at tile positions [c*E, c*E + E) of the current level order.
//
// Everything a level needs besides its bits follows from the tile's digit
// counts and is tabulated for all levels before the first one: per group
// (the codes of one node of the level) its destination R in the level's bit
// array, its tile position and length, its destination words (prefix over
// the groups), and per child the partition base.  A level then costs two
// barriers: after its bits and ones counts, and after its partition.
//
// A workgroup takes p.tiles_per_wg consecutive tiles in turn.  In one node
// the runs of consecutive tiles are adjacent, so a run's last, partial word
// is carried in LDS (per level and group) into the next tile, which
// completes it: such words are written once with a plain store.  Only a
// word that a workgroup shares with another (at the ends of its span, or
// where the tiles' nodes change) is OR-ed in atomically.  One owner thread
// per group (thread 4q) handles the partial first and last word of the
// group's run; all other threads write its full words.
//
// Waves per SIMD that the tile's LDS allows (three, four, four workgroups
// of eight waves per CU for 4, 2, 1 chunk rows): the register budget
// follows.
template <int ROWS> struct PassOcc;
template <> struct PassOcc<4> { static constexpr int waves = 6; };
template <> struct PassOcc<2> { static constexpr int waves = 8; };
template <> struct PassOcc<1> { static constexpr int waves = 8; };

// Offsets of level i's entries in the per-level tables: groups (2^i per
// level), group word prefixes (2^i + 1) and children (2^(i+1)).
__device__ __forceinline__ uint32_t grp_off(uint32_t i) { return (1u << i) - 1u; }
__device__ __forceinline__ uint32_t gw_off(uint32_t i) { return (1u << i) - 1u + i; }
__device__ __forceinline__ uint32_t child_off(uint32_t i) { return (2u << i) - 2u; }
constexpr uint32_t kLevelGroups = 255;   // groups of levels 0 .. 7
constexpr uint32_t kLevelGw = 263;       // group word prefixes of levels 0 .. 7
constexpr uint32_t kLevelChildren = 510; // children of levels 0 .. 7

constexpr uint32_t kNoCarry = 0xFF;

__device__ __forceinline__ uint32_t low_mask(uint32_t nbits) {
  return nbits >= 32 ? 0xFFFFFFFFu : (1u << nbits) - 1u;
}
// 32 bits of the tile's bit string from bit sp on (zeros past its end)
template <uint32_t T>
__device__ __forceinline__ uint32_t tile_bits(const uint32_t* bits, uint32_t sp) {
  const uint32_t sw = sp >> 5;
  const uint64_t win = static_cast<uint64_t>(bits[sw]) |
                       (sw + 1 < T / 32 ? static_cast<uint64_t>(bits[sw + 1]) << 32 : 0ull);
  return static_cast<uint32_t>(win >> (sp & 31));
}

template <int CW, int ROWS>
__global__ __launch_bounds__(kPassThreads)
__attribute__((amdgpu_waves_per_eu(PassOcc<ROWS>::waves)))
void pass_kernel(PassParams p) {
  using Code = typename CodeOf<CW>::type;
  constexpr uint32_t kChunkRows = ROWS;
  constexpr uint32_t kTileChunks = kPassThreads * ROWS;
  constexpr uint32_t E = 16 / CW;                 // codes per chunk
  constexpr uint32_t T = kTileChunks * E;          // codes per tile
  constexpr uint32_t LPW = 32 / E;                 // lanes per 32-bit bit word
  static_assert(T <= 65535, "tile positions are kept in 16 bits");
  // the tile in the current level's order; rewritten in place by the
  // partition (every thread holds its chunks in registers by then)
  __shared__ uint4 s_buf[kTileChunks];
  __shared__ uint32_t s_bits[kTileChunks * E / 32]; // level bits of the tile
  __shared__ uint32_t s_hc[kMaxBins + 1];          // codes of the tile with digit < x
  __shared__ uint64_t s_ec[kMaxBins + 1];          // node's codes with digit < x in earlier tiles
  __shared__ uint64_t s_pc[kMaxBins + 1];          // global start of (node, digit x)
  __shared__ uint64_t s_wave[kPassThreads / 64];   // ones per wave, 16-bit field per row
  __shared__ uint32_t s_thr[9];
  __shared__ uint64_t s_lw[8], s_lnw[8];           // word offset and words per level
  __shared__ uint64_t s_gR[kLevelGroups];          // group destination bit
  __shared__ uint32_t s_gSL[kLevelGroups];         // group tile start | length << 16
  __shared__ uint32_t s_gw[kLevelGw];              // destination words before each group
  __shared__ uint16_t s_base[kLevelChildren];      // partition base per child
  __shared__ uint64_t s_cD[kLevelGroups];          // carried word (index into p.words)
  __shared__ uint32_t s_cV[kLevelGroups];          // its bits so far
  __shared__ uint8_t s_cLo[kLevelGroups];          // first of them; kNoCarry: none
  __shared__ uint8_t s_cHi[kLevelGroups];          // bit after the last of them
  __shared__ uint32_t s_even[8];                   // level's groups fit 512 >> i threads each

  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < kLevelGroups) s_cLo[tid] = kNoCarry;
  const uint64_t span = p.tiles_per_wg ? p.tiles_per_wg : 1;
  const uint64_t tbeg = static_cast<uint64_t>(blockIdx.x) * span;
  const uint64_t tend = tbeg + span < p.num_tiles ? tbeg + span : p.num_tiles;
  for (uint64_t t = tbeg; t < tend; ++t) {
  const TileGeo g = tile_geo(p, t, T);
  const uint32_t k = p.node_k != nullptr ? p.node_k[g.node] : p.k;
  const uint32_t nb = 1u << k;
  const uint32_t row = p.row_bins ? p.row_bins : nb;        // hist / excl row stride
  const uint32_t ts = p.thr_stride ? p.thr_stride : k + 1;  // thr entries per node

  // ---- per-tile tables (one wave each)
  if (wave == 0) {
    uint32_t v[4], tot;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t d = 4 * lane + j;
      v[j] = d < nb ? p.hist[t * row + d] : 0u;
    }
    wave_excl_scan4<uint32_t>(v, tot);
#pragma unroll
    for (int j = 0; j < 4; ++j) s_hc[4 * lane + j] = v[j];
    if (lane == 0) s_hc[kMaxBins] = tot;
  } else if (wave == 1) {
    const uint64_t f = p.node_first_tile != nullptr ? p.node_first_tile[g.node] : 0;
    uint64_t v[4], tot;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t d = 4 * lane + j;
      v[j] = d < nb ? p.excl[t * row + d] - p.excl[f * row + d] : 0ull;
    }
    wave_excl_scan4<uint64_t>(v, tot);
#pragma unroll
    for (int j = 0; j < 4; ++j) s_ec[4 * lane + j] = v[j];
    if (lane == 0) s_ec[kMaxBins] = tot;
  } else if (wave == 2) {
    const uint64_t* pc = p.pc + static_cast<uint64_t>(g.node) * (row + 1);
    for (uint32_t x = lane; x <= nb; x += 64) s_pc[x] = pc[x];
  } else if (wave == 3) {
    // read here, before the first bit-array store: a global load inside the
    // level loop would make the wave wait for its stores of earlier levels
    if (lane <= k) s_thr[lane] = p.thr[static_cast<uint64_t>(g.node) * ts + lane];
    if (lane < k) {
      s_lw[lane] = p.node_lw != nullptr ? p.node_lw[g.node * 8ull + lane] : p.level_word[lane];
      s_lnw[lane] = p.node_lnw != nullptr ? p.node_lnw[g.node * 8ull + lane] : p.level_nwords[lane];
    }
  }
  // ---- stage the tile
  const Code* src = static_cast<const Code*>(
      (p.node_remap != nullptr && p.node_remap[g.node]) ? p.codes_out : p.codes);
#pragma unroll
  for (uint32_t r = 0; r < kChunkRows; ++r) {
    const uint32_t c = r * kPassThreads + tid;
    const uint32_t p0 = c * E;
    uint4 q = make_uint4(0, 0, 0, 0);
    if (p0 + E <= g.len) q = *reinterpret_cast<const uint4*>(src + g.start + p0);
    else if (p0 < g.len) q = load_partial_chunk<Code>(src, g.start + p0, g.len - p0);
    s_buf[c] = q;
  }
  // (digits >= nb never occur, so s_hc[x] and s_ec[x] for x >= nb hold the totals)
  __syncthreads();

  // ---- level tables: wave i for level i.  Group q of level i covers digits
  // [a, b) = [q, q+1) << (k - i), its children split at m = a + (b - a) / 2;
  // it occupies tile positions [s_hc[a], s_hc[b]) of the level order and
  // level bits [R, R + len) with R = s_pc[a] + s_ec[b] - s_ec[a].  Its ones
  // (codes of child 2q+1) follow those of the earlier groups, so the
  // partition bases need no bits: child 2q takes the zeros from Oa (ones
  // before the group), child 2q+1 the ones from s_hc[m] - Oa.
  for (uint32_t i = wave; i < k; i += kPassThreads / 64) {
    const uint32_t gshift = k - i;
    const uint32_t ngroups = s_thr[i] >> gshift;
    uint32_t nw[2], ones[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t q = 2 * lane + j;
      nw[j] = 0;
      ones[j] = 0;
      if (q < ngroups) {
        const uint32_t a = q << gshift, b = (q + 1) << gshift, m = a + (1u << (gshift - 1));
        const uint32_t len = s_hc[b] - s_hc[a];
        const uint64_t R = s_pc[a] + s_ec[b] - s_ec[a];
        nw[j] = len ? static_cast<uint32_t>(((R & 31) + len + 31) >> 5) : 0u;
        ones[j] = s_hc[b] - s_hc[m];
        s_gR[grp_off(i) + q] = R;
        s_gSL[grp_off(i) + q] = s_hc[a] | (len << 16);
      }
    }
    // word prefixes in the low half, ones prefixes in the high half
    const uint64_t v = (static_cast<uint64_t>(ones[0] + ones[1]) << 32) | (nw[0] + nw[1]);
    const uint64_t incl = detail::wave_incl_scan_2x32(v);
    const uint32_t wex = static_cast<uint32_t>(incl - v);           // words before q0
    const uint32_t oex = static_cast<uint32_t>((incl - v) >> 32);   // ones before q0
    // groups of at most 3 * (512 >> i) words: written by 512 >> i threads each
    uint32_t mx = nw[0] > nw[1] ? nw[0] : nw[1];
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_xor(mx, off, 64);
      mx = mx > y ? mx : y;
    }
    if (lane == 0) s_even[i] = mx <= 3u * (kPassThreads >> i) ? 1u : 0u;
    uint32_t* gw = s_gw + gw_off(i);
    if (2 * lane < ngroups) gw[2 * lane] = wex;
    if (2 * lane + 1 < ngroups) gw[2 * lane + 1] = wex + nw[0];
    if (lane == 63) gw[ngroups] = static_cast<uint32_t>(incl);
    uint16_t* base = s_base + child_off(i);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t q = 2 * lane + j;
      const uint32_t Oa = oex + (j ? ones[0] : 0u);
      if (q < ngroups) {
        const uint32_t m = (q << gshift) + (1u << (gshift - 1));
        base[2 * q] = static_cast<uint16_t>(Oa);
        base[2 * q + 1] = static_cast<uint16_t>(s_hc[m] - Oa);
      }
    }
  }
  __syncthreads();

  for (uint32_t i = 0; i < k; ++i) {
    const uint32_t npres = s_hc[s_thr[i]];            // tile codes present at level l0+i
    const uint32_t cbit = p.dshift + k - 1 - i;        // code bit of the level
    const uint32_t gshift = k - i;                     // digit >> gshift = group
    // ---- level bits of the chunks; ones per chunk (16-bit field per row)
    uint4 v[kChunkRows];
    uint64_t packed = 0;
#pragma unroll
    for (uint32_t r = 0; r < kChunkRows; ++r) {
      const uint32_t c = r * kPassThreads + tid;
      v[r] = s_buf[c];
      uint32_t mm = chunk_mask<Code>(v[r], cbit & 63);
      const uint32_t p0 = c * E;
      if (p0 + E > npres) mm &= (p0 >= npres) ? 0u : ((1u << (npres - p0)) - 1u);
      // 32-bit words of the tile's bit string: the chunks of LPW neighbouring lanes
      uint32_t w32 = mm << (E * (lane % LPW));
#pragma unroll
      for (uint32_t off = 1; off < LPW; off <<= 1) w32 |= __shfl_xor(w32, off, 64);
      if (lane % LPW == 0) s_bits[p0 >> 5] = w32;
      packed |= static_cast<uint64_t>(__popc(mm)) << (16 * r);
    }
    // wave-local exclusive prefixes (fields stay below 2^16: no carries)
    const uint64_t incl = detail::wave_incl_scan_2x32(packed);
    const uint64_t lex = incl - packed;
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    // ones before each of my chunks: earlier rows + earlier waves of my row,
    // summed as packed 16-bit fields (one per row; no field overflows)
    uint64_t all_waves = 0, waves_before = 0;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(wave);
#pragma unroll
    for (uint32_t w = 0; w < kPassThreads / 64; ++w) {
      const uint64_t x = s_wave[w];
      all_waves += x;
      if (w < wv) waves_before += x;
    }
    uint32_t O[kChunkRows];
    {
      uint32_t tot_ones = 0;
      const uint64_t mine = waves_before + lex;
#pragma unroll
      for (uint32_t r = 0; r < kChunkRows; ++r) {
        O[r] = tot_ones + (static_cast<uint32_t>(mine >> (16 * r)) & 0xFFFFu);
        tot_ones += static_cast<uint32_t>(all_waves >> (16 * r)) & 0xFFFFu;
      }
    }
    // ---- write the level's runs.  Each thread takes destination words,
    // finds their group by binary search over the word prefixes and
    // assembles the word from the tile's bit string (plain store when the
    // word lies inside the run, atomicOr on the two edge words shared with
    // neighbouring runs).
    {
      const uint32_t ngroups = s_thr[i] >> gshift;
      const uint32_t* gw = s_gw + gw_off(i);
      const uint64_t* gR = s_gR + grp_off(i);
      const uint32_t* gSL = s_gSL + grp_off(i);
      const uint32_t total_words = gw[ngroups];
      const uint64_t lnw = s_lnw[i];
      uint32_t* words = p.words + s_lw[i];
      // full words of the runs
      if (s_even[i]) {
        // group tid >> (9 - i), its words j = tid mod 2^(9-i) + m * 2^(9-i)
        const uint32_t tsh = 9u - i, q = tid >> tsh;
        if (q < ngroups) {
          const uint32_t nwq = gw[q + 1] - gw[q];
          const uint64_t R = gR[q];
          const uint32_t sl = gSL[q];
          const uint32_t src0 = sl & 0xFFFFu, len = sl >> 16;
          const uint32_t r5 = static_cast<uint32_t>(R) & 31u;
          for (uint32_t j = tid & ((1u << tsh) - 1u); j < nwq; j += 1u << tsh) {
            if ((j == 0 && r5 != 0) || r5 + len < 32u * (j + 1)) continue;  // partial: owner's
            const uint64_t D = (R >> 5) + j;
            if (D >= lnw) atomicOr(p.err, 1u);
            else words[D] = tile_bits<T>(s_bits, src0 + 32u * j - r5);
          }
        }
      } else
      for (uint32_t w = tid; w < total_words; w += kPassThreads) {
        uint32_t lo = 0, hi = ngroups, wlo = 0;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          const uint32_t x = gw[mid];
          if (x <= w) { lo = mid; wlo = x; } else { hi = mid; }
        }
        const uint64_t R = gR[lo];
        const uint32_t sl = gSL[lo];
        const uint32_t src0 = sl & 0xFFFFu, len = sl >> 16;
        const uint32_t j = w - wlo;                   // word of the run
        const uint32_t r5 = static_cast<uint32_t>(R) & 31u;
        if ((j == 0 && r5 != 0) || r5 + len < 32u * (j + 1)) continue;  // partial: owner's
        const uint64_t D = (R >> 5) + j;
        const uint32_t sp = src0 + 32u * j - r5;
        if (D >= lnw) atomicOr(p.err, 1u);
        else words[D] = tile_bits<T>(s_bits, sp);
      }
      // partial first and last words: group q's owner, thread 4q
      if ((tid & 3u) == 0 && (tid >> 2) < ngroups) {
        const uint32_t q = tid >> 2;
        const uint32_t sl = gSL[q];
        const uint32_t src0 = sl & 0xFFFFu, len = sl >> 16;
        if (len != 0) {
          const uint32_t ci = grp_off(i) + q;
          const uint64_t R = gR[q];
          const uint32_t r5 = static_cast<uint32_t>(R) & 31u;
          const uint64_t D0 = R >> 5;
          const uint64_t aD0 = s_lw[i] + D0;
          uint32_t clo = s_cLo[ci], chi = 0, cur = 0;
          if (clo != kNoCarry) {
            if (s_cD[ci] == aD0 && s_cHi[ci] == r5) {
              cur = s_cV[ci];  // this run continues the carried bits
            } else {           // not continued here: the word is shared
              if (s_cV[ci]) atomicOr(&p.words[s_cD[ci]], s_cV[ci]);
              clo = kNoCarry;
            }
          }
          if (r5 != 0 || len < 32u) {
            // first word: the run's bits [r5, min(32, r5 + len)) join the carry
            const uint32_t hn = len < 32u - r5 ? len : 32u - r5;
            cur |= (tile_bits<T>(s_bits, src0) & low_mask(hn)) << r5;
            if (clo == kNoCarry) clo = r5;
            if (r5 + len >= 32u) {  // complete
              if (D0 >= lnw) atomicOr(p.err, 1u);
              else if (clo == 0) words[D0] = cur;
              else if (cur) atomicOr(&words[D0], cur);
              clo = kNoCarry;
            } else {
              chi = r5 + len;
            }
          }
          const uint32_t e5 = static_cast<uint32_t>(R + len) & 31u;
          if (r5 + len > 32u && e5 != 0) {
            // last word: bits [0, e5), carried into the next tile
            cur = tile_bits<T>(s_bits, src0 + len - e5) & low_mask(e5);
            clo = 0;
            chi = e5;
            s_cD[ci] = s_lw[i] + ((R + len) >> 5);
          } else if (clo != kNoCarry) {
            s_cD[ci] = aD0;
          }
          if (clo != kNoCarry && ((R + len) >> 5) >= lnw) {
            atomicOr(p.err, 1u);
            clo = kNoCarry;
          }
          s_cV[ci] = cur;
          s_cLo[ci] = static_cast<uint8_t>(clo);
          s_cHi[ci] = static_cast<uint8_t>(chi);
        }
      }
    }
    const bool partition = (i + 1 < k) || (p.dst != nullptr);
    if (!partition) break;
    // ---- partition: child c of group c >> 1 takes its zeros (c even) or
    // ones (c odd) in order from its base (level tables)
    const uint16_t* s_bs = s_base + child_off(i);
    Code* both = reinterpret_cast<Code*>(s_buf);
    const uint32_t cwidth = i + 1;  // child index = top i+1 digit bits
    if (npres == T && i == 0) {
      // one group, two children: the bases stay in registers
      const uint32_t base0 = s_bs[0], base1 = s_bs[1];
#pragma unroll
      for (uint32_t r = 0; r < kChunkRows; ++r) {
        const uint32_t p0 = (r * kPassThreads + tid) * E;
        uint32_t o = O[r], z = p0 - O[r];
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) {
          const uint32_t b = child_of<Code>(v[r], j, cbit, 1);
          both[(b ? base1 + o : base0 + z)] = chunk_elem<Code>(v[r], j);
          o += b;
          z += b ^ 1u;
        }
      }
    } else if (npres == T) {
      // every code of the tile is present at this level: no per-code tests
#pragma unroll
      for (uint32_t r = 0; r < kChunkRows; ++r) {
        const uint32_t p0 = (r * kPassThreads + tid) * E;
        uint32_t child[E], base[E];
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) child[j] = child_of<Code>(v[r], j, cbit, cwidth);
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) base[j] = s_bs[child[j]];
        uint32_t o = O[r], z = p0 - O[r];
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) {
          const uint32_t b = child[j] & 1u;
          both[base[j] + (b ? o : z)] = chunk_elem<Code>(v[r], j);
          o += b;
          z += b ^ 1u;
        }
      }
    } else {
#pragma unroll
      for (uint32_t r = 0; r < kChunkRows; ++r) {
        const uint32_t p0 = (r * kPassThreads + tid) * E;
        if (p0 >= npres) continue;
        uint32_t o = O[r], z = p0 - O[r];
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) {
          if (p0 + j < npres) {
            const uint32_t child = child_of<Code>(v[r], j, cbit, cwidth);
            const uint32_t b = child & 1u;
            both[s_bs[child] + (b ? o : z)] = chunk_elem<Code>(v[r], j);
            o += b;
            z += b ^ 1u;
          }
        }
      }
    }
    __syncthreads();
  }

  // ---- codes of level l0+k for the next pass, grouped by digit
  if (p.dst != nullptr) {
    // s_pc becomes the output offset per digit: the (node, digit) start in
    // the padded output + its codes in earlier tiles - its first tile position
    const uint64_t* ob = p.obase + static_cast<uint64_t>(g.node) * nb;
    for (uint32_t d = tid; d < nb; d += kPassThreads)
      s_pc[d] = ob[d] + (s_ec[d + 1] - s_ec[d]) - s_hc[d];
    __syncthreads();
    const uint32_t nout = s_hc[s_thr[k]];
    const Code* cb = reinterpret_cast<const Code*>(s_buf);
    Code* dst = static_cast<Code*>(p.dst);
    const uint32_t dmask = nb - 1;
    for (uint32_t x = tid; x < nout; x += kPassThreads) {
      const Code c = cb[x];
      const uint32_t d = static_cast<uint32_t>(static_cast<uint64_t>(c) >> p.dshift) & dmask;
      const uint64_t o = s_pc[d] + x;
      if (o < p.dst_cap) dst[o] = c;
      else atomicOr(p.err, 1u);
    }
  }
  __syncthreads();  // the next tile rewrites the tables and the buffer
  }  // tiles of the span
  // words still carried are shared with the next span
  if (tid < kLevelGroups && s_cLo[tid] != kNoCarry && s_cV[tid] != 0)
    atomicOr(&p.words[s_cD[tid]], s_cV[tid]);
}

}  // namespace

namespace {
std::atomic<uint32_t> g_pass_span{0};  // 0: automatic
}  // namespace

void set_pass_tiles_per_wg(uint32_t span) { g_pass_span.store(span); }

uint32_t pass_span(uint64_t num_tiles) {
  const uint32_t forced = g_pass_span.load();
  if (forced != 0) return forced;
  // spans of up to 16 tiles while the grid still gives every CU about four
  // rounds of resident workgroups (256 CUs, 3 per CU); at 6e9 symbols 16
  // tiles took the two-level pass from 3.99 to 3.79 ms and left the
  // eight-level pass as it was (8 -> 16: 15.06 -> 15.15 ms)
  const uint64_t span = num_tiles / (256 * 3 * 4);
  return static_cast<uint32_t>(span < 1 ? 1 : span > 16 ? 16 : span);
}

void launch_span_words(const PassParams& p, hipStream_t s) {
  if (p.num_tiles == 0) return;
  if (p.tiles_per_wg == 0 || p.tile_start != nullptr || p.k > 8)
    throw Error(ECL_ERROR_INTERNAL, "span words need a single-node pass with a set span");
  const uint64_t spans = (p.num_tiles + p.tiles_per_wg - 1) / p.tiles_per_wg;
  hipLaunchKernelGGL(span_words_kernel, dim3(static_cast<uint32_t>(spans)), dim3(256), 0, s, p);
  ECL_HIP(hipGetLastError());
}

uint64_t pass_tile_codes(int cw, uint32_t rows) {
  return kPassThreads * rows * (16u / static_cast<uint32_t>(cw));
}

void launch_tile_hist(const PassParams& p, hipStream_t s) {
  if (p.num_tiles == 0) return;
  if (p.num_tiles > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "text too large");
  ProfScope ps(K_TILE_HIST, s);
  const dim3 grid(static_cast<uint32_t>(p.num_tiles)), block(kPassThreads);
  const bool remap = p.codes_out != nullptr;
  if (p.k <= 2 && p.dshift == 0 && p.src_esize == 1 && p.cw == 1 && !remap && p.map_kind == 0 &&
      p.node_remap == nullptr) {
    hipLaunchKernelGGL(tile_hist_small_kernel, grid, block, 0, s, p);
    ECL_HIP(hipGetLastError());
    return;
  }
#define ECL_TH(TIN, CWV, MK)                                                                 \
  do {                                                                                       \
    if (remap) hipLaunchKernelGGL((tile_hist_kernel<TIN, CWV, MK, true>), grid, block, 0, s, p); \
    else hipLaunchKernelGGL((tile_hist_kernel<TIN, CWV, MK, false>), grid, block, 0, s, p);      \
  } while (0)
#define ECL_TH_MK(TIN, CWV)                          \
  do {                                               \
    if (p.map_kind == 0) ECL_TH(TIN, CWV, 0);        \
    else if (p.map_kind == 1) ECL_TH(TIN, CWV, 1);   \
    else ECL_TH(TIN, CWV, 2);                        \
  } while (0)
#define ECL_TH_MK02(TIN, CWV)                        \
  do {                                               \
    if (p.map_kind == 0) ECL_TH(TIN, CWV, 0);        \
    else ECL_TH(TIN, CWV, 2);                        \
  } while (0)
  if (p.src_esize == 1) {
    ECL_TH_MK(uint8_t, 1);
  } else if (p.src_esize == 2) {
    if (p.cw == 1) ECL_TH_MK(uint16_t, 1); else ECL_TH_MK(uint16_t, 2);
  } else if (p.src_esize == 4) {
    if (p.cw == 1) ECL_TH_MK02(uint32_t, 1);
    else if (p.cw == 2) ECL_TH_MK02(uint32_t, 2);
    else ECL_TH_MK02(uint32_t, 4);
  } else {
    if (p.cw == 1) ECL_TH_MK02(uint64_t, 1);
    else if (p.cw == 2) ECL_TH_MK02(uint64_t, 2);
    else if (p.cw == 4) ECL_TH_MK02(uint64_t, 4);
    else ECL_TH_MK02(uint64_t, 8);
  }
#undef ECL_TH_MK02
#undef ECL_TH_MK
#undef ECL_TH
  ECL_HIP(hipGetLastError());
}

uint64_t scan_partial_entries(uint64_t tiles, uint32_t k) {
  return ((tiles + kScanTiles - 1) / kScanTiles) << k;
}

void launch_tile_scan(const PassParams& p, hipStream_t s) {
  if (p.num_tiles == 0) return;
  ProfScope ps(K_TILE_SCAN, s);
  const uint32_t nb = p.row_bins ? p.row_bins : (1u << p.k);  // row stride
  const uint64_t nblk = (p.num_tiles + kScanTiles - 1) / kScanTiles;
  hipLaunchKernelGGL(scan_partial_kernel, dim3(static_cast<uint32_t>(nblk)), dim3(256), 0, s,
                     p.hist, p.num_tiles, nb, p.partial);
  ECL_HIP(hipGetLastError());
  hipLaunchKernelGGL(scan_blocks_kernel, dim3(nb), dim3(256), 0, s, p.partial, nblk, nb);
  ECL_HIP(hipGetLastError());
  hipLaunchKernelGGL(scan_final_kernel, dim3(static_cast<uint32_t>(nblk)), dim3(256), 0, s,
                     p.hist, p.partial, p.num_tiles, nb, p.excl);
  ECL_HIP(hipGetLastError());
}

void launch_gather_rows(const uint64_t* excl, const uint32_t* rows, uint64_t nrows, uint32_t nb,
                        uint64_t* out, hipStream_t s) {
  if (nrows == 0) return;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(static_cast<uint32_t>(nrows)), dim3(256), 0, s,
                     excl, rows, nb, out);
  ECL_HIP(hipGetLastError());
}

void launch_digit_starts(const uint64_t* totals, uint32_t nb, uint64_t* pc, hipStream_t s) {
  if (nb > kMaxBins) throw Error(ECL_ERROR_INTERNAL, "too many digits");
  hipLaunchKernelGGL(digit_starts_kernel, dim3(1), dim3(64), 0, s, totals, nb, pc);
  ECL_HIP(hipGetLastError());
}

void launch_pass(const PassParams& params, hipStream_t s) {
  if (params.num_tiles == 0) return;
  ProfScope ps(K_PASS, s);
  PassParams p = params;
  if (p.tiles_per_wg == 0) p.tiles_per_wg = pass_span(p.num_tiles);
  const uint64_t span = p.tiles_per_wg;
  const dim3 grid(static_cast<uint32_t>((p.num_tiles + span - 1) / span)), block(kPassThreads);
  if (p.tile_codes != pass_tile_codes(p.cw, p.tile_rows))
    throw Error(ECL_ERROR_INTERNAL, "tile size does not match the tile rows");
#define ECL_PASS_ROWS(CWV)                                                                  \
  do {                                                                                      \
    if (p.tile_rows == 4) hipLaunchKernelGGL((pass_kernel<CWV, 4>), grid, block, 0, s, p);  \
    else if (p.tile_rows == 2) hipLaunchKernelGGL((pass_kernel<CWV, 2>), grid, block, 0, s, p); \
    else hipLaunchKernelGGL((pass_kernel<CWV, 1>), grid, block, 0, s, p);                   \
  } while (0)
  switch (p.cw) {
    case 1: ECL_PASS_ROWS(1); break;
    case 2: ECL_PASS_ROWS(2); break;
    case 4: ECL_PASS_ROWS(4); break;
    default: ECL_PASS_ROWS(8); break;
  }
#undef ECL_PASS_ROWS
  ECL_HIP(hipGetLastError());
}

}  // namespace lib
}  // namespace ecl
