// kernels_build.hip -- construction kernels for gfx950 (wave64).
//
//  * histogram_kernel: LDS-privatised u32 bins, remaps to codes in the same
//    pass (replaces computeGlobalHistogramKernel, tree/wavelet_tree.cuh:
//    170-227).
//  * level_kernel: one pass per level.  Writes the level's bits with one
//    ballot per 64 symbols and scatters every symbol to its position in the
//    next level's order (stable one-bit partition inside each node).  The
//    global ones-prefix needed by the partition comes from a decoupled
//    look-back over tiles.  Replaces the per-level cub::DeviceRadixSort +
//    fillLevelKernel pair (wavelet_tree.cuh:241-323, 806-818).
//  * rs_build_kernel: one pass per bit array.  L2 counts, L1 prefix (look-
//    back), total ones and both select-sample streams in a single read of
//    the bits.  Replaces calculateL2EntriesKernel + cub InclusiveSum +
//    addNumOnesKernel + calculateSelectSamplesKernel (src/bitarray/
//    rank_select.cu:412-591).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace ecl {
namespace lib {
namespace {

using detail::ballot64;
using detail::lane_id;
using detail::mbcnt64;

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

__device__ __forceinline__ void status_store(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t status_load(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back, run by all 64 lanes of one wave.  Publishes this
// tile's aggregate, then reads the status of up to 64*J predecessors per
// step (lane i covers predecessors 8i+1 .. 8i+J nearest-first) until it
// meets an inclusive prefix, publishes its own inclusive prefix and returns
// the exclusive one.  Publication can only advance by one window per
// status-load latency, so the window is wide and the tiles are large.
// Tiles take their ids in launch order from an atomic counter: every
// predecessor is resident or finished, so the wait ends.
template <int J>
__device__ uint64_t lookback(uint64_t* status, uint64_t tile, uint64_t agg) {
  const uint32_t lane = lane_id();
  if (tile == 0) {
    if (lane == 0) status_store(&status[0], kFlagPre | agg);
    return 0;
  }
  if (lane == 0) status_store(&status[tile], kFlagAgg | agg);
  uint64_t excl = 0;
  int64_t top = static_cast<int64_t>(tile) - 1;
  while (true) {
    uint64_t v[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int64_t idx = top - static_cast<int64_t>(lane) * J - j;
      v[j] = idx >= 0 ? status_load(&status[idx]) : kFlagPre;
    }
    int first_pre = J;
    bool empty = false;
#pragma unroll
    for (int j = J - 1; j >= 0; --j)
      if ((v[j] >> 62) == 2) first_pre = j;
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (j <= first_pre && (v[j] >> 62) == 0) empty = true;
    const uint64_t pre_mask = ballot64(first_pre < J);
    const uint32_t stop = pre_mask ? static_cast<uint32_t>(__builtin_ctzll(pre_mask)) : 64u;
    if (ballot64(lane <= stop && empty)) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint64_t sum = 0;
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (lane < stop || (lane == stop && j <= first_pre)) sum += v[j] & kValMask;
    excl += detail::wave_sum_u64(sum);
    if (pre_mask) break;
    top -= 64 * J;
  }
  if (lane == 0) status_store(&status[tile], kFlagPre | (excl + agg));
  return excl;
}
constexpr int kLookbackWords = 8;  // window of 512 tiles

template <typename T>
__device__ __forceinline__ uint64_t load_elem(const T* p, uint64_t i) {
  return static_cast<uint64_t>(p[i]);
}

// ---------------------------------------------------------------------------
__global__ void fill_u64_kernel(uint64_t* p, uint64_t n, uint64_t v) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       i < n; i += stride)
    p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       i < n; i += stride)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// Histogram.  Each wave owns `copies` replicas of the bins in LDS (lane l
// uses replica l % copies) so that small alphabets do not serialise on a
// few addresses; large alphabets fall back to global u64 atomics.
// map_kind: 0 identity, 1 LUT, 2 binary search over the sorted alphabet.
template <typename T, int MapKind, bool UseLds>
__global__ __launch_bounds__(256) void histogram_kernel(HistParams p,
                                                        uint32_t copies) {
  extern __shared__ uint32_t s_bins[];
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid >> 6, lane = tid & 63;
  const uint64_t sigma = p.sigma;
  uint32_t* my_bins = nullptr;
  if constexpr (UseLds) {
    const uint32_t total = 4u * copies * static_cast<uint32_t>(sigma);
    for (uint32_t i = tid; i < total; i += 256) s_bins[i] = 0;
    __syncthreads();
    my_bins = s_bins + (wave * copies + (lane % copies)) * sigma;
  }
  const T* text = static_cast<const T*>(p.text);
  T* out = static_cast<T*>(p.codes_out);
  const T* alpha = static_cast<const T*>(p.alphabet);
  auto count = [&](uint64_t x) -> T {
    uint64_t c;
    if constexpr (MapKind == 0) {
      c = x;
    } else if constexpr (MapKind == 1) {
      c = p.lut[x];
    } else {
      uint64_t lo = 0, hi = sigma;  // lower_bound
      while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (static_cast<uint64_t>(alpha[mid]) < x) lo = mid + 1; else hi = mid;
      }
      c = lo;
    }
    if (c >= sigma) c = sigma - 1;  // symbol outside the alphabet
    if constexpr (UseLds) {
      atomicAdd(&my_bins[c], 1u);
    } else {
      atomicAdd(reinterpret_cast<unsigned long long*>(&p.hist[c]), 1ull);
    }
    return static_cast<T>((p.pow2 || c < p.codes_start) ? c : p.codes[c - p.codes_start]);
  };
  // 16-byte vectors: E symbols per lane per step (text is 16 B aligned)
  constexpr uint32_t E = 16 / sizeof(T);
  const uint64_t nvec = p.n / E;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256u;
  const uint4* tv = reinterpret_cast<const uint4*>(text);
  uint4* ov = reinterpret_cast<uint4*>(out);
  for (uint64_t i = blockIdx.x * 256ull + tid; i < nvec; i += stride) {
    union { uint4 v; T e[E]; } in, res;
    in.v = tv[i];
#pragma unroll
    for (uint32_t k = 0; k < E; ++k) res.e[k] = count(static_cast<uint64_t>(in.e[k]));
    if (out != nullptr) ov[i] = res.v;
  }
  for (uint64_t i = nvec * E + blockIdx.x * 256ull + tid; i < p.n; i += stride) {
    const T code = count(static_cast<uint64_t>(text[i]));
    if (out != nullptr) out[i] = code;
  }
  if constexpr (UseLds) {
    __syncthreads();
    const uint32_t reps = 4u * copies;
    for (uint64_t c = tid; c < sigma; c += 256) {
      uint64_t sum = 0;
      for (uint32_t r = 0; r < reps; ++r) sum += s_bins[r * sigma + c];
      if (sum) atomicAdd(reinterpret_cast<unsigned long long*>(&p.hist[c]),
                         static_cast<unsigned long long>(sum));
    }
  }
}

// ---------------------------------------------------------------------------
// Level pass.  One workgroup (256 threads) per tile of 32 KiB of symbols
// (2 tiles per 65536-bit L1 block for u8, 4 for u16, 8 for u32, 16 for u64),
// staged in LDS so the level is read from HBM once.
//  phase 1: 16-byte vector loads (chunk c = E symbols), bit mask of every
//           chunk packed in registers, u64 bit words assembled over the
//           lanes of one word with xor-shuffles and stored; chunk ones
//           counts go to LDS and are scanned (in-tile prefixes).
//  look-back on the tile's ones count -> global ones before the tile; the
//           first tile of an L1 block writes the L1 entry, every tile writes
//           its L2 entries relative to the block start.
//  phase 2: each thread walks its chunks from LDS: select samples where a
//           multiple of 4096 ones / zeros falls, and the stable one-bit
//           partition dest = base(node, bit) + rank_bit(position).
constexpr uint32_t kTileBytes = 32768;
constexpr int kLevelLookbackWords = 16;  // window of 1024 tiles

template <typename T>
struct LevelCfg {
  static constexpr uint32_t E = 16 / sizeof(T);               // symbols per chunk
  static constexpr uint32_t TILE = kTileBytes / sizeof(T);     // symbols per tile
  static constexpr uint32_t CHUNKS = TILE / E;                 // 2048
  static constexpr uint32_t ITERS = CHUNKS / 256;              // 8
  static constexpr uint32_t TILES_PER_L1 = 65536 / TILE;       // 2..16
  static constexpr uint32_t LPW = 64 / E;                      // lanes per u64 word
};

// bit `bit` of each of the E symbols of a 16-byte chunk, symbol k at bit k
template <typename T>
__device__ __forceinline__ uint32_t chunk_mask(const uint4& v, uint32_t bit) {
  const uint32_t d[4] = {v.x, v.y, v.z, v.w};
  uint32_t m = 0;
  if constexpr (sizeof(T) == 1) {
    // four bytes of a dword: isolate the bit of each byte, gather the four
    // bits into a nibble with one multiply (no carries: the partial
    // products land on distinct bit positions)
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

// symbol k of a 16-byte chunk, by shifts (k is a constant after unrolling)
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
// the same chunk assembled from scalar loads of the symbols below n
template <typename T>
__device__ __forceinline__ uint4 load_partial_chunk(const T* src, uint64_t s0, uint64_t n) {
  uint32_t d[4] = {0, 0, 0, 0};
  constexpr uint32_t E = 16 / sizeof(T);
#pragma unroll
  for (uint32_t k = 0; k < E; ++k) {
    if (s0 + k < n) {
      const uint64_t x = static_cast<uint64_t>(src[s0 + k]);
      const uint32_t bitpos = k * 8 * static_cast<uint32_t>(sizeof(T));
      if constexpr (sizeof(T) == 8) {
        d[2 * k] = static_cast<uint32_t>(x);
        d[2 * k + 1] = static_cast<uint32_t>(x >> 32);
      } else {
        d[bitpos >> 5] |= static_cast<uint32_t>(x) << (bitpos & 31);
      }
    }
  }
  return make_uint4(d[0], d[1], d[2], d[3]);
}

template <typename T>
__global__ __launch_bounds__(256) void level_kernel(LevelParams p) {
  using C = LevelCfg<T>;
  __shared__ uint4 s_tile[C::CHUNKS];        // 32 KiB of symbols
  __shared__ uint16_t s_pre[C::CHUNKS];      // in-tile exclusive ones prefix per chunk
  __shared__ uint32_t s_wave[4];
  __shared__ uint64_t s_tile_id, s_G;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63, wave = tid >> 6;
  if (tid == 0)
    s_tile_id = atomicAdd(reinterpret_cast<unsigned long long*>(p.counter), 1ull);
  __syncthreads();
  const uint64_t tile = s_tile_id;
  const uint64_t tbase = tile * C::TILE;  // first symbol of the tile
  const T* src = static_cast<const T*>(p.src);
  const uint64_t ntiles = (p.n + C::TILE - 1) / C::TILE;

  // ---- phase 1: load, stage, bit words, chunk counts
  uint4 v[C::ITERS];
#pragma unroll
  for (uint32_t r = 0; r < C::ITERS; ++r) {
    const uint64_t s0 = tbase + static_cast<uint64_t>(r * 256 + tid) * C::E;
    if (s0 + C::E <= p.n) {
      v[r] = reinterpret_cast<const uint4*>(src + s0)[0];
    } else {
      v[r] = load_partial_chunk<T>(src, s0, p.n);
    }
  }
#pragma unroll
  for (uint32_t r = 0; r < C::ITERS; ++r) {
    const uint32_t c = r * 256 + tid;
    const uint64_t s0 = tbase + static_cast<uint64_t>(c) * C::E;
    s_tile[c] = v[r];
    uint32_t m = chunk_mask<T>(v[r], p.bit);
    if (s0 + C::E > p.n) m &= (s0 >= p.n) ? 0u : ((1u << (p.n - s0)) - 1u);
    s_pre[c] = static_cast<uint16_t>(__popc(m));
    uint64_t w = static_cast<uint64_t>(m) << (C::E * (lane % C::LPW));
#pragma unroll
    for (uint32_t off = 1; off < C::LPW; off <<= 1) w |= __shfl_xor(w, off, 64);
    const uint64_t widx = (tbase + static_cast<uint64_t>(c) * C::E) / 64;
    if (lane % C::LPW == 0 && widx < p.nwords64) p.words64[widx] = w;
  }
  __syncthreads();
  // in-tile exclusive scan of the chunk counts: thread t owns chunks 8t..8t+7
  uint32_t cnt[8], sum = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) { cnt[k] = s_pre[tid * 8 + k]; sum += cnt[k]; }
  const uint32_t incl = detail::wave_incl_scan_u32(sum);
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = incl - sum, ttot = 0;
#pragma unroll
  for (uint32_t w = 0; w < 4; ++w) {
    before += (w < wave) ? s_wave[w] : 0u;
    ttot += s_wave[w];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) { s_pre[tid * 8 + k] = static_cast<uint16_t>(before); before += cnt[k]; }

  // ---- look-back, L1 entry, global prefixes of the L2 blocks.  The look-back
  // communicates only through its status words (value and flag in one atomic
  // word), so no fence is needed; L2 entries are made block-relative by a
  // separate pass (l2_fixup_kernel) instead of reading other tiles' counts.
  if (wave == 0) {
    const uint64_t g = lookback<kLevelLookbackWords>(p.status, tile, ttot);
    if (lane == 0) s_G = g;
  }
  __syncthreads();
  const uint64_t G = s_G;
  if (tid == 0) {
    if (tile % C::TILES_PER_L1 == 0) {
      const uint64_t k = tile / C::TILES_PER_L1;
      if (k < p.num_l1) p.l1[k] = G;
    }
    if (tile == ntiles - 1) *p.total_ones = G + ttot;
  }
  {  // ones before each 512-symbol block of this tile (512/E chunks apart)
    constexpr uint32_t kBlocks = C::TILE / 512;
    if (tid < kBlocks) {
      const uint64_t k = tbase / 512 + tid;
      if (k < p.num_l2) p.l2_global[k] = G + s_pre[tid * (512 / C::E)];
    }
  }

  // ---- phase 2: samples and scatter, one chunk at a time per thread
  T* dst = static_cast<T*>(p.dst);
  uint64_t cached_pfx = ~0ull, base0 = 0, base1 = 0;
  for (uint32_t r = 0; r < C::ITERS; ++r) {
    const uint32_t c = r * 256 + tid;
    const uint64_t s0 = tbase + static_cast<uint64_t>(c) * C::E;
    if (s0 >= p.n) break;
    const uint4 q = s_tile[c];
    const uint32_t valid = (p.n - s0 < C::E) ? static_cast<uint32_t>(p.n - s0) : C::E;
    uint32_t m = chunk_mask<T>(q, p.bit);
    if (valid < C::E) m &= (1u << valid) - 1u;
    const uint32_t pc = __popc(m);
    uint64_t r1 = G + s_pre[c];  // ones before the chunk
    if (p.cap) {
      const uint64_t k1 = (r1 + pc) / detail::kSampleRate;
      if (k1 > r1 / detail::kSampleRate)
        p.samples[k1 - 1] =
            s0 + detail::select_in_word(m, static_cast<uint32_t>(k1 * detail::kSampleRate - r1));
      const uint64_t zb = s0 - r1;
      const uint64_t k0 = (zb + valid - pc) / detail::kSampleRate;
      if (k0 > zb / detail::kSampleRate) {
        const uint32_t zm = ~m & ((valid >= 32) ? ~0u : ((1u << valid) - 1u));
        p.samples[p.cap - k0] =
            s0 + detail::select_in_word(zm, static_cast<uint32_t>(k0 * detail::kSampleRate - zb));
      }
    }
    if (dst != nullptr) {
      uint64_t r0 = s0 - r1;
#pragma unroll
      for (uint32_t k = 0; k < C::E; ++k) {
        if (k >= valid) break;
        const T x = chunk_elem<T>(q, k);
        const uint64_t pfx = detail::shr64(static_cast<uint64_t>(x), p.pshift);
        if (pfx != cached_pfx) {
          cached_pfx = pfx;
          if (pfx < p.num_prefixes) {
            base0 = p.table[2 * pfx];
            base1 = p.table[2 * pfx + 1];
          } else {
            base0 = base1 = detail::kSkip;
          }
        }
        const bool b = (m >> k) & 1u;
        const uint64_t base = b ? base1 : base0;
        const uint64_t dest = base + (b ? r1 : r0);
        if (base != detail::kSkip && dest < p.n_next) dst[dest] = x;
        r1 += b;
        r0 += !b;
      }
    }
  }
}

// L2 entry = ones before the 512-bit block minus ones before its L1 block
__global__ void l2_fixup_kernel(const uint64_t* __restrict__ l2_global,
                                const uint64_t* __restrict__ l1, uint16_t* __restrict__ l2,
                                uint64_t num_l2) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < num_l2;
       k += stride)
    l2[k] = static_cast<uint16_t>(l2_global[k] - l1[k >> 7]);
}

// ---------------------------------------------------------------------------
// RankSelect directory, one L1 block (1024 u64 words) per workgroup of 256
// threads, 4 words per thread.
__global__ __launch_bounds__(256) void rs_build_kernel(RsBuildParams p) {
  __shared__ uint64_t s_tile;
  __shared__ uint64_t s_excl;
  __shared__ uint32_t s_wave[4];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63, wave = tid >> 6;
  if (tid == 0)
    s_tile = atomicAdd(reinterpret_cast<unsigned long long*>(p.counter), 1ull);
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t nwords = (p.m + 63) / 64;
  const uint64_t w0 = tile * 1024 + static_cast<uint64_t>(tid) * 4;
  uint64_t w[4];
  uint32_t c = 0;
  if (w0 + 3 < nwords) {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p.words64 + w0);
    const ulonglong2 a = q[0], b2 = q[1];
    w[0] = a.x; w[1] = a.y; w[2] = b2.x; w[3] = b2.y;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (w0 + k < nwords) ? p.words64[w0 + k] : 0ull;
  }
  const uint32_t tail = static_cast<uint32_t>(p.m & 63);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (tail && w0 + k == nwords - 1) w[k] &= (1ull << tail) - 1;
    c += static_cast<uint32_t>(__popcll(w[k]));
  }
  // block exclusive scan of c
  const uint32_t incl = detail::wave_incl_scan_u32(c);
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t wexcl = 0, ttot = 0;
#pragma unroll
  for (uint32_t ww = 0; ww < 4; ++ww) {
    const uint32_t x = s_wave[ww];
    wexcl += (ww < wave) ? x : 0u;
    ttot += x;
  }
  const uint32_t excl = wexcl + incl - c;  // ones in the L1 block before me
  // L2 entry of 512-bit block j = threads 2j, 2j+1
  if ((tid & 1) == 0) {
    const uint64_t k = tile * 128 + (tid >> 1);
    if (k < p.num_l2) p.l2[k] = static_cast<uint16_t>(excl);
  }
  if (wave == 0) {
    const uint64_t g = lookback<kLookbackWords>(p.status, tile, ttot);
    if (lane == 0) s_excl = g;
  }
  __syncthreads();
  const uint64_t g = s_excl;
  if (tid == 0 && tile < p.num_l1) p.l1[tile] = g;
  if (tid == 0 && tile == p.num_l1 - 1) *p.total_ones = g + ttot;

  // select samples: at most one multiple of 4096 ones (and of zeros) falls
  // inside a thread's 256 bits
  const uint64_t b0 = w0 * 64;
  if (p.cap == 0 || b0 >= p.m) return;
  const uint64_t b1 = (b0 + 256 < p.m) ? b0 + 256 : p.m;
  const uint64_t ones_before = g + excl;
  const uint64_t k1 = (ones_before + c) / detail::kSampleRate;
  if (k1 > ones_before / detail::kSampleRate) {
    uint32_t r = static_cast<uint32_t>(k1 * detail::kSampleRate - ones_before);
    uint64_t pos = b0;
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
      const uint32_t pc = static_cast<uint32_t>(__popcll(w[k]));
      if (r <= pc) {
        pos += detail::select_in_word(w[k], r);
        break;
      }
      r -= pc;
      pos += 64;
    }
    p.samples[k1 - 1] = pos;
  }
  const uint64_t zeros_before = b0 - ones_before;
  const uint64_t zc = (b1 - b0) - c;
  const uint64_t k0 = (zeros_before + zc) / detail::kSampleRate;
  if (k0 > zeros_before / detail::kSampleRate) {
    uint32_t r = static_cast<uint32_t>(k0 * detail::kSampleRate - zeros_before);
    uint64_t pos = b0;
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
      uint64_t z = ~w[k];
      const uint64_t wb = b0 + 64ull * k;
      if (wb + 64 > p.m) z &= (p.m > wb) ? ((1ull << (p.m - wb)) - 1) : 0ull;
      const uint32_t pc = static_cast<uint32_t>(__popcll(z));
      if (r <= pc) {
        pos += detail::select_in_word(z, r);
        break;
      }
      r -= pc;
      pos += 64;
    }
    p.samples[p.cap - k0] = pos;
  }
}

// ---------------------------------------------------------------------------
template <typename T>
__global__ void gen_text_kernel(T* out, uint64_t n, uint64_t seed,
                                uint64_t sigma, const uint64_t* thr) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       i < n; i += stride) {
    const uint64_t h = detail::hash64(seed, i);
    uint64_t s;
    if (thr == nullptr) {
      s = detail::bounded64(h, sigma);
    } else {
      uint64_t lo = 0, hi = sigma - 1;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (h < thr[mid]) hi = mid; else lo = mid + 1;
      }
      s = lo;
    }
    out[i] = static_cast<T>(s);
  }
}

// one u32 word per thread; words beyond nbits are cleared
__global__ void gen_bits_kernel(uint32_t* words, uint64_t nbits,
                                uint64_t nwords_total, uint64_t seed,
                                uint64_t threshold) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t w = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       w < nwords_total; w += stride) {
    uint32_t x = 0;
    const uint64_t base = w * 32;
    if (base < nbits) {
#pragma unroll 4
      for (uint32_t b = 0; b < 32; ++b)
        if (base + b < nbits && detail::hash64(seed, base + b) < threshold)
          x |= 1u << b;
    }
    words[w] = x;
  }
}

template <typename T>
__global__ void presence_kernel(const T* text, uint64_t n, uint32_t* flags) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
       i < n; i += stride)
    flags[static_cast<uint64_t>(text[i])] = 1u;
}

uint32_t grid_for(uint64_t work, uint32_t block, uint32_t max_blocks) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return static_cast<uint32_t>(g);
}

}  // namespace

void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
  if (n == 0) return;
  ProfScope ps(K_FILL, s);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(n, 256, 4096)), dim3(256),
                     0, s, p, n, v);
  ECL_HIP(hipGetLastError());
}
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return;
  ProfScope ps(K_FILL, s);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 256, 4096)), dim3(256),
                     0, s, p, n, v);
  ECL_HIP(hipGetLastError());
}

void launch_histogram(const HistParams& p, hipStream_t s) {
  if (p.n == 0) return;
  ProfScope ps(K_HIST, s);
  // replicas per wave: keep 4 * copies * sigma u32 bins within 32 KiB
  uint32_t copies = 0;
  if (p.sigma <= 2048) {
    copies = 16;
    while (copies > 1 && 4ull * copies * p.sigma * 4 > 32768) copies >>= 1;
  }
  const bool lds = copies > 0;
  // each block handles at most 2^31 symbols so u32 bins cannot overflow
  uint64_t min_blocks = (p.n >> 31) + 1;
  uint32_t grid = grid_for(p.n / (16 / p.esize), 256 * 4, 8192);
  if (grid < min_blocks) grid = static_cast<uint32_t>(min_blocks);
  const size_t shmem = lds ? 4ull * copies * p.sigma * sizeof(uint32_t) : 0;
#define ECL_HIST_LAUNCH(TT, MK)                                                   \
  do {                                                                            \
    if (lds)                                                                      \
      hipLaunchKernelGGL((histogram_kernel<TT, MK, true>), dim3(grid), dim3(256), \
                         shmem, s, p, copies);                                    \
    else                                                                          \
      hipLaunchKernelGGL((histogram_kernel<TT, MK, false>), dim3(grid),           \
                         dim3(256), 0, s, p, 0u);                                 \
  } while (0)
#define ECL_HIST_TYPE(TT)                     \
  do {                                        \
    if (p.map_kind == 0) ECL_HIST_LAUNCH(TT, 0); \
    else if (p.map_kind == 1) ECL_HIST_LAUNCH(TT, 1); \
    else ECL_HIST_LAUNCH(TT, 2);              \
  } while (0)
  switch (p.esize) {
    case 1: ECL_HIST_TYPE(uint8_t); break;
    case 2: ECL_HIST_TYPE(uint16_t); break;
    case 4: ECL_HIST_TYPE(uint32_t); break;
    default: ECL_HIST_TYPE(uint64_t); break;
  }
#undef ECL_HIST_TYPE
#undef ECL_HIST_LAUNCH
  ECL_HIP(hipGetLastError());
}

uint64_t level_tiles(uint64_t n, int esize) {
  const uint64_t tile = kTileBytes / esize;
  return (n + tile - 1) / tile;
}

void launch_level(const LevelParams& p, hipStream_t s) {
  if (p.n == 0) return;
  const uint64_t tiles = level_tiles(p.n, p.esize);
  if (tiles > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "level too large");
  ProfScope ps(K_LEVEL, s);
  switch (p.esize) {
    case 1: hipLaunchKernelGGL(level_kernel<uint8_t>, dim3(tiles), dim3(256), 0, s, p); break;
    case 2: hipLaunchKernelGGL(level_kernel<uint16_t>, dim3(tiles), dim3(256), 0, s, p); break;
    case 4: hipLaunchKernelGGL(level_kernel<uint32_t>, dim3(tiles), dim3(256), 0, s, p); break;
    default: hipLaunchKernelGGL(level_kernel<uint64_t>, dim3(tiles), dim3(256), 0, s, p); break;
  }
  ECL_HIP(hipGetLastError());
  hipLaunchKernelGGL(l2_fixup_kernel, dim3(grid_for(p.num_l2, 256, 16384)), dim3(256), 0, s,
                     p.l2_global, p.l1, p.l2, p.num_l2);
  ECL_HIP(hipGetLastError());
}

void launch_rs_build(const RsBuildParams& p, hipStream_t s) {
  if (p.m == 0) return;
  if (p.num_l1 > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit array too large");
  ProfScope ps(K_RS_BUILD, s);
  hipLaunchKernelGGL(rs_build_kernel, dim3(static_cast<uint32_t>(p.num_l1)), dim3(256), 0, s, p);
  ECL_HIP(hipGetLastError());
}

void launch_gen_text(void* out, int esize, uint64_t n, uint64_t seed,
                     uint64_t sigma, const uint64_t* thr, hipStream_t s) {
  if (n == 0) return;
  ProfScope ps(K_GEN, s);
  const dim3 g(grid_for(n, 256, 65536)), b(256);
  switch (esize) {
    case 1: hipLaunchKernelGGL(gen_text_kernel<uint8_t>, g, b, 0, s, static_cast<uint8_t*>(out), n, seed, sigma, thr); break;
    case 2: hipLaunchKernelGGL(gen_text_kernel<uint16_t>, g, b, 0, s, static_cast<uint16_t*>(out), n, seed, sigma, thr); break;
    case 4: hipLaunchKernelGGL(gen_text_kernel<uint32_t>, g, b, 0, s, static_cast<uint32_t*>(out), n, seed, sigma, thr); break;
    default: hipLaunchKernelGGL(gen_text_kernel<uint64_t>, g, b, 0, s, static_cast<uint64_t*>(out), n, seed, sigma, thr); break;
  }
  ECL_HIP(hipGetLastError());
}

void launch_gen_bits(uint32_t* words, uint64_t nbits, uint64_t nwords_total,
                     uint64_t seed, uint64_t threshold, hipStream_t s) {
  if (nwords_total == 0) return;
  ProfScope ps(K_GEN, s);
  hipLaunchKernelGGL(gen_bits_kernel, dim3(grid_for(nwords_total, 256, 65536)),
                     dim3(256), 0, s, words, nbits, nwords_total, seed, threshold);
  ECL_HIP(hipGetLastError());
}

void launch_presence(const void* text, int esize, uint64_t n, uint32_t* flags,
                     hipStream_t s) {
  if (n == 0) return;
  const dim3 g(grid_for(n, 256, 16384)), b(256);
  switch (esize) {
    case 1: hipLaunchKernelGGL(presence_kernel<uint8_t>, g, b, 0, s, static_cast<const uint8_t*>(text), n, flags); break;
    case 2: hipLaunchKernelGGL(presence_kernel<uint16_t>, g, b, 0, s, static_cast<const uint16_t*>(text), n, flags); break;
    default: throw Error(ECL_ERROR_INTERNAL, "presence kernel only for 8/16-bit texts");
  }
  ECL_HIP(hipGetLastError());
}

}  // namespace lib
}  // namespace ecl
