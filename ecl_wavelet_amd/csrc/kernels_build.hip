// kernels_build.hip -- construction kernels for gfx950 (wave64).
//
//  * histogram_kernel: symbol counts with LDS-privatised u32 bins (the count
//    half of computeGlobalHistogramKernel, tree/wavelet_tree.cuh:170-227);
//    used when a multi-pass tree needs its level sizes before the passes.
//  * rs_build_kernel: one pass per bit array.  L2 counts, L1 prefix (look-
//    back over workgroups of 8 L1 blocks), total ones and both select-sample
//    streams in a single read of the bits.  Replaces calculateL2EntriesKernel
//    + cub InclusiveSum + addNumOnesKernel + calculateSelectSamplesKernel
//    (src/bitarray/rank_select.cu:412-591).
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
// Tiles of several independent chains may share one status array: the
// first tile of each chain publishes its inclusive prefix at once, so every
// later tile of the chain meets an inclusive prefix before leaving it.
template <int J>
__device__ uint64_t lookback(uint64_t* status, uint64_t tile, uint64_t agg,
                             uint64_t first = 0) {
  const uint32_t lane = lane_id();
  if (tile == first) {
    if (lane == 0) status_store(&status[tile], kFlagPre | agg);
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
// RankSelect directory: 8 L1 blocks (8 x 1024 u64 words) per workgroup of
// 512 threads; wave w owns L1 block 8 * tile + w, 16 words per lane.  One
// look-back per workgroup gives the ones before its first L1 block.  One
// launch covers any number of bit arrays: global tile t belongs to array
// tile_array[t] and is tile t - first_tile of that array's chain.
constexpr uint32_t kRsBlocksPerTile = 8;

__global__ __launch_bounds__(512) void rs_build_kernel(const RsArrayDesc* __restrict__ arrays,
                                                       const uint32_t* __restrict__ tile_array,
                                                       uint64_t* status, uint64_t* counter) {
  __shared__ uint64_t s_tile;
  __shared__ uint64_t s_excl;
  __shared__ uint32_t s_wave[kRsBlocksPerTile];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63, wave = tid >> 6;
  if (tid == 0)
    s_tile = atomicAdd(reinterpret_cast<unsigned long long*>(counter), 1ull);
  __syncthreads();
  const uint64_t gtile = s_tile;
  const RsArrayDesc p = arrays[tile_array[gtile]];
  const uint64_t tile = gtile - p.first_tile;
  const uint64_t kb = tile * kRsBlocksPerTile + wave;  // this wave's L1 block
  const uint64_t nwords = (p.m + 63) / 64;
  const uint64_t w0 = kb * 1024 + static_cast<uint64_t>(lane) * 16;
  uint64_t w[16];
  if (w0 + 15 < nwords) {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p.words64 + w0);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const ulonglong2 x = q[j];
      w[2 * j] = x.x;
      w[2 * j + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) w[j] = (w0 + j < nwords) ? p.words64[w0 + j] : 0ull;
  }
  const uint32_t tail = static_cast<uint32_t>(p.m & 63);
  uint32_t c = 0, c_first = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (tail && w0 + j == nwords - 1) w[j] &= (1ull << tail) - 1;
    const uint32_t pc = static_cast<uint32_t>(__popcll(w[j]));
    c += pc;
    if (j < 8) c_first += pc;
  }
  // ones in the L1 block before this lane; lane covers L2 blocks 2l, 2l+1
  const uint32_t incl = detail::wave_incl_scan_u32(c);
  const uint32_t excl = incl - c;
  {
    const uint64_t k2 = kb * detail::kL2PerL1 + 2ull * lane;
    if (k2 < p.num_l2) p.l2[k2] = static_cast<uint16_t>(excl);
    if (k2 + 1 < p.num_l2) p.l2[k2 + 1] = static_cast<uint16_t>(excl + c_first);
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t wbefore = 0, ttot = 0;
#pragma unroll
  for (uint32_t ww = 0; ww < kRsBlocksPerTile; ++ww) {
    const uint32_t x = s_wave[ww];
    wbefore += (ww < wave) ? x : 0u;
    ttot += x;
  }
  if (wave == 0) {
    const uint64_t g = lookback<kLookbackWords>(status, gtile, ttot, p.first_tile);
    if (lane == 0) s_excl = g;
  }
  __syncthreads();
  const uint64_t gb = s_excl + wbefore;  // ones before L1 block kb
  if (lane == 0 && kb < p.num_l1) p.l1[kb] = gb;
  if (lane == 0 && kb + 1 == p.num_l1) *p.total_ones = gb + s_wave[wave];

  // select samples: at most one multiple of 4096 ones (and of zeros) falls
  // inside a lane's 1024 bits
  const uint64_t b0 = w0 * 64;
  if (p.cap == 0 || b0 >= p.m) return;
  const uint64_t b1 = (b0 + 1024 < p.m) ? b0 + 1024 : p.m;
  const uint64_t ones_before = gb + excl;
  const uint64_t k1 = (ones_before + c) / detail::kSampleRate;
  if (k1 > ones_before / detail::kSampleRate) {
    uint32_t r = static_cast<uint32_t>(k1 * detail::kSampleRate - ones_before);
    uint64_t pos = b0;
#pragma unroll 1
    for (int j = 0; j < 16; ++j) {
      const uint32_t pc = static_cast<uint32_t>(__popcll(w[j]));
      if (r <= pc) {
        pos += detail::select_in_word(w[j], r);
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
    for (int j = 0; j < 16; ++j) {
      uint64_t z = ~w[j];
      const uint64_t wb = b0 + 64ull * j;
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

uint64_t rs_build_tiles(uint64_t num_l1) {
  return (num_l1 + kRsBlocksPerTile - 1) / kRsBlocksPerTile;
}

void launch_rs_build(const RsArrayDesc* arrays, const uint32_t* tile_array, uint64_t tiles,
                     uint64_t* status, uint64_t* counter, hipStream_t s) {
  if (tiles == 0) return;
  if (tiles > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit arrays too large");
  ProfScope ps(K_RS_BUILD, s);
  hipLaunchKernelGGL(rs_build_kernel, dim3(static_cast<uint32_t>(tiles)), dim3(512), 0, s,
                     arrays, tile_array, status, counter);
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
