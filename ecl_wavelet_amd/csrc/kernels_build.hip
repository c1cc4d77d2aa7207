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
// RankSelect directory in three steps, none of which waits on another
// workgroup:
//  1. rs_local_kernel: one wave per 65536-bit L1 block (16 u64 words per
//     lane): popcounts, in-block prefix = the block's L2 entries, block total.
//  2. rs_chunk_sum / rs_chunk_scan / rs_l1_write kernels: the exclusive scan
//     of every array's block totals = its L1 entries, and its total (chunk
//     sums, their scan per array, chunk scans from their bases).
//  3. rs_samples_kernel: one wave per L1 block, one lane per select sample
//     that falls inside the block: binary search over the block's L2 entries,
//     then one 64-byte L2 block is read and the bit is selected in-word.
// Global L1 block g belongs to array block_array[g] (written by
// rs_block_map_kernel), whose first block is arrays[...].first_tile.
constexpr uint32_t kRsWaves = 8;  // L1 blocks per workgroup

// one workgroup per array: its blocks' entries of block_array
__global__ __launch_bounds__(256) void rs_block_map_kernel(const RsArrayDesc* __restrict__ arrays,
                                                           uint32_t* __restrict__ block_array) {
  const RsArrayDesc& a = arrays[blockIdx.x];
  const uint64_t first = a.first_tile, cnt = a.cap_l1 ? a.cap_l1 : a.num_l1;
  for (uint64_t b = threadIdx.x; b < cnt; b += 256) block_array[first + b] = blockIdx.x;
}

__global__ __launch_bounds__(512) void rs_local_kernel(const RsArrayDesc* __restrict__ arrays,
                                                       const uint32_t* __restrict__ block_array,
                                                       uint64_t nblocks,
                                                       uint32_t* __restrict__ block_ones) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t gb = static_cast<uint64_t>(blockIdx.x) * kRsWaves + wave;
  if (gb >= nblocks) return;
  const RsArrayDesc p = arrays[block_array[gb]];
  const uint64_t kb = gb - p.first_tile;  // block within its array
  if (kb >= p.num_l1) return;             // reserved past the array's size
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
  // ones in the block before this lane; the lane covers L2 blocks 2l, 2l+1
  const uint32_t incl = detail::wave_incl_scan_u32(c);
  const uint32_t excl = incl - c;
  const uint64_t k2 = kb * detail::kL2PerL1 + 2ull * lane;
  if (k2 < p.num_l2) p.l2[k2] = static_cast<uint16_t>(excl);
  if (k2 + 1 < p.num_l2) p.l2[k2 + 1] = static_cast<uint16_t>(excl + c_first);
  if (lane == 63) block_ones[gb] = incl;
}

// L1 entries = exclusive prefix of each array's block totals, in three
// launches with every workgroup busy: the sum of every chunk of
// kRsScanChunk blocks (kept for the moment in the L1 entry of the chunk's
// first block), the exclusive scan of each array's chunk sums in place (and
// the array's total), then every chunk's own scan from its base.

// the array of global scan chunk j: the last whose first chunk is <= j
__device__ __forceinline__ uint32_t rs_chunk_array(const RsArrayDesc* __restrict__ arrays,
                                                   uint32_t narrays, uint64_t j) {
  uint32_t lo = 0, hi = narrays;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (arrays[mid].first_chunk <= j) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void rs_chunk_sum_kernel(const RsArrayDesc* __restrict__ arrays,
                                                           uint32_t narrays,
                                                           const uint32_t* __restrict__ block_ones) {
  __shared__ uint32_t s_w[4];
  const uint32_t tid = threadIdx.x;
  const RsArrayDesc& p = arrays[rs_chunk_array(arrays, narrays, blockIdx.x)];
  const uint64_t b0 = (blockIdx.x - p.first_chunk) * kRsScanChunk;
  if (b0 >= p.num_l1) return;  // reserved past the array's size
  const uint32_t n = static_cast<uint32_t>(
      p.num_l1 - b0 < kRsScanChunk ? p.num_l1 - b0 : kRsScanChunk);
  const uint32_t* cnt = block_ones + p.first_tile + b0;
  uint32_t sum = 0;
  for (uint32_t i = tid; i < n; i += 256) sum += cnt[i];
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if ((tid & 63) == 0) s_w[tid >> 6] = sum;
  __syncthreads();
  if (tid == 0) p.l1[b0] = static_cast<uint64_t>(s_w[0]) + s_w[1] + s_w[2] + s_w[3];
}

// one wave per array: chunk sums -> chunk bases (exclusive), and the total
__global__ __launch_bounds__(64) void rs_chunk_scan_kernel(const RsArrayDesc* __restrict__ arrays) {
  const RsArrayDesc p = arrays[blockIdx.x];
  const uint32_t lane = threadIdx.x;
  const uint64_t nch = rs_scan_chunks(p.num_l1);
  uint64_t run = 0;
  for (uint64_t base = 0; base < nch; base += 64) {
    const uint64_t c = base + lane;
    const uint64_t v = c < nch ? p.l1[c * kRsScanChunk] : 0ull;
    uint64_t incl = v;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint64_t y = __shfl_up(incl, off, 64);
      if (lane >= off) incl += y;
    }
    if (c < nch) p.l1[c * kRsScanChunk] = run + incl - v;
    run += __shfl(incl, 63, 64);
  }
  if (lane == 0) *p.total_ones = run;
}

// one workgroup per chunk: the chunk's L1 entries from its base, eight
// consecutive blocks per thread
__global__ __launch_bounds__(256) void rs_l1_write_kernel(const RsArrayDesc* __restrict__ arrays,
                                                          uint32_t narrays,
                                                          const uint32_t* __restrict__ block_ones) {
  __shared__ uint64_t s_base;
  __shared__ uint32_t s_w[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const RsArrayDesc& p = arrays[rs_chunk_array(arrays, narrays, blockIdx.x)];
  const uint64_t b0 = (blockIdx.x - p.first_chunk) * kRsScanChunk;
  if (b0 >= p.num_l1) return;
  const uint32_t n = static_cast<uint32_t>(
      p.num_l1 - b0 < kRsScanChunk ? p.num_l1 - b0 : kRsScanChunk);
  if (tid == 0) s_base = p.l1[b0];
  const uint32_t* cnt = block_ones + p.first_tile + b0;
  const uint32_t i0 = tid * 8;
  uint32_t x[8], mine = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    x[j] = i0 + j < n ? cnt[i0 + j] : 0u;
    mine += x[j];
  }
  const uint32_t incl = detail::wave_incl_scan_u32(mine);
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();  // the base is read before any entry of the chunk is written
  uint64_t run = s_base + incl - mine;
  for (uint32_t w = 0; w < wave; ++w) run += s_w[w];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (i0 + j < n) p.l1[b0 + i0 + j] = run;
    run += x[j];
  }
}

// position of the r-th (1-based) `value` bit inside the 8 words at q (one
// 64-byte L2 block, loaded at once; bits at or past valid_bits ignored)
template <int Value>
__device__ __forceinline__ uint32_t select_in_block(const uint64_t* q, uint32_t r,
                                                    uint64_t valid_bits) {
  const ulonglong2* v = reinterpret_cast<const ulonglong2*>(q);
  uint64_t x[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const ulonglong2 t = v[j];
    x[2 * j] = t.x;
    x[2 * j + 1] = t.y;
  }
  uint32_t pos = 0, found = 0, at = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    uint64_t y = Value ? x[j] : ~x[j];
    if (valid_bits < 64ull * (j + 1))
      y &= valid_bits > 64ull * j ? ((1ull << (valid_bits - 64ull * j)) - 1) : 0ull;
    const uint32_t pc = static_cast<uint32_t>(__popcll(y));
    if (!found && r <= pc) {
      found = 1;
      at = pos + detail::select_in_word(y, r);
    }
    if (!found) {
      r -= pc;
      pos += 64;
    }
  }
  return at;
}

__global__ __launch_bounds__(512) void rs_samples_kernel(const RsArrayDesc* __restrict__ arrays,
                                                         const uint32_t* __restrict__ block_array,
                                                         uint64_t nblocks) {
  __shared__ uint16_t s_l2[kRsWaves][detail::kL2PerL1];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t gb = static_cast<uint64_t>(blockIdx.x) * kRsWaves + wave;
  if (gb >= nblocks) return;
  const RsArrayDesc p = arrays[block_array[gb]];
  if (p.cap == 0) return;
  const uint64_t kb = gb - p.first_tile;
  if (kb >= p.num_l1) return;  // reserved past the array's size
  const uint64_t bits0 = kb * detail::kL1Bits;                     // first bit of the block
  const uint64_t bits = p.m - bits0 < detail::kL1Bits ? p.m - bits0 : detail::kL1Bits;
  const uint64_t nl2 = (bits + detail::kL2Bits - 1) / detail::kL2Bits;  // L2 blocks here
  // the block's L2 entries are loaded together with its L1 entries
  for (uint32_t x = lane; x < detail::kL2PerL1; x += 64)
    s_l2[wave][x] = x < nl2 ? p.l2[kb * detail::kL2PerL1 + x] : 0xFFFFu;
  const uint64_t o0 = p.l1[kb];
  const uint64_t o1 = kb + 1 < p.num_l1 ? p.l1[kb + 1] : *p.total_ones;
  const uint64_t z0 = bits0 - o0, z1 = bits0 + bits - o1;
  // samples in this block: ones ranks (o0, o1], zeros ranks (z0, z1]
  const uint64_t first1 = o0 / detail::kSampleRate + 1, last1 = o1 / detail::kSampleRate;
  const uint64_t first0 = z0 / detail::kSampleRate + 1, last0 = z1 / detail::kSampleRate;
  const uint64_t n1 = last1 >= first1 ? last1 - first1 + 1 : 0;
  const uint64_t n0 = last0 >= first0 ? last0 - first0 + 1 : 0;
  if (n1 + n0 == 0) return;
  // the wave's LDS writes complete before its reads (in-order LDS queue)
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  // lanes 0 .. n1-1: ones samples; n1 .. n1+n0-1: zeros samples (<= 32 in all)
  if (lane < n1 + n0) {
    const bool one = lane < n1;
    const uint64_t j = one ? first1 + lane : first0 + (lane - n1);  // sample number
    const uint32_t r = static_cast<uint32_t>(j * detail::kSampleRate - (one ? o0 : z0));
    // last L2 block whose prefix (of this value) is below r
    uint32_t lo = 0, hi = static_cast<uint32_t>(nl2);
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      const uint32_t before = one ? s_l2[wave][mid] : mid * 512u - s_l2[wave][mid];
      if (before < r) lo = mid; else hi = mid;
    }
    const uint32_t before = one ? s_l2[wave][lo] : lo * 512u - s_l2[wave][lo];
    const uint64_t wb = (bits0 >> 6) + lo * 8ull;  // first word of the L2 block
    const uint64_t valid = p.m - (wb << 6);
    const uint32_t pos = one ? select_in_block<1>(p.words64 + wb, r - before, valid)
                             : select_in_block<0>(p.words64 + wb, r - before, valid);
    const uint64_t at = (wb << 6) + pos;
    if (one) p.samples[j - 1] = at;
    else p.samples[p.cap - j] = at;
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

// Piece k (0..7) of query line j of an array: piece 0 holds the ones
// before the line (from the array's directory) and words 30j, 30j+1; piece
// k > 0 holds words 30j+4k-2 .. 30j+4k+1 (zeros past the array).
__device__ __forceinline__ uint4 query_line_piece(const LevelView& v, uint64_t j, uint32_t k) {
  const uint64_t nw = (v.size + 31) / 32;
  const uint64_t first = 30 * j + (k == 0 ? 0 : 4 * k - 2);  // even: 8-byte aligned pairs
  const uint32_t cnt = k == 0 ? 2 : 4;
  const uint64_t* w64 = reinterpret_cast<const uint64_t*>(v.words);
  uint32_t w[4] = {0, 0, 0, 0};
  if (first + cnt <= nw) {
    const uint64_t a = w64[first / 2];
    w[0] = static_cast<uint32_t>(a);
    w[1] = static_cast<uint32_t>(a >> 32);
    if (cnt == 4) {
      const uint64_t b = w64[first / 2 + 1];
      w[2] = static_cast<uint32_t>(b);
      w[3] = static_cast<uint32_t>(b >> 32);
    }
  } else {
    for (uint32_t x = 0; x < cnt; ++x) w[x] = first + x < nw ? v.words[first + x] : 0u;
  }
  if (k == 0) {
    const uint64_t s = j * detail::kLineBits;  // < size for every line of the array
    bool unused = false;
    const uint64_t head = s == 0 ? 0 : detail::rank1_inner(v, s, &unused);
    return make_uint4(static_cast<uint32_t>(head), static_cast<uint32_t>(head >> 32), w[0], w[1]);
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// One thread per 16-byte piece of the query lines of one array; consecutive
// threads read and write consecutive memory.
__global__ __launch_bounds__(256) void query_lines_kernel(LevelView v, uint4* __restrict__ lines,
                                                          uint64_t nlines) {
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  const uint64_t j = t >> 3;
  if (j >= nlines) return;
  lines[t] = query_line_piece(v, j, static_cast<uint32_t>(t & 7));
}

// segment i of the blob to its destination, one workgroup for all segments
__global__ __launch_bounds__(256) void upload_kernel(UploadBlob b) {
  for (uint32_t i = 0; i < b.nseg; ++i) {
    uint8_t* dst = reinterpret_cast<uint8_t*>(b.dst[i]);
    for (uint32_t x = b.off[i] + threadIdx.x; x < b.off[i + 1]; x += 256)
      dst[x - b.off[i]] = b.data[x];
  }
}

uint32_t grid_for(uint64_t work, uint32_t block, uint32_t max_blocks) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return static_cast<uint32_t>(g);
}

}  // namespace

// Select hint h of an array (h < nh1: the ones' hint h; else the zeros'
// hint h - nh1; layout at LevelView::sel_hints), written to out[0..3];
// returns whether its bracket fits (at most kHintBlocks blocks after b0).
__device__ __forceinline__ bool select_hint(const LevelView& v, uint64_t h, uint64_t nh1,
                                            uint4* __restrict__ out) {
  const bool one = h < nh1;
  const uint64_t j = one ? h : h - nh1;
  const uint64_t total = one ? v.ones : v.size - v.ones;
  const uint64_t nsamp = total / detail::kSampleRate;
  const uint64_t lo = j == 0 ? 0 : (one ? v.samples[j - 1] : v.samples[v.sample_cap - j]);
  const uint64_t hi = j < nsamp ? (one ? v.samples[j] : v.samples[v.sample_cap - 1 - j])
                                : v.size - 1;
  const uint64_t b0 = lo >> 9;
  const uint64_t span = (hi >> 9) - b0;
  auto before = [&](uint64_t blk) -> uint64_t {  // Values before block blk of the array
    const uint64_t ones = v.l1[blk >> 7] + v.l2[blk];
    return one ? ones : (blk << 9) - ones;
  };
  const uint64_t c0 = before(b0);
  uint32_t e[24];
  e[0] = span > detail::kHintBlocks ? 0xFFFFu : static_cast<uint32_t>(span);
  for (uint32_t u = 1; u < 24; ++u)
    e[u] = u <= span ? static_cast<uint32_t>(before(b0 + u) - c0) : 0xFFFFu;
  out[0] = make_uint4(static_cast<uint32_t>(lo), static_cast<uint32_t>(lo >> 32),
                      static_cast<uint32_t>(c0), static_cast<uint32_t>(c0 >> 32));
  for (int k = 0; k < 3; ++k)
    out[1 + k] = make_uint4(e[8 * k] | (e[8 * k + 1] << 16), e[8 * k + 2] | (e[8 * k + 3] << 16),
                            e[8 * k + 4] | (e[8 * k + 5] << 16), e[8 * k + 6] | (e[8 * k + 7] << 16));
  return span <= detail::kHintBlocks;
}

// One thread per select hint of one array: hints [0, nh1) for the ones,
// [nh1, nh1 + nh0) for the zeros.  block_fit[2 b + 1] / block_fit[2 b]
// count the hints of workgroup b for the ones / zeros whose bracket fits.
__global__ __launch_bounds__(256) void select_hints_kernel(LevelView v, uint4* __restrict__ hints,
                                                           uint64_t nh1, uint64_t nh0,
                                                           uint32_t* __restrict__ block_fit) {
  __shared__ uint32_t s_fit[2];
  if (threadIdx.x < 2) s_fit[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  bool fits1 = false, fits0 = false;
  if (t < nh1 + nh0) {
    const bool fits = select_hint(v, t, nh1, hints + t * 4);
    fits1 = t < nh1 && fits;
    fits0 = t >= nh1 && fits;
  }
  const uint64_t m1 = __ballot(fits1), m0 = __ballot(fits0);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&s_fit[1], static_cast<uint32_t>(__popcll(m1)));
    atomicAdd(&s_fit[0], static_cast<uint32_t>(__popcll(m0)));
  }
  __syncthreads();
  if (threadIdx.x < 2) block_fit[2 * blockIdx.x + threadIdx.x] = s_fit[threadIdx.x];
}

// One workgroup: sums the per-workgroup fit counts of one array's hints and
// sets the array's hint_values in its device view (a Value reads the hints
// where at least half of its brackets fit one).
__global__ __launch_bounds__(256) void hint_values_kernel(const uint32_t* __restrict__ block_fit,
                                                          uint64_t nblocks, uint64_t nh1,
                                                          uint64_t nh0, LevelView* d_view) {
  __shared__ uint64_t s_sum[2][4];
  uint64_t f1 = 0, f0 = 0;
  for (uint64_t b = threadIdx.x; b < nblocks; b += 256) {
    f1 += block_fit[2 * b + 1];
    f0 += block_fit[2 * b];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    f1 += __shfl_xor(static_cast<unsigned long long>(f1), off, 64);
    f0 += __shfl_xor(static_cast<unsigned long long>(f0), off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s_sum[1][threadIdx.x >> 6] = f1;
    s_sum[0][threadIdx.x >> 6] = f0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t t1 = s_sum[1][0] + s_sum[1][1] + s_sum[1][2] + s_sum[1][3];
    const uint64_t t0 = s_sum[0][0] + s_sum[0][1] + s_sum[0][2] + s_sum[0][3];
    d_view->hint_values = (2 * t1 >= nh1 ? 2u : 0u) | (2 * t0 >= nh0 ? 1u : 0u);
  }
}

uint64_t select_hint_blocks(uint64_t nh1, uint64_t nh0) { return (nh1 + nh0 + 255) / 256; }

void launch_select_hints(const LevelView& v, uint4* hints, uint64_t nh1, uint64_t nh0,
                         uint32_t* block_fit, LevelView* d_view, hipStream_t s) {
  if (nh1 + nh0 == 0) return;
  const uint64_t blocks = select_hint_blocks(nh1, nh0);
  if (blocks > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit array too large");
  LevelView plain = v;
  plain.lines = nullptr;
  plain.sel_hints = nullptr;
  hipLaunchKernelGGL(select_hints_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s,
                     plain, hints, nh1, nh0, block_fit);
  ECL_HIP(hipGetLastError());
  hipLaunchKernelGGL(hint_values_kernel, dim3(1), dim3(256), 0, s, block_fit, blocks, nh1, nh0,
                     d_view);
  ECL_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Query lines and select hints of many arrays (the levels of a forest) in
// three launches.  Array a owns lines [line_start[a], line_start[a + 1]) and
// hints [hint_start[a], hint_start[a + 1]) of the shared buffers; a thread
// finds its array by binary search over the starts (empty arrays share their
// start with the next array, and the search takes the last array whose start
// is not above the item, which owns it).
__device__ __forceinline__ uint32_t owner_job(const uint64_t* __restrict__ start, uint32_t njobs,
                                              uint64_t x) {
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (start[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void multi_query_lines_kernel(
    const LinesJob* __restrict__ jobs, const uint64_t* __restrict__ line_start, uint32_t njobs,
    uint4* __restrict__ lines, uint64_t total_lines) {
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  const uint64_t j = t >> 3;
  if (j >= total_lines) return;
  const uint32_t a = owner_job(line_start, njobs, j);
  lines[t] = query_line_piece(jobs[a].v, j - line_start[a], static_cast<uint32_t>(t & 7));
}

// fit[2 a + 1] / fit[2 a]: hints of array a for the ones / zeros that fit
__global__ __launch_bounds__(256) void multi_select_hints_kernel(
    const LinesJob* __restrict__ jobs, const uint64_t* __restrict__ hint_start, uint32_t njobs,
    uint4* __restrict__ hints, uint64_t total_hints, uint32_t* __restrict__ fit) {
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= total_hints) return;
  const uint32_t a = owner_job(hint_start, njobs, t);
  const uint64_t h = t - hint_start[a];
  const uint64_t nh1 = jobs[a].nh1;
  if (select_hint(jobs[a].v, h, nh1, hints + t * 4))
    atomicAdd(&fit[2 * static_cast<uint64_t>(a) + (h < nh1 ? 1 : 0)], 1u);
}

// one thread per array: its device view gets the lines, the hints and the
// Values that read the hints (at least half of their brackets fit one)
__global__ __launch_bounds__(256) void multi_lines_views_kernel(
    const LinesJob* __restrict__ jobs, const uint64_t* __restrict__ line_start,
    const uint64_t* __restrict__ hint_start, uint32_t njobs, const uint4* lines,
    const uint4* hints, const uint32_t* __restrict__ fit) {
  const uint64_t a = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (a >= njobs) return;
  const LinesJob& J = jobs[a];
  LevelView* d = J.d_view;
  d->lines = J.nl ? lines + line_start[a] * 8 : nullptr;
  d->sel_hints = J.nh1 + J.nh0 ? hints + hint_start[a] * 4 : nullptr;
  d->hint_values = (2ull * fit[2 * a + 1] >= J.nh1 ? 2u : 0u) | (2ull * fit[2 * a] >= J.nh0 ? 1u : 0u);
}

void launch_multi_lines(const LinesJob* jobs, const uint64_t* line_start,
                        const uint64_t* hint_start, uint32_t njobs, uint4* lines,
                        uint64_t total_lines, uint4* hints, uint64_t total_hints, uint32_t* fit,
                        hipStream_t s) {
  if (njobs == 0) return;
  const uint64_t lb = (total_lines * 8 + 255) / 256, hb = (total_hints + 255) / 256;
  if (lb > 0x7FFFFFFFull || hb > 0x7FFFFFFFull)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit arrays too large");
  if (lb)
    hipLaunchKernelGGL(multi_query_lines_kernel, dim3(static_cast<uint32_t>(lb)), dim3(256), 0, s,
                       jobs, line_start, njobs, lines, total_lines);
  if (hb)
    hipLaunchKernelGGL(multi_select_hints_kernel, dim3(static_cast<uint32_t>(hb)), dim3(256), 0, s,
                       jobs, hint_start, njobs, hints, total_hints, fit);
  hipLaunchKernelGGL(multi_lines_views_kernel, dim3((njobs + 255) / 256), dim3(256), 0, s, jobs,
                     line_start, hint_start, njobs, lines, hints, fit);
  ECL_HIP(hipGetLastError());
}

// One thread per select line of one array (layout at LevelView::sel_lines):
// lines [0, n1) for the ones at rate r1, [n1, n1 + n0) for the zeros at r0.
// The line's start comes from a select over the array's directory (and its
// hints when they exist); its 960 bits are copied from the 64-bit word that
// holds it (zeros past the array).
// floor(x / r) for r >= 1, with magic = floor((2^64 - 1) / r) + 1 for r > 1
// (exact for x < 2^64 / r; the counts here are below 2^44)
__device__ __forceinline__ uint64_t div_rate(uint64_t x, uint32_t r, uint64_t magic) {
  return r == 1 ? x : __umul64hi(x, magic);
}

// Select lines, one workgroup per 256 consecutive 512-bit blocks and Value
// (the ones' workgroups first).  Line j of Value V starts at the (j r)-th V
// (line 0 at bit 0), so the lines that start in block b are those with
// cb < j r <= ce, cb and ce the V's before blocks b and b + 1 (from L1 +
// L2); thread t takes block b0 + t.  The lines of one workgroup are
// consecutive, [first, first + T), and their bits lie in its blocks and the
// next two.  The workgroup stages those 258 blocks in LDS with coalesced
// loads; each thread places its lines with one select inside its block and
// records their starts; then the workgroup writes the lines out
// contiguously, one 16-byte piece per lane.  (A line per thread written
// with its own 16-byte stores touched 64 lines per store instruction: 1.08
// ms per level of 6e9 bits; assembling each line in LDS by one thread put
// the lanes' stores 128 bytes apart, on the same LDS banks: 1.24 ms.)
constexpr uint32_t kSlBlocks = kSelLineChunk;       // blocks (threads) per workgroup
constexpr uint32_t kSlSrcWords = (kSlBlocks + 2) * 16;
constexpr uint32_t kSlLines = 1024;                 // line starts recorded per round

template <int Value>
__device__ __forceinline__ void select_lines_wg(const LevelView& v, uint4* __restrict__ out,
                                                uint64_t nl2, uint64_t chunk, uint64_t base,
                                                uint32_t r, uint64_t magic,
                                                uint32_t* s_src, uint64_t* s_pos) {
  const uint32_t tid = threadIdx.x;
  const uint64_t b0 = chunk * kSlBlocks;
  const uint64_t total = Value ? v.ones : v.size - v.ones;
  auto before = [&](uint64_t blk) -> uint64_t {  // Values before block blk
    if (blk >= nl2) return total;
    const uint64_t ones = detail::gld(v.l1, blk >> 7) + detail::gld(v.l2, blk);
    return Value ? ones : (blk << 9) - ones;
  };
  // the blocks' words (zeros past the array's words, as a line's tail)
  const uint64_t nw = (v.size + 31) / 32;
  for (uint32_t i = tid; i < kSlSrcWords / 4; i += kSlBlocks) {
    const uint64_t wi = b0 * 16 + 4ull * i;
    uint4 q;
    if (wi + 4 <= nw) {
      q = detail::gld(reinterpret_cast<const uint4*>(v.words + wi), 0);
    } else {
      q.x = wi < nw ? v.words[wi] : 0u;
      q.y = wi + 1 < nw ? v.words[wi + 1] : 0u;
      q.z = wi + 2 < nw ? v.words[wi + 2] : 0u;
      q.w = 0u;  // wi + 3 >= nw here
    }
    reinterpret_cast<uint4*>(s_src)[i] = q;
  }
  // this thread's lines [jlo, jend] and the workgroup's [first, first + T)
  const uint64_t b = b0 + tid;
  const uint64_t bend = b0 + kSlBlocks < nl2 ? b0 + kSlBlocks : nl2;
  const uint64_t first = b0 == 0 ? 0 : div_rate(before(b0), r, magic) + 1;
  const uint64_t T = div_rate(before(bend), r, magic) + 1 - first;
  uint64_t jlo = 1, jend = 0, cb = 0;  // empty
  if (b < nl2) {
    cb = before(b);
    jlo = b == 0 ? 0 : div_rate(cb, r, magic) + 1;
    jend = div_rate(before(b + 1), r, magic);
  }
  __syncthreads();
  const uint32_t blk_word = tid * 16;  // this block's words in s_src
  for (uint64_t r0 = 0; r0 < T; r0 += kSlLines) {
    for (uint64_t j = jlo; j <= jend; ++j) {
      const uint64_t slot = j - first;
      if (slot < r0) continue;
      if (slot >= r0 + kSlLines) break;
      uint64_t pos = 0;
      if (j != 0) {
        uint32_t rank = static_cast<uint32_t>(j * r - cb);  // 1-based, <= 512
        uint32_t bit = 0;
#pragma unroll 1
        for (uint32_t k = 0; k < 8; ++k) {
          uint64_t x = static_cast<uint64_t>(s_src[blk_word + 2 * k]) |
                       (static_cast<uint64_t>(s_src[blk_word + 2 * k + 1]) << 32);
          if (!Value) x = ~x;
          const uint32_t c = static_cast<uint32_t>(__popcll(x));
          if (rank <= c) {
            bit = 64 * k + detail::select_in_word(x, rank);
            break;
          }
          rank -= c;
        }
        pos = (b << 9) + bit;
      }
      s_pos[slot - r0] = pos;
    }
    __syncthreads();
    // piece k of a line: k = 0 the position and words 0, 1; k > 0 words
    // 4k - 2 .. 4k + 1 of the 30 from the 64-bit word holding the position
    const uint64_t nl = T - r0 < kSlLines ? T - r0 : kSlLines;
    uint4* dst = out + (base + first + r0) * 8;
    for (uint32_t i = tid; i < nl * 8; i += kSlBlocks) {
      const uint64_t pos = s_pos[i >> 3];
      const uint32_t k = i & 7;
      const uint32_t lw = static_cast<uint32_t>((pos >> 6) * 2 - b0 * 16);
      uint4 q;
      if (k == 0) {
        q = make_uint4(static_cast<uint32_t>(pos), static_cast<uint32_t>(pos >> 32), s_src[lw],
                       s_src[lw + 1]);
      } else {
        const uint32_t w = lw + 4 * k - 2;
        q = make_uint4(s_src[w], s_src[w + 1], s_src[w + 2], s_src[w + 3]);
      }
      dst[i] = q;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kSlBlocks) void select_lines_kernel(LevelView v, uint4* __restrict__ out,
                                                                 uint64_t nl2, uint64_t nchunks,
                                                                 uint64_t n1, uint32_t r1, uint32_t r0,
                                                                 uint64_t magic1, uint64_t magic0) {
  __shared__ __align__(16) uint32_t s_src[kSlSrcWords];
  __shared__ uint64_t s_pos[kSlLines];
  const bool one = blockIdx.x < nchunks;
  const uint64_t chunk = one ? blockIdx.x : blockIdx.x - nchunks;
  if (one) {
    if (r1 != 0) select_lines_wg<1>(v, out, nl2, chunk, 0, r1, magic1, s_src, s_pos);
  } else {
    if (r0 != 0) select_lines_wg<0>(v, out, nl2, chunk, n1, r0, magic0, s_src, s_pos);
  }
}

void launch_select_lines(const LevelView& v, uint4* out, uint64_t n1, uint64_t n0, uint32_t r1,
                         uint32_t r0, hipStream_t s) {
  if (n1 + n0 == 0 || v.size == 0) return;
  const uint64_t nl2 = (v.size + detail::kL2Bits - 1) / detail::kL2Bits;
  const uint64_t nchunks = (nl2 + kSlBlocks - 1) / kSlBlocks;
  if (2 * nchunks > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit array too large");
  hipLaunchKernelGGL(select_lines_kernel, dim3(static_cast<uint32_t>(2 * nchunks)), dim3(kSlBlocks),
                     0, s, v, out, nl2, nchunks, n1, r1, r0, select_line_magic(r1),
                     select_line_magic(r0));
  ECL_HIP(hipGetLastError());
}

// The select lines of many arrays: a workgroup finds its array by a search
// over the arrays' first workgroups (an array without lines has none, and
// shares its start with the next one) and builds one chunk of one Value, as
// select_lines_kernel does.
__global__ __launch_bounds__(kSlBlocks) void multi_select_lines_kernel(
    const SelLinesJob* __restrict__ jobs, const uint64_t* __restrict__ wg_start, uint32_t njobs,
    uint4* __restrict__ out) {
  __shared__ __align__(16) uint32_t s_src[kSlSrcWords];
  __shared__ uint64_t s_pos[kSlLines];
  const uint32_t a = owner_job(wg_start, njobs, blockIdx.x);
  const SelLinesJob& J = jobs[a];
  const LevelView v = J.v;
  const uint64_t w = blockIdx.x - wg_start[a];
  const bool one = w < J.nchunks;
  uint4* dst = out + J.line_off * 8;
  if (one) {
    if (J.r1 != 0)
      select_lines_wg<1>(v, dst, J.nl2, w, 0, J.r1, select_line_magic(J.r1), s_src, s_pos);
  } else {
    if (J.r0 != 0)
      select_lines_wg<0>(v, dst, J.nl2, w - J.nchunks, J.n1, J.r0, select_line_magic(J.r0), s_src,
                         s_pos);
  }
}

void launch_multi_select_lines(const SelLinesJob* jobs, const uint64_t* wg_start, uint32_t njobs,
                               uint64_t total_wgs, uint4* out, hipStream_t s) {
  if (njobs == 0 || total_wgs == 0) return;
  if (total_wgs > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit arrays too large");
  hipLaunchKernelGGL(multi_select_lines_kernel, dim3(static_cast<uint32_t>(total_wgs)),
                     dim3(kSlBlocks), 0, s, jobs, wg_start, njobs, out);
  ECL_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Quad lines of one quad level q, from binary levels a = 2q and b = 2q + 1.
// In node N of level 2q (positions [S_N, S_N+1)) the codes with bit 0 go,
// in order, to positions [S_N, S_N + Z_N) of level 2q + 1 and those with bit
// 1 to [S_N + Z_N, S_N+1), Z_N the node's zeros.  So the code at position p
// with o ones and z zeros of level 2q before it in its node has its level
// 2q + 1 bit at S_N + z (bit 0) or S_N + Z_N + o (bit 1): a word of 32
// positions inside one node takes its second bits from two contiguous runs
// of level 2q + 1.  Line headers: the d's before position P are the d's
// before its node (CD) plus, for d = 1 and 3, the ones of level 2q + 1 in
// the two runs up to P (two ranks), and for d = 0 and 2 the rest of the
// node's zeros and ones before P.
// A workgroup builds kQuadLinesPerWg lines in LDS and writes them out
// contiguously.
constexpr uint32_t kQuadLinesPerWg = 64;
constexpr uint32_t kQuadWordsPerLine = kQuadDigits / 32;  // 14 words of 32 positions

// last node N with S[N] <= pos (pos < n)
__device__ __forceinline__ uint64_t quad_node(const uint64_t* S, uint64_t nodes, uint64_t pos) {
  uint64_t lo = 0, hi = nodes;  // S[lo] <= pos < S[hi]
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (detail::gld(S, mid) <= pos) lo = mid; else hi = mid;
  }
  return lo;
}
// bits [x, x + 32) of a bit array of nw words (zeros past them)
__device__ __forceinline__ uint32_t bits32(const uint32_t* w, uint64_t nw, uint64_t x) {
  const uint64_t k = x >> 5;
  const uint32_t sh = static_cast<uint32_t>(x & 31);
  const uint32_t lo = k < nw ? detail::gld(w, k) : 0u;
  if (sh == 0) return lo;
  const uint32_t hi = k + 1 < nw ? detail::gld(w, k + 1) : 0u;
  return (lo >> sh) | (hi << (32 - sh));
}

// node holding pos, searched forward from a node N0 with S[N0] <= pos
__device__ __forceinline__ uint64_t quad_node_from(const uint64_t* S, uint64_t nodes, uint64_t N0,
                                                   uint64_t pos) {
  if (pos < detail::gld(S, N0 + 1)) return N0;
  uint64_t lo = N0 + 1, hi = nodes;  // S[lo] <= pos < S[hi]
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (detail::gld(S, mid) <= pos) lo = mid; else hi = mid;
  }
  return lo;
}

// Thread t takes the 4 words (128 positions) from word 4t of the
// workgroup's range; a block scan of their ones gives every word's rank from
// one directory rank at the range's start, and the node search runs once per
// workgroup (a thread searches only past a node start).
__device__ __forceinline__ void quad_lines_wg(const QuadBuildParams& qp, uint64_t blk) {
  constexpr uint32_t kWords = kQuadLinesPerWg * kQuadWordsPerLine;  // 896
  __shared__ __align__(16) uint32_t s_lines[kQuadLinesPerWg * 32];
  __shared__ uint32_t s_wrank[kWords];  // ones of level 2q before each word, from the range start
  __shared__ uint32_t s_wave[4];
  // from quad_wg_kernel: ones of level 2q before the range, the node holding
  // its start (node0), node0's start and end, the ones of level 2q before
  // node0 and the zeros in it
  __shared__ uint64_t s_info[kQuadWgInfo];
  __shared__ __align__(16) uint8_t s_lut[4096];
  reinterpret_cast<uint4*>(s_lut)[threadIdx.x] =
      detail::gld(reinterpret_cast<const uint4*>(qp.lut), threadIdx.x);
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint64_t line0 = blk * kQuadLinesPerWg;
  const uint64_t nl = qp.nlines - line0 < kQuadLinesPerWg ? qp.nlines - line0 : kQuadLinesPerWg;
  const uint32_t nwords = static_cast<uint32_t>(nl) * kQuadWordsPerLine;
  const uint64_t n = qp.n, nw = (n + 31) / 32;
  const uint64_t k0 = line0 * kQuadWordsPerLine;  // first word (a multiple of 4)
  const uint32_t* A = qp.a.words;
  const uint32_t* B = qp.b.words;
  auto cd = [&](uint64_t N, uint32_t d) { return detail::gld(qp.CD, 4 * N + d); };
  // this thread's words (zero past the array) and their ones
  uint32_t a4[4] = {0u, 0u, 0u, 0u};
  const uint64_t kt = k0 + 4ull * tid;
  if (4 * tid < nwords && kt < nw) {
    if (kt + 4 <= nw) {
      const uint4 q = detail::gld(reinterpret_cast<const uint4*>(A + kt), 0);
      a4[0] = q.x; a4[1] = q.y; a4[2] = q.z; a4[3] = q.w;
    } else {
      for (uint32_t i = 0; i < 4; ++i) a4[i] = kt + i < nw ? detail::gld(A, kt + i) : 0u;
    }
  }
  const uint32_t pc = __popc(a4[0]) + __popc(a4[1]) + __popc(a4[2]) + __popc(a4[3]);
  uint32_t incl = pc;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) s_wave[wave] = incl;
  if (tid < kQuadWgInfo) s_info[tid] = detail::gld(qp.wg_info, blk * kQuadWgInfo + tid);
  __syncthreads();
  uint32_t run = incl - pc;
  for (uint32_t w = 0; w < wave; ++w) run += s_wave[w];
  for (uint32_t i = 0; i < 4; ++i) {
    if (4 * tid + i < nwords) s_wrank[4 * tid + i] = run;
    run += __popc(a4[i]);
  }
  const uint64_t rank0 = s_info[0], node0 = s_info[1];
  const uint64_t SN0 = s_info[2], SN10 = s_info[3], OB0 = s_info[4], ZN0 = s_info[5];
  __syncthreads();
  // digits: 32 positions per word.  Words inside node0 (nearly all) take its
  // parameters from LDS and fetch their two runs of level 2q + 1 first, all
  // four words' runs at once; the others go the general way.
  uint32_t zb[4], ob[4];
  bool fast[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    const uint32_t w = 4 * tid + i;
    const uint64_t P0 = 32 * (k0 + w);
    fast[i] = false;
    zb[i] = ob[i] = 0;
    if (w < nwords && P0 < n) {
      const uint32_t valid = n - P0 < 32 ? static_cast<uint32_t>(n - P0) : 32u;
      if (P0 + valid <= SN10) {
        const uint64_t o0 = rank0 + s_wrank[w] - OB0;
        const uint64_t z0 = (P0 - SN0) - o0;
        zb[i] = bits32(B, nw, SN0 + z0);
        ob[i] = bits32(B, nw, SN0 + ZN0 + o0);
        fast[i] = true;
      }
    }
  }
  for (uint32_t i = 0; i < 4; ++i) {
    const uint32_t w = 4 * tid + i;
    if (w >= nwords) break;
    const uint64_t P0 = 32 * (k0 + w);
    uint64_t D = 0;
    if (P0 < n) {
      const uint32_t valid = n - P0 < 32 ? static_cast<uint32_t>(n - P0) : 32u;
      const uint32_t a = a4[i];  // bits past n are zero
      const uint64_t rA = rank0 + s_wrank[w];  // ones of level 2q before P0
      uint32_t zbi = zb[i], obi = ob[i];
      bool one_node = fast[i];
      uint64_t N = node0, SN = SN0, SN1 = SN10;
      if (!one_node) {
        N = quad_node_from(qp.S, qp.nodes, node0, P0);
        SN = detail::gld(qp.S, N);
        SN1 = detail::gld(qp.S, N + 1);
        if (P0 + valid <= SN1) {
          const uint64_t o0 = rA - (cd(N, 2) + cd(N, 3));
          const uint64_t z0 = (P0 - SN) - o0;
          const uint64_t ZN = (cd(N + 1, 0) - cd(N, 0)) + (cd(N + 1, 1) - cd(N, 1));
          zbi = bits32(B, nw, SN + z0);
          obi = bits32(B, nw, SN + ZN + o0);
          one_node = true;
        }
      }
      if (one_node) {
        // one node: the second bits come from the two runs, 4 positions per
        // table lookup (at most 28 of a run's 32 bits precede a nibble)
        uint32_t iz = 0, io = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
          const uint32_t a4 = (a >> (4 * k)) & 15u;
          const uint32_t z4 = (zbi >> iz) & 15u, o4 = (obi >> io) & 15u;
          D |= static_cast<uint64_t>(s_lut[a4 | (z4 << 4) | (o4 << 8)]) << (8 * k);
          const uint32_t c = __popc(a4);
          io += c;
          iz += 4u - c;
        }
        // positions past n (the last word) hold no digits
        if (valid < 32) D &= (1ull << (2 * valid)) - 1;
      } else {
        // the word crosses node starts: position by position
        uint32_t before = 0;  // ones of the word before position j
        for (uint32_t j = 0; j < valid; ++j) {
          const uint64_t pos = P0 + j;
          while (pos >= SN1) {
            ++N;
            SN = SN1;
            SN1 = detail::gld(qp.S, N + 1);
          }
          const uint64_t o = rA + before - (cd(N, 2) + cd(N, 3));
          const uint64_t z = (pos - SN) - o;
          const uint64_t ZN = (cd(N + 1, 0) - cd(N, 0)) + (cd(N + 1, 1) - cd(N, 1));
          const uint32_t aj = (a >> j) & 1u;
          const uint64_t pb = SN + (aj ? ZN + o : z);
          const uint32_t bj = (detail::gld(B, pb >> 5) >> (pb & 31)) & 1u;
          D |= static_cast<uint64_t>((aj << 1) | bj) << (2 * j);
          before += aj;
        }
      }
    }
    const uint32_t li = w / kQuadWordsPerLine, wi = w - li * kQuadWordsPerLine;
    uint32_t* dst = s_lines + li * 32 + 4 + 2 * wi;
    dst[0] = static_cast<uint32_t>(D);
    dst[1] = static_cast<uint32_t>(D >> 32);
  }
  // headers: the counts of digits 0, 1, 2 before the line's first position
  if (tid < nl) {
    const uint64_t P = (line0 + tid) * kQuadDigits;
    uint64_t c0, c1, c2;
    if (P >= n) {
      c0 = cd(qp.nodes, 0);
      c1 = cd(qp.nodes, 1);
      c2 = cd(qp.nodes, 2);
    } else {
      const uint64_t N = quad_node_from(qp.S, qp.nodes, node0, P);
      const uint64_t SN = detail::gld(qp.S, N);
      const uint64_t o = rank0 + s_wrank[tid * kQuadWordsPerLine] - (cd(N, 2) + cd(N, 3));
      const uint64_t z = (P - SN) - o;
      const uint64_t cnt1 = cd(N + 1, 1) - cd(N, 1);
      const uint64_t ZN = (cd(N + 1, 0) - cd(N, 0)) + cnt1;
      const uint64_t rB = cd(N, 1) + cd(N, 3);  // ones of level 2q + 1 before S_N
      const uint64_t d1 = detail::level_rank<1>(qp.b, SN + z) - rB;
      const uint64_t d3 = detail::level_rank<1>(qp.b, SN + ZN + o) - (rB + cnt1);
      c0 = cd(N, 0) + (z - d1);
      c1 = cd(N, 1) + d1;
      c2 = cd(N, 2) + (o - d3);
    }
    uint32_t* h = s_lines + tid * 32;
    const uint64_t h0 = c0 | (c1 << 40);
    const uint64_t h1 = (c1 >> 24) | (c2 << 16);
    h[0] = static_cast<uint32_t>(h0);
    h[1] = static_cast<uint32_t>(h0 >> 32);
    h[2] = static_cast<uint32_t>(h1);
    h[3] = static_cast<uint32_t>(h1 >> 32);
  }
  __syncthreads();
  uint4* out = qp.out + line0 * 8;
  for (uint32_t i = tid; i < nl * 8; i += blockDim.x) out[i] = reinterpret_cast<const uint4*>(s_lines)[i];
}
__global__ __launch_bounds__(256) void quad_lines_kernel(QuadBuildParams qp) {
  quad_lines_wg(qp, blockIdx.x);
}

// Per build workgroup g (positions from P = 448 * 64 g): the ones of level
// 2q before P, the node N holding P (the last position when P >= n), N's
// start and end, the ones of level 2q before N and the zeros in N -- the
// chain of dependent loads (directory rank, node search, tables) that would
// otherwise hold up every build workgroup, here one thread each, all at once.
__device__ __forceinline__ void quad_wg_one(const QuadBuildParams& qp, uint64_t g) {
  const uint64_t P = g * kQuadLinesPerWg * kQuadDigits;
  const uint64_t N = quad_node(qp.S, qp.nodes, P < qp.n ? P : qp.n - 1);
  auto cd = [&](uint64_t M, uint32_t d) { return detail::gld(qp.CD, 4 * M + d); };
  uint64_t* o = qp.wg_info + g * kQuadWgInfo;
  o[0] = detail::level_rank<1>(qp.a, P);
  o[1] = N;
  o[2] = detail::gld(qp.S, N);
  o[3] = detail::gld(qp.S, N + 1);
  o[4] = cd(N, 2) + cd(N, 3);
  o[5] = (cd(N + 1, 0) - cd(N, 0)) + (cd(N + 1, 1) - cd(N, 1));
}
__global__ __launch_bounds__(256) void quad_wg_kernel(QuadBuildParams qp, uint64_t nwg) {
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g < nwg) quad_wg_one(qp, g);
}

uint64_t quad_build_workgroups(uint64_t nlines) {
  return (nlines + kQuadLinesPerWg - 1) / kQuadLinesPerWg;
}

void launch_quad_lines(const QuadBuildParams& qp, hipStream_t s) {
  if (qp.nlines == 0) return;
  const uint64_t blocks = quad_build_workgroups(qp.nlines);
  if (blocks > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit array too large");
  hipLaunchKernelGGL(quad_wg_kernel, dim3(static_cast<uint32_t>((blocks + 255) / 256)), dim3(256),
                     0, s, qp, blocks);
  hipLaunchKernelGGL(quad_lines_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, qp);
  ECL_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Quad select lines of one digit d of one quad level, one workgroup per 256
// quad lines: line j starts at the (j r)-th d (line 0 at position 0), so
// the lines that start in quad line L are those with cb < j r <= ce, cb and
// ce the d's before quad lines L and L + 1 (their headers).  The workgroup
// stages its quad lines and the next one in LDS; each thread places the
// lines that start in its quad line with one select inside it; then the
// workgroup writes them out contiguously, one 16-byte piece per lane.  A
// line's 448 digits, from the 32-digit word holding its start, lie in that
// quad line and the next.
constexpr uint32_t kQsQuadLines = 256;
// line starts recorded per round: 512 (4 KB) instead of 1024 lets a CU hold
// 4 workgroups instead of 3 (11.1 vs 12.7 ms per build at n = 6e9, sigma =
// 256; 10.6 ms with the write-out's word index in 32 bits)
constexpr uint32_t kQsPos = 512;

__device__ __forceinline__ uint64_t qs_digit_word(const uint4* s_ql, uint32_t l, uint32_t wi) {
  const uint4 q = s_ql[l * 8 + 1 + (wi >> 1)];
  return (wi & 1) ? ((uint64_t)q.z | ((uint64_t)q.w << 32)) : ((uint64_t)q.x | ((uint64_t)q.y << 32));
}

__device__ __forceinline__ void quad_select_lines_wg(const QuadSelBuildParams& qp, uint64_t blk) {
  __shared__ __align__(16) uint4 s_ql[(kQsQuadLines + 1) * 8];
  __shared__ uint64_t s_pos[kQsPos];
  const uint32_t tid = threadIdx.x;
  const uint64_t L0 = blk * kQsQuadLines;
  for (uint32_t i = tid; i < (kQsQuadLines + 1) * 8; i += blockDim.x) {
    const uint64_t gi = L0 * 8 + i;
    s_ql[i] = gi < qp.nql * 8 ? detail::gld(qp.qlines, gi) : make_uint4(0, 0, 0, 0);
  }
  __syncthreads();
  const uint32_t lend = static_cast<uint32_t>(qp.nql - L0 < kQsQuadLines ? qp.nql - L0 : kQsQuadLines);
  // the four digits' lines in turn, over the same staged quad lines
  for (uint32_t d = 0; d < 4; ++d) {
    const uint32_t r = qp.r[d];
    if (r == 0) continue;
    const uint64_t magic = qp.magic[d];
    // d's before staged quad line l (the total past the last one)
    auto before = [&](uint32_t l) -> uint64_t {
      if (L0 + l >= qp.nql) return qp.total[d];
      const uint4 h = s_ql[l * 8];
      const uint64_t h0 = (uint64_t)h.x | ((uint64_t)h.y << 32);
      const uint64_t h1 = (uint64_t)h.z | ((uint64_t)h.w << 32);
      const uint64_t m40 = (1ull << 40) - 1;
      const uint64_t c0 = h0 & m40, c1 = (h0 >> 40) | ((h1 & 0xFFFFull) << 24),
                     c2 = (h1 >> 16) & m40;
      return d == 0 ? c0 : d == 1 ? c1 : d == 2 ? c2 : (L0 + l) * kQuadDigits - c0 - c1 - c2;
    };
    const uint64_t first = L0 == 0 ? 0 : div_rate(before(0), r, magic) + 1;
    const uint64_t T = div_rate(before(lend), r, magic) + 1 - first;
    uint64_t jlo = 1, jend = 0, cb = 0;  // empty
    if (tid < lend) {
      cb = before(tid);
      jlo = L0 + tid == 0 ? 0 : div_rate(cb, r, magic) + 1;
      jend = div_rate(before(tid + 1), r, magic);
    }
    for (uint64_t r0 = 0; r0 < T; r0 += kQsPos) {
      for (uint64_t j = jlo; j <= jend; ++j) {
        const uint64_t slot = j - first;
        if (slot < r0) continue;
        if (slot >= r0 + kQsPos) break;
        uint64_t pos = 0;
        if (j != 0) {
          uint32_t rank = static_cast<uint32_t>(j * r - cb);  // 1-based, <= 448
          for (uint32_t k = 0; k < kQuadDigits / 32; ++k) {
            const uint64_t m = quad_match(qs_digit_word(s_ql, tid, k), d);
            const uint32_t c = static_cast<uint32_t>(__popcll(m));
            if (rank <= c) {
              pos = (L0 + tid) * kQuadDigits + 32 * k + (detail::select_in_word(m, rank) >> 1);
              break;
            }
            rank -= c;
          }
        }
        s_pos[slot - r0] = pos;
      }
      __syncthreads();
      // piece 0: the position; piece k > 0: digit words 2k - 2, 2k - 1 from
      // the word holding the position
      const uint64_t nl = T - r0 < kQsPos ? T - r0 : kQsPos;
      uint4* dst = qp.out[d] + (first + r0) * 8;
      for (uint32_t i = tid; i < nl * 8; i += blockDim.x) {
        const uint64_t pos = s_pos[i >> 3];
        const uint32_t k = i & 7;
        uint4 q = make_uint4(static_cast<uint32_t>(pos), static_cast<uint32_t>(pos >> 32), 0u, 0u);
        if (k != 0) {
          // digit word from the first staged quad line (< 257 * 14)
          const uint32_t w0 = static_cast<uint32_t>((pos >> 5) - L0 * 14) + 2 * (k - 1);
          const uint32_t a = w0 / 14, b = (w0 + 1) / 14;
          const uint64_t x = qs_digit_word(s_ql, a, w0 - a * 14);
          const uint64_t y = qs_digit_word(s_ql, b, w0 + 1 - b * 14);
          q = make_uint4(static_cast<uint32_t>(x), static_cast<uint32_t>(x >> 32),
                         static_cast<uint32_t>(y), static_cast<uint32_t>(y >> 32));
        }
        dst[i] = q;
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void quad_select_lines_kernel(QuadSelBuildParams qp) {
  quad_select_lines_wg(qp, blockIdx.x);
}

void launch_quad_select_lines(const QuadSelBuildParams& qp, hipStream_t s) {
  if (qp.nql == 0) return;
  const uint64_t blocks = (qp.nql + kQsQuadLines - 1) / kQsQuadLines;
  if (blocks > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit array too large");
  hipLaunchKernelGGL(quad_select_lines_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, qp);
  ECL_HIP(hipGetLastError());
}

uint64_t quad_select_build_workgroups(uint64_t nql) {
  return (nql + kQsQuadLines - 1) / kQsQuadLines;
}

// The quad lines and quad select lines of many quad levels (those of the
// trees of a forest) in three launches.  Job a (one quad level of one tree)
// owns the build workgroups [wg_start[a], wg_start[a + 1]) of the quad-line
// launches and [sel_start[a], sel_start[a + 1]) of the select-line launch;
// a workgroup (a thread, for the per-workgroup information) finds its job
// by a search over the starts and runs the one-level code on its part.
__global__ __launch_bounds__(256) void multi_quad_wg_kernel(const QuadBuildParams* __restrict__ jobs,
                                                            const uint64_t* __restrict__ wg_start,
                                                            uint32_t njobs, uint64_t total_wg) {
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total_wg) return;
  const uint32_t a = owner_job(wg_start, njobs, t);
  const QuadBuildParams qp = jobs[a];
  quad_wg_one(qp, t - wg_start[a]);
}
__global__ __launch_bounds__(256) void multi_quad_lines_kernel(const QuadBuildParams* __restrict__ jobs,
                                                               const uint64_t* __restrict__ wg_start,
                                                               uint32_t njobs) {
  const uint32_t a = owner_job(wg_start, njobs, blockIdx.x);
  const QuadBuildParams qp = jobs[a];
  quad_lines_wg(qp, blockIdx.x - wg_start[a]);
}
__global__ __launch_bounds__(256) void multi_quad_select_lines_kernel(
    const QuadSelBuildParams* __restrict__ jobs, const uint64_t* __restrict__ sel_start,
    uint32_t njobs) {
  const uint32_t a = owner_job(sel_start, njobs, blockIdx.x);
  // (read in place: a copy's per-digit arrays, indexed at run time, would
  // live in scratch memory)
  quad_select_lines_wg(jobs[a], blockIdx.x - sel_start[a]);
}

void launch_multi_quad_layouts(const QuadBuildParams* jobs, const QuadSelBuildParams* sel_jobs,
                               const uint64_t* wg_start, const uint64_t* sel_start,
                               uint32_t njobs, uint64_t total_wg, uint64_t total_sel,
                               hipStream_t s) {
  if (njobs == 0) return;
  if (total_wg > 0x7FFFFFFFull || total_sel > 0x7FFFFFFFull)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit arrays too large");
  if (total_wg) {
    hipLaunchKernelGGL(multi_quad_wg_kernel, dim3(static_cast<uint32_t>((total_wg + 255) / 256)),
                       dim3(256), 0, s, jobs, wg_start, njobs, total_wg);
    hipLaunchKernelGGL(multi_quad_lines_kernel, dim3(static_cast<uint32_t>(total_wg)), dim3(256), 0,
                       s, jobs, wg_start, njobs);
  }
  if (total_sel)
    hipLaunchKernelGGL(multi_quad_select_lines_kernel, dim3(static_cast<uint32_t>(total_sel)),
                       dim3(256), 0, s, sel_jobs, sel_start, njobs);
  ECL_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Hex lines of one hex level h from the quad lines of quad levels 2h (qa)
// and 2h + 1 (qb).  Four lanes build a hex line, lane q the digits
// [32 q, 32 q + 32) (one 32-digit word of qa) and the counts of hex digits
// 4q .. 4q + 3 before the line.

// digit word g of a quad level (14 per line, after the line's 16-byte header)
__device__ __forceinline__ uint64_t quad_word(const uint4* ql, uint64_t g) {
  const uint64_t line = __umul64hi(g >> 1, 0x2492492492492493ull);  // g / 14
  const uint64_t k = g - line * 14;
  return detail::gld(reinterpret_cast<const uint64_t*>(ql), line * 16 + 2 + k);
}
// the d's before position e of a quad level, for each d (one lane): the
// header and words 0..7 load together, words 8..13 only when needed
__device__ __forceinline__ void quad_count_words(uint4 v, uint32_t k0, uint32_t kw, uint64_t pm,
                                                 uint32_t (&n)[4]) {
#pragma unroll
  for (uint32_t j = 0; j < 2; ++j) {
    const uint32_t k = k0 + j;
    const uint64_t w = j ? ((uint64_t)v.z | ((uint64_t)v.w << 32))
                         : ((uint64_t)v.x | ((uint64_t)v.y << 32));
    const uint64_t m = k < kw ? 0x5555555555555555ull : (k == kw ? pm : 0ull);
    const uint64_t lo = w & m, hi = (w >> 1) & m;
    n[1] += static_cast<uint32_t>(__popcll(lo & ~hi));
    n[2] += static_cast<uint32_t>(__popcll(hi & ~lo));
    n[3] += static_cast<uint32_t>(__popcll(lo & hi));
    n[0] += static_cast<uint32_t>(__popcll(~(lo | hi) & m));
  }
}
__device__ __forceinline__ void quad_counts(const uint4* ql, uint64_t e, uint64_t (&c)[4]) {
  const uint64_t line = __umul64hi(e >> 6, 0x2492492492492493ull);  // e / 448
  const uint32_t o = static_cast<uint32_t>(e - line * kQuadDigits);
  const uint4* lp = ql + line * 8;
  const uint32_t kw = o >> 5, sh = 2 * (o & 31);
  const uint64_t pm = (sh ? ~0ull >> (64 - sh) : 0ull) & 0x5555555555555555ull;
  const uint4 hd = detail::gld(lp, 0), v1 = detail::gld(lp, 1), v2 = detail::gld(lp, 2),
              v3 = detail::gld(lp, 3), v4 = detail::gld(lp, 4);
  uint32_t n[4] = {0, 0, 0, 0};
  quad_count_words(v1, 0, kw, pm, n);
  quad_count_words(v2, 2, kw, pm, n);
  quad_count_words(v3, 4, kw, pm, n);
  quad_count_words(v4, 6, kw, pm, n);
  if (kw >= 8) {
    const uint4 v5 = detail::gld(lp, 5), v6 = detail::gld(lp, 6), v7 = detail::gld(lp, 7);
    quad_count_words(v5, 8, kw, pm, n);
    quad_count_words(v6, 10, kw, pm, n);
    quad_count_words(v7, 12, kw, pm, n);
  }
  const uint64_t h0 = (uint64_t)hd.x | ((uint64_t)hd.y << 32);
  const uint64_t h1 = (uint64_t)hd.z | ((uint64_t)hd.w << 32);
  const uint64_t m40 = (1ull << 40) - 1;
  c[0] = h0 & m40;
  c[1] = (h0 >> 40) | ((h1 & 0xFFFFull) << 24);
  c[2] = (h1 >> 16) & m40;
  c[3] = line * kQuadDigits - c[0] - c[1] - c[2];
  c[0] += n[0];
  c[1] += n[1];
  c[2] += n[2];
  c[3] += n[3];
}
// 32 digits of a quad level from position x on (nwords: the level's words)
__device__ __forceinline__ uint64_t quad_window(const uint4* ql, uint64_t nwords, uint64_t x) {
  const uint64_t g = x >> 5;
  const uint32_t sh = 2 * static_cast<uint32_t>(x & 31);
  const uint64_t a = quad_word(ql, g);
  if (sh == 0) return a;
  const uint64_t b = g + 1 < nwords ? quad_word(ql, g + 1) : 0ull;
  return (a >> sh) | (b << (64 - sh));
}

// One lane's part of a hex line, everything read from global memory: the
// path of workgroups whose positions hold a node start.  t = 4 l + q.
__device__ __forceinline__ void hex_lane(const HexBuildParams& hp, uint64_t t) {
  const uint64_t l = t >> 2;  // hex line
  const uint32_t q = static_cast<uint32_t>(t & 3);
  // the four lanes of a line are all present or all absent
  if (l >= hp.nlines) return;
  const uint64_t n = hp.n;
  const uint64_t P = l * kHexDigits + 32 * q;  // this lane's first position
  const uint64_t Pc = P < n ? P : n;
  // node of Pc: the last node starting at or before it
  uint64_t N = 0;
  {
    uint64_t hi = hp.nodes;
    while (hi - N > 1) {
      const uint64_t mid = (N + hi) >> 1;
      if (detail::gld(hp.S, mid) <= Pc) N = mid;
      else hi = mid;
    }
  }
  uint64_t g[4];
  quad_counts(hp.qa, Pc, g);
  // ---- counts of hex digits 4q .. 4q + 3 before the line's first position
  // (lane 0's node and quad counts)
  const uint32_t src = (threadIdx.x & 63u) & ~3u;
  const uint64_t N0 = __shfl(N, src, 64);
  const uint64_t g00 = __shfl(g[0], src, 64), g01 = __shfl(g[1], src, 64),
                 g02 = __shfl(g[2], src, 64), g03 = __shfl(g[3], src, 64);
  {
    const uint64_t gq = q == 0 ? g00 : q == 1 ? g01 : q == 2 ? g02 : g03;
    const uint64_t e = detail::gld(hp.qbase, 4 * N0 + q) + gq;
    uint64_t c2[4];
    quad_counts(hp.qb, e, c2);
    uint64_t cnt[4];
#pragma unroll
    for (uint32_t d1 = 0; d1 < 4; ++d1)
      cnt[d1] = detail::gld(hp.chd, 16 * N0 + 4 * q + d1) + c2[d1] -
                detail::gld(hp.qcd, 4 * (4 * N0 + q) + d1);
    hp.out[l * 8 + (q < 2 ? q : 2 + q)] =
        make_uint4(static_cast<uint32_t>(cnt[0]), static_cast<uint32_t>(cnt[1]),
                   static_cast<uint32_t>(cnt[2]), static_cast<uint32_t>(cnt[3]));
    if ((l & ((1ull << kHexSbShift) - 1)) == 0) {
#pragma unroll
      for (uint32_t d1 = 0; d1 < 4; ++d1) hp.sb[(l >> kHexSbShift) * 16 + 4 * q + d1] = cnt[d1];
    }
  }
  // ---- the 32 digits: position P + j with quad digit d0 goes to position
  // qbase[4N + d0] + (d0's before it) of quad level 2h + 1
  const uint32_t valid = P < n ? (n - P < 32 ? static_cast<uint32_t>(n - P) : 32u) : 0u;
  const uint64_t w = valid ? quad_word(hp.qa, P >> 5) : 0ull;
  uint64_t lo = 0, hi = 0;
  const uint64_t nend = N + 1 < hp.nodes ? detail::gld(hp.S, N + 1) : n;
  if (valid && nend >= P + valid) {
    // one node: the positions of each d0 take consecutive digits of one run
    const uint64_t nwords = hp.nql * 14;
    uint64_t w0 = quad_window(hp.qb, nwords, detail::gld(hp.qbase, 4 * N + 0) + g[0]);
    uint64_t w1 = quad_window(hp.qb, nwords, detail::gld(hp.qbase, 4 * N + 1) + g[1]);
    uint64_t w2 = quad_window(hp.qb, nwords, detail::gld(hp.qbase, 4 * N + 2) + g[2]);
    uint64_t w3 = quad_window(hp.qb, nwords, detail::gld(hp.qbase, 4 * N + 3) + g[3]);
#pragma unroll
    for (uint32_t j = 0; j < 32; ++j) {
      if (j < valid) {
        const uint32_t d0 = static_cast<uint32_t>(w >> (2 * j)) & 3u;
        const uint64_t wd = d0 == 0 ? w0 : d0 == 1 ? w1 : d0 == 2 ? w2 : w3;
        const uint64_t x = 4 * d0 + (static_cast<uint32_t>(wd) & 3u);
        w0 = d0 == 0 ? w0 >> 2 : w0;
        w1 = d0 == 1 ? w1 >> 2 : w1;
        w2 = d0 == 2 ? w2 >> 2 : w2;
        w3 = d0 == 3 ? w3 >> 2 : w3;
        if (j < 16) lo |= x << (4 * j);
        else hi |= x << (4 * (j - 16));
      }
    }
  } else {
    // a node starts inside the word: position by position
    for (uint32_t j = 0; j < valid; ++j) {
      const uint64_t pj = P + j;
      while (N + 1 < hp.nodes && detail::gld(hp.S, N + 1) <= pj) ++N;
      const uint32_t d0 = static_cast<uint32_t>(w >> (2 * j)) & 3u;
      const uint64_t gd = d0 == 0 ? g[0] : d0 == 1 ? g[1] : d0 == 2 ? g[2] : g[3];
      const uint64_t e = detail::gld(hp.qbase, 4 * N + d0) + gd;
      g[0] += d0 == 0;
      g[1] += d0 == 1;
      g[2] += d0 == 2;
      g[3] += d0 == 3;
      const uint64_t d1 = (quad_word(hp.qb, e >> 5) >> (2 * (e & 31))) & 3u;
      const uint64_t x = 4 * d0 + d1;
      if (j < 16) lo |= x << (4 * j);
      else hi |= x << (4 * (j - 16));
    }
  }
  hp.out[l * 8 + (q < 2 ? 2 + q : 4 + q)] =
      make_uint4(static_cast<uint32_t>(lo), static_cast<uint32_t>(lo >> 32),
                 static_cast<uint32_t>(hi), static_cast<uint32_t>(hi >> 32));
}

// The workgroup form: 64 hex lines (8192 positions) per workgroup, one
// 32-position word per thread.  When the positions lie in one node N, the
// workgroup stages the quad lines of qa holding them and, per digit d0, the
// quad lines of qb holding the run its d0's map to (from qbase[4N + d0] +
// the d0's before the first position on); a block scan of the words' digit
// counts gives every word its place in each run, and a block scan of the
// words' hex-digit counts gives every line its header.  Otherwise (a node
// starts inside) the lanes take hex_lane.
constexpr uint32_t kHexWgLines = 64;
constexpr uint32_t kHexWgPos = kHexWgLines * kHexDigits;         // 8192
constexpr uint32_t kHexQaLines = kHexWgPos / kQuadDigits + 2;   // 20
constexpr uint32_t kHexQbLines = 32;  // the four runs need at most 26

// exclusive scans over the workgroup (4 waves) of u64 values made of 16-bit
// fields whose sums stay below 2^16; s_w holds 4 * J u64
template <int J>
__device__ __forceinline__ void wg_excl_scan_fields(uint64_t (&v)[J], uint64_t* s_w,
                                                    uint64_t (&total)[J]) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t incl[J];
#pragma unroll
  for (int j = 0; j < J; ++j) incl[j] = detail::wave_incl_scan_2x32(v[j]);
  if (lane == 63) {
#pragma unroll
    for (int j = 0; j < J; ++j) s_w[J * wave + j] = incl[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < J; ++j) {
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
      const uint64_t x = s_w[J * w + j];
      tot += x;
      before += w < wave ? x : 0ull;
    }
    total[j] = tot;
    v[j] = before + incl[j] - v[j];
  }
  __syncthreads();
}

// Occupancy target: 7 waves per SIMD (70 VGPRs, no spills; 8 waves spill 7
// and are no faster, 12.0 vs 12.1 ms for both levels of n = 6e9, sigma =
// 256; the compiler's own choice is 5 waves).
__device__ __forceinline__ void hex_lines_wg(const HexBuildParams& hp, uint64_t blk) {
  __shared__ __align__(16) uint4 s_qa[kHexQaLines * 8];
  __shared__ __align__(16) uint4 s_qb[kHexQbLines * 8];
  __shared__ uint64_t s_w[16];
  __shared__ uint64_t s_info[1];  // staged qb lines
  __shared__ uint64_t s_R[4], s_rb0[4], s_hx0[16];
  __shared__ uint32_t s_rl[4];
  __shared__ uint32_t s_dlut[256];
  const uint32_t tid = threadIdx.x;
  // deposit table for 4 positions with first digits p (2 bits each): per
  // position i the index 4 d0_i + r_i of its second digit among the next 4
  // of each run (r_i = the earlier positions with the same d0), and per run
  // the digits the 4 positions take (3 bits each, from bit 16)
  {
    uint32_t e = 0, cnt[4] = {0, 0, 0, 0};
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t d0 = (tid >> (2 * i)) & 3u;
      const uint32_t r = d0 == 0 ? cnt[0] : d0 == 1 ? cnt[1] : d0 == 2 ? cnt[2] : cnt[3];
      e |= (4 * d0 + r) << (4 * i);
      cnt[0] += d0 == 0;
      cnt[1] += d0 == 1;
      cnt[2] += d0 == 2;
      cnt[3] += d0 == 3;
    }
    s_dlut[tid] = e | (cnt[0] << 16) | (cnt[1] << 19) | (cnt[2] << 22) | (cnt[3] << 25);
  }
  const uint64_t L0 = blk * kHexWgLines;
  const uint64_t P0 = L0 * kHexDigits;
  const uint64_t n = hp.n;
  const uint64_t E = P0 + kHexWgPos < n ? P0 + kHexWgPos : n;  // <= P0 only when P0 >= n
  // node of P0 and of E - 1: the node starts S[1 .. nodes - 1] at or before
  // them, counted over the workgroup (S is nondecreasing, S[0] = 0)
  uint32_t na0 = 0, na1 = 0;
  for (uint64_t k0 = 1; k0 < hp.nodes; k0 += blockDim.x) {
    const uint64_t k = k0 + tid;
    const uint64_t sk = k < hp.nodes ? detail::gld(hp.S, k) : ~0ull;
    na0 += __syncthreads_count(sk <= P0);
    na1 += __syncthreads_count(E > P0 && sk <= E - 1);
  }
  if (P0 >= n || na0 != na1) {  // the whole workgroup
    hex_lane(hp, blk * blockDim.x + tid);
    return;
  }
  const uint64_t N = na0;
  // table entries used below, issued ahead
  const uint64_t qb_d = tid < 4 ? detail::gld(hp.qbase, 4 * N + tid) : 0ull;
  const uint64_t chd_x = tid < 16 ? detail::gld(hp.chd, 16 * N + tid) : 0ull;
  const uint64_t qcd_x = tid < 16 ? detail::gld(hp.qcd, 4 * (4 * N + (tid >> 2)) + (tid & 3)) : 0ull;
  // ---- stage the quad lines of qa holding [P0, E)
  const uint64_t a0 = __umul64hi(P0 >> 6, 0x2492492492492493ull);
  const uint64_t a1 = __umul64hi((E - 1) >> 6, 0x2492492492492493ull);
  const uint32_t na = static_cast<uint32_t>(a1 - a0 + 1);
  for (uint32_t i = tid; i < na * 8; i += blockDim.x) s_qa[i] = detail::gld(hp.qa, a0 * 8 + i);
  __syncthreads();
  // ---- this thread's word and its digit counts (4 x 16-bit fields)
  const uint64_t P = P0 + 32ull * tid;
  const uint32_t valid = P < E ? (E - P < 32 ? static_cast<uint32_t>(E - P) : 32u) : 0u;
  uint64_t w = 0;
  if (valid) {
    const uint64_t qi = __umul64hi(P >> 6, 0x2492492492492493ull);
    w = qs_digit_word(s_qa, static_cast<uint32_t>(qi - a0),
                      static_cast<uint32_t>(P - qi * kQuadDigits) >> 5);
  }
  const uint64_t vm = (valid == 32 ? ~0ull : (1ull << (2 * valid)) - 1) & 0x5555555555555555ull;
  uint64_t pre[1], ctot[1];
  {
    const uint64_t lo = w & vm, hi = (w >> 1) & vm;
    pre[0] = static_cast<uint64_t>(__popcll(~(lo | hi) & vm)) |
             (static_cast<uint64_t>(__popcll(lo & ~hi)) << 16) |
             (static_cast<uint64_t>(__popcll(hi & ~lo)) << 32) |
             (static_cast<uint64_t>(__popcll(lo & hi)) << 48);
  }
  const uint64_t cword = pre[0];           // the word's own d's
  wg_excl_scan_fields<1>(pre, s_w, ctot);  // pre: the d's in [P0, P)
  // ---- the d's before P0 and the runs R_d of qb (threads 0..3)
  if (tid < 4) {
    const uint32_t d = tid;
    const uint4 h = s_qa[0];
    const uint64_t h0 = (uint64_t)h.x | ((uint64_t)h.y << 32);
    const uint64_t h1 = (uint64_t)h.z | ((uint64_t)h.w << 32);
    const uint64_t m40 = (1ull << 40) - 1;
    const uint64_t c0 = h0 & m40, c1 = (h0 >> 40) | ((h1 & 0xFFFFull) << 24), c2 = (h1 >> 16) & m40;
    uint64_t g = d == 0 ? c0 : d == 1 ? c1 : d == 2 ? c2 : a0 * kQuadDigits - c0 - c1 - c2;
    const uint32_t kw = static_cast<uint32_t>(P0 - a0 * kQuadDigits) >> 5;
    for (uint32_t k = 0; k < kw; ++k) g += static_cast<uint64_t>(__popcll(quad_match(qs_digit_word(s_qa, 0, k), d)));
    s_R[d] = qb_d + g;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t at = 0;
    for (uint32_t d = 0; d < 4; ++d) {
      const uint64_t R = s_R[d], C = (ctot[0] >> (16 * d)) & 0xFFFFull;
      const uint64_t b0 = __umul64hi(R >> 6, 0x2492492492492493ull);
      uint64_t b1 = __umul64hi((R + C + 31) >> 6, 0x2492492492492493ull);
      if (b1 > hp.nql - 1) b1 = hp.nql - 1;
      s_rb0[d] = b0;
      s_rl[d] = at;
      at += static_cast<uint32_t>(b1 - b0 + 1);
    }
    s_info[0] = at;
  }
  __syncthreads();
  const uint32_t nst = static_cast<uint32_t>(s_info[0]);
  if (nst > kHexQbLines) {  // not reached (at most 26); the whole workgroup
    hex_lane(hp, blk * blockDim.x + tid);
    return;
  }
  for (uint32_t i = tid; i < nst * 8; i += blockDim.x) {
    const uint32_t li = i >> 3;
    const uint32_t d = li >= s_rl[3] ? 3u : li >= s_rl[2] ? 2u : li >= s_rl[1] ? 1u : 0u;
    s_qb[i] = detail::gld(hp.qb, (s_rb0[d] + (li - s_rl[d])) * 8 + (i & 7));
  }
  __syncthreads();
  // ---- hex digits 4 x + d1 before P0 (threads 0..15): the d1's of qb before R_d0
  if (tid < 16) {
    const uint32_t d0 = tid >> 2, d1 = tid & 3;
    const uint64_t R = s_R[d0], b0 = s_rb0[d0];
    const uint32_t li = s_rl[d0];
    const uint4 h = s_qb[li * 8];
    const uint64_t h0 = (uint64_t)h.x | ((uint64_t)h.y << 32);
    const uint64_t h1 = (uint64_t)h.z | ((uint64_t)h.w << 32);
    const uint64_t m40 = (1ull << 40) - 1;
    const uint64_t c0 = h0 & m40, c1 = (h0 >> 40) | ((h1 & 0xFFFFull) << 24), c2 = (h1 >> 16) & m40;
    uint64_t g = d1 == 0 ? c0 : d1 == 1 ? c1 : d1 == 2 ? c2 : b0 * kQuadDigits - c0 - c1 - c2;
    const uint32_t o = static_cast<uint32_t>(R - b0 * kQuadDigits);
    const uint32_t kw = o >> 5, sh = 2 * (o & 31);
    for (uint32_t k = 0; k < kw; ++k) g += static_cast<uint64_t>(__popcll(quad_match(qs_digit_word(s_qb, li, k), d1)));
    if (sh) g += static_cast<uint64_t>(__popcll(quad_match(qs_digit_word(s_qb, li, kw), d1) & (~0ull >> (64 - sh))));
    s_hx0[tid] = chd_x + g - qcd_x;
  }
  __syncthreads();
  // ---- the word's 32 hex digits: its d0's take the next digits of run d0
  auto window = [&](uint32_t d) -> uint64_t {
    const uint64_t x = s_R[d] + ((pre[0] >> (16 * d)) & 0xFFFFull);
    const uint64_t xl = __umul64hi(x >> 6, 0x2492492492492493ull);
    const uint32_t o = static_cast<uint32_t>(x - xl * kQuadDigits);
    const uint32_t k = o >> 5, sh = 2 * (o & 31);
    const uint32_t li = s_rl[d] + static_cast<uint32_t>(xl - s_rb0[d]);
    const uint64_t a = qs_digit_word(s_qb, li, k);
    if (sh == 0) return a;
    // the next word (staged whenever one of its digits is used)
    const uint64_t b = k < 13 ? qs_digit_word(s_qb, li, k + 1) : qs_digit_word(s_qb, li + 1, 0);
    return (a >> sh) | (b << (64 - sh));
  };
  const uint64_t W[4] = {window(0), window(1), window(2), window(3)};
  // 4 positions at a time: the next 4 digits of each window side by side in
  // C, each position's digit picked by its table index, then each window
  // advanced by the digits its positions took.  Positions past the valid
  // ones take digits too (from the padding) and are masked off.
  uint64_t lo = 0, hi = 0;
  {
    uint64_t S0 = W[0], S1 = W[1], S2 = W[2], S3 = W[3];
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) {
      const uint32_t L = s_dlut[static_cast<uint32_t>(w >> (8 * c)) & 0xFFu];
      const uint32_t C = (static_cast<uint32_t>(S0) & 0xFFu) | ((static_cast<uint32_t>(S1) & 0xFFu) << 8) |
                         ((static_cast<uint32_t>(S2) & 0xFFu) << 16) |
                         ((static_cast<uint32_t>(S3) & 0xFFu) << 24);
      uint32_t out = 0;
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i) {
        const uint32_t idx = (L >> (4 * i)) & 15u;
        out |= ((idx & 12u) | ((C >> (2 * idx)) & 3u)) << (4 * i);
      }
      S0 >>= 2 * ((L >> 16) & 7u);
      S1 >>= 2 * ((L >> 19) & 7u);
      S2 >>= 2 * ((L >> 22) & 7u);
      S3 >>= 2 * ((L >> 25) & 7u);
      if (c < 4) lo |= static_cast<uint64_t>(out) << (16 * c);
      else hi |= static_cast<uint64_t>(out) << (16 * (c - 4));
    }
    lo &= valid >= 16 ? ~0ull : (1ull << (4 * valid)) - 1;
    hi &= valid >= 32 ? ~0ull : valid > 16 ? (1ull << (4 * (valid - 16))) - 1 : 0ull;
  }
  // ---- the word's hex-digit counts (field d1 of hh[d0]): the d1's among the
  // digits its d0's used from window d0; scanned
  uint64_t hh[4], htot[4];
#pragma unroll
  for (uint32_t d0 = 0; d0 < 4; ++d0) {
    const uint32_t c = static_cast<uint32_t>(cword >> (16 * d0)) & 0xFFFFu;  // <= 32
    const uint64_t m = c >= 32 ? ~0ull : (1ull << (2 * c)) - 1;
    uint64_t f = 0;
#pragma unroll
    for (uint32_t d1 = 0; d1 < 4; ++d1)
      f |= static_cast<uint64_t>(__popcll(quad_match(W[d0], d1) & m)) << (16 * d1);
    hh[d0] = f;
  }
  wg_excl_scan_fields<4>(hh, s_w, htot);  // hh: the hex digits in [P0, P)
  // ---- write: the digits, and the counts of hex digits 4q .. 4q + 3 before
  // the line (its first word's prefix, from lane 4 (tid / 4))
  const uint64_t l = L0 + (tid >> 2);
  const uint32_t q = tid & 3;
  const uint32_t src = (tid & 63u) & ~3u;
  const uint64_t f0 = __shfl(hh[0], src, 64), f1 = __shfl(hh[1], src, 64),
                 f2 = __shfl(hh[2], src, 64), f3 = __shfl(hh[3], src, 64);
  const uint64_t mine = q == 0 ? f0 : q == 1 ? f1 : q == 2 ? f2 : f3;
  if (l < hp.nlines) {
    hp.out[l * 8 + (q < 2 ? 2 + q : 4 + q)] =
        make_uint4(static_cast<uint32_t>(lo), static_cast<uint32_t>(lo >> 32),
                   static_cast<uint32_t>(hi), static_cast<uint32_t>(hi >> 32));
    uint64_t cnt[4];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) cnt[j] = s_hx0[4 * q + j] + ((mine >> (16 * j)) & 0xFFFFull);
    hp.out[l * 8 + (q < 2 ? q : 2 + q)] =
        make_uint4(static_cast<uint32_t>(cnt[0]), static_cast<uint32_t>(cnt[1]),
                   static_cast<uint32_t>(cnt[2]), static_cast<uint32_t>(cnt[3]));
    if ((l & ((1ull << kHexSbShift) - 1)) == 0) {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) hp.sb[(l >> kHexSbShift) * 16 + 4 * q + j] = cnt[j];
    }
  }
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void hex_lines_kernel(
    HexBuildParams hp) {
  hex_lines_wg(hp, blockIdx.x);
}
// The hex lines of many hex levels (a forest's): job a owns the build
// workgroups [wg_start[a], wg_start[a + 1]); a workgroup finds its job by a
// search over the starts and runs the one-level code on its part.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void multi_hex_lines_kernel(
    const HexBuildParams* __restrict__ jobs, const uint64_t* __restrict__ wg_start, uint32_t njobs) {
  const uint32_t a = owner_job(wg_start, njobs, blockIdx.x);
  const HexBuildParams hp = jobs[a];
  hex_lines_wg(hp, blockIdx.x - wg_start[a]);
}

uint64_t hex_build_workgroups(uint64_t nlines) { return (nlines + kHexWgLines - 1) / kHexWgLines; }

void launch_multi_hex_lines(const HexBuildParams* jobs, const uint64_t* wg_start, uint32_t njobs,
                            uint64_t total_wg, hipStream_t s) {
  if (njobs == 0 || total_wg == 0) return;
  if (total_wg > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit arrays too large");
  hipLaunchKernelGGL(multi_hex_lines_kernel, dim3(static_cast<uint32_t>(total_wg)), dim3(256), 0, s,
                     jobs, wg_start, njobs);
  ECL_HIP(hipGetLastError());
}

void launch_hex_lines(const HexBuildParams& hp, hipStream_t s) {
  if (hp.nlines == 0) return;
  const uint64_t blocks = (hp.nlines + kHexWgLines - 1) / kHexWgLines;
  if (blocks > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit array too large");
  hipLaunchKernelGGL(hex_lines_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, hp);
  ECL_HIP(hipGetLastError());
}

void launch_query_lines(const LevelView& v, uint4* lines, uint64_t nlines, hipStream_t s) {
  if (nlines == 0) return;
  const uint64_t blocks = (nlines * 8 + 255) / 256;
  if (blocks > 0x7FFFFFFFull) throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit array too large");
  LevelView plain = v;
  plain.lines = nullptr;  // headers come from the directory
  plain.sel_hints = nullptr;
  hipLaunchKernelGGL(query_lines_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s,
                     plain, lines, nlines);
  ECL_HIP(hipGetLastError());
}

void launch_upload(const UploadBlob& b, hipStream_t s) {
  if (b.nseg == 0) return;
  hipLaunchKernelGGL(upload_kernel, dim3(1), dim3(256), 0, s, b);
  ECL_HIP(hipGetLastError());
}

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

void launch_rs_build(const RsArrayDesc* arrays, uint64_t narrays, uint64_t nblocks,
                     uint64_t nchunks, uint32_t* scratch, hipStream_t s) {
  if (nblocks == 0) return;
  const uint64_t groups = (nblocks + kRsWaves - 1) / kRsWaves;
  if (groups > 0x7FFFFFFFull || narrays > 0x7FFFFFFFull || nchunks > 0x7FFFFFFFull)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "bit arrays too large");
  uint32_t* block_array = scratch;
  uint32_t* block_ones = scratch + nblocks;
  ProfScope ps(K_RS_BUILD, s);
  hipLaunchKernelGGL(rs_block_map_kernel, dim3(static_cast<uint32_t>(narrays)), dim3(256), 0, s,
                     arrays, block_array);
  ECL_HIP(hipGetLastError());
  hipLaunchKernelGGL(rs_local_kernel, dim3(static_cast<uint32_t>(groups)), dim3(512), 0, s,
                     arrays, block_array, nblocks, block_ones);
  ECL_HIP(hipGetLastError());
  const uint32_t na = static_cast<uint32_t>(narrays);
  hipLaunchKernelGGL(rs_chunk_sum_kernel, dim3(static_cast<uint32_t>(nchunks)), dim3(256), 0, s,
                     arrays, na, block_ones);
  hipLaunchKernelGGL(rs_chunk_scan_kernel, dim3(na), dim3(64), 0, s, arrays);
  hipLaunchKernelGGL(rs_l1_write_kernel, dim3(static_cast<uint32_t>(nchunks)), dim3(256), 0, s,
                     arrays, na, block_ones);
  ECL_HIP(hipGetLastError());
  hipLaunchKernelGGL(rs_samples_kernel, dim3(static_cast<uint32_t>(groups)), dim3(512), 0, s,
                     arrays, block_array, nblocks);
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
