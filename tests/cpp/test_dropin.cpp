// test_dropin.cpp -- the C++ drop-in API exercised the way the reference's
// own tests use it (tests/test_bit_array.cu, test_rank_select.cu,
// test_wavelet_tree.cu, example/example.cu): objects passed by value into
// probe kernels, spans returned by the host batch API, exact exception
// strings.  Expected values are closed forms or direct scans of the text.
// Prints one line per check group and "ALL OK" when every check passed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <limits>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include <stdlib.h>
#include <unistd.h>

#include "ecl_wavelet/bitarray/rank_select.cuh"
#include "ecl_wavelet/tree/wavelet_tree.cuh"
#include "ecl_wavelet/utils/test_benchmark_utils.cuh"

static int g_fail = 0;
static int g_checks = 0;
#define EXPECT_EQ(a, b)                                                              \
  do {                                                                               \
    ++g_checks;                                                                      \
    auto _a = (a);                                                                   \
    auto _b = (b);                                                                   \
    if (!(_a == _b)) {                                                               \
      if (g_fail < 20)                                                               \
        std::printf("FAIL %s:%d: %s == %s (%llu vs %llu)\n", __FILE__, __LINE__, #a, #b, \
                    (unsigned long long)_a, (unsigned long long)_b);                 \
      ++g_fail;                                                                      \
    }                                                                                \
  } while (0)

namespace ecl {

// ---- utils in device code (test_utils.cu:17-57) -----------------------------
// One launch evaluates the device helpers for every i below the limit.
template <typename T>
__global__ void utilsKernel(uint32_t limit, uint64_t* prev, uint64_t* clog) {
  for (uint32_t i = threadIdx.x; i < limit; i += blockDim.x) {
    prev[i] = utils::getPrevPowTwo<T>(static_cast<T>(i));
    clog[i] = i ? utils::ceilLog2<T>(static_cast<T>(i)) : 0;
  }
}

// ---- probe kernels (test_rank_select.cu:21-111) ------------------------------
__global__ void writeWordKernel(BitArray ba, size_t a, size_t i, uint32_t v) { ba.writeWord(a, i, v); }
__global__ void accessKernel(BitArray ba, size_t a, size_t i, size_t* out) { *out = ba.access(a, i); }
__global__ void twoWordsKernel(BitArray ba, size_t a, size_t i, size_t* out) { *out = ba.twoWords(a, i); }
__global__ void partialWordKernel(BitArray ba, uint32_t w, uint8_t b, size_t* out) { *out = ba.partialWord(w, b); }
// one group of NT lanes; every lane writes what it got
template <int NT, int V>
__global__ void rankKernel(RankSelect rs, uint32_t a, size_t i, size_t* out) {
  RankResult r = rs.rank<NT, true, V>(a, i, rs.bit_array_.getOffset(a));
  out[threadIdx.x] = r.rank;
}
template <int NT, int V>
__global__ void selectKernel(RankSelect rs, uint32_t a, size_t i, size_t* out) {
  out[threadIdx.x] = rs.select<V, NT>(a, i, rs.bit_array_.getOffset(a));
}
// many groups of NT lanes, group g answering query g (test_rank_select.cu:
// 21-111 launch one group per query the same way); lane t of group g writes
// out[g * NT + t]
template <int NT, int V>
__global__ void rankGroups(RankSelect rs, uint32_t a, const size_t* idx, size_t m, size_t* out) {
  const size_t step = size_t(gridDim.x) * blockDim.x / NT;
  for (size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / NT; g < m; g += step) {
    RankResult r = rs.rank<NT, true, V>(a, idx[g], rs.bit_array_.getOffset(a));
    out[g * NT + threadIdx.x % NT] = r.rank;
  }
}
template <int NT, int V>
__global__ void selectGroups(RankSelect rs, uint32_t a, const size_t* idx, size_t m, size_t* out) {
  const size_t step = size_t(gridDim.x) * blockDim.x / NT;
  for (size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / NT; g < m; g += step)
    out[g * NT + threadIdx.x % NT] = rs.select<V, NT>(a, idx[g], rs.bit_array_.getOffset(a));
}
__global__ void l1Kernel(RankSelect rs, uint32_t a, size_t i, size_t* out) { *out = rs.getL1Entry(a, i); }
__global__ void l2Kernel(RankSelect rs, uint32_t a, size_t i, size_t* out) { *out = rs.getL2Entry(a, i); }
__global__ void numL1Kernel(RankSelect rs, uint32_t a, size_t* out) { *out = rs.getNumL1Blocks(a); }
__global__ void numL2Kernel(RankSelect rs, uint32_t a, size_t* out) { *out = rs.getNumL2Blocks(a); }
__global__ void lastL2Kernel(RankSelect rs, uint32_t a, size_t* out) { *out = rs.getNumLastL2Blocks(a); }
__global__ void sampleKernel(RankSelect rs, uint32_t a, size_t j, bool v, size_t* out) {
  *out = v ? rs.getSelectSample<1>(a, j) : rs.getSelectSample<0>(a, j);
}
__global__ void totalKernel(RankSelect rs, uint32_t a, bool v, size_t* out) {
  *out = v ? rs.getTotalNumVals<1>(a) : rs.getTotalNumVals<0>(a);
}
// user kernels over a WaveletTree passed by value (README.md:5), TPQ lanes
// per query: lane t of the group answering query g writes out[g * TPQ + t]
template <typename T, int TPQ>
__global__ void userAccess(WaveletTree<T> wt, const size_t* idx, size_t m, size_t* out) {
  const size_t step = size_t(gridDim.x) * blockDim.x / TPQ;
  for (size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / TPQ; g < m; g += step)
    out[g * TPQ + threadIdx.x % TPQ] =
        wt.template deviceAccess<TPQ>(idx[g], utils::isPowTwo(wt.getAlphabetSize()));
}
template <typename T, int TPQ>
__global__ void userRank(WaveletTree<T> wt, const RankSelectQuery<T>* q, size_t m, size_t* out) {
  const size_t step = size_t(gridDim.x) * blockDim.x / TPQ;
  for (size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / TPQ; g < m; g += step)
    out[g * TPQ + threadIdx.x % TPQ] =
        wt.template deviceRank<TPQ>(q[g], utils::isPowTwo(wt.getAlphabetSize()));
}
template <typename T, int TPQ>
__global__ void userSelect(WaveletTree<T> wt, const RankSelectQuery<T>* q, size_t m, size_t* out) {
  const size_t step = size_t(gridDim.x) * blockDim.x / TPQ;
  for (size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / TPQ; g < m; g += step)
    out[g * TPQ + threadIdx.x % TPQ] =
        wt.template deviceSelect<TPQ>(q[g], utils::isPowTwo(wt.getAlphabetSize()));
}
template <typename T, bool P2>
__global__ void nodePosKernel(WaveletTree<T> wt, T s, uint8_t l, size_t* out) {
  *out = wt.template getNodePosAtLevel<P2>(s, l);
}
template <typename T>
__global__ void baAccess(BitArray ba, size_t a, size_t i, size_t* out) { *out = ba.access(a, i); }

}  // namespace ecl

using namespace ecl;

static size_t* g_res = nullptr;

static const int kGroupSizes[] = {1, 2, 4, 8, 16, 32, 64};

// one group of nt lanes; every lane must return the same answer
template <int NT>
static void launch_one(bool sel, RankSelect& rs, int v, uint32_t a, size_t i) {
  if (sel) {
    if (v) selectKernel<NT, 1><<<1, NT>>>(rs, a, i, g_res);
    else selectKernel<NT, 0><<<1, NT>>>(rs, a, i, g_res);
  } else {
    if (v) rankKernel<NT, 1><<<1, NT>>>(rs, a, i, g_res);
    else rankKernel<NT, 0><<<1, NT>>>(rs, a, i, g_res);
  }
}
static size_t run_one(bool sel, RankSelect& rs, int nt, int v, uint32_t a, size_t i) {
  switch (nt) {
    case 1: launch_one<1>(sel, rs, v, a, i); break;
    case 2: launch_one<2>(sel, rs, v, a, i); break;
    case 4: launch_one<4>(sel, rs, v, a, i); break;
    case 8: launch_one<8>(sel, rs, v, a, i); break;
    case 16: launch_one<16>(sel, rs, v, a, i); break;
    case 32: launch_one<32>(sel, rs, v, a, i); break;
    default: launch_one<64>(sel, rs, v, a, i); break;
  }
  kernelCheck();
  for (int t = 1; t < nt; ++t) EXPECT_EQ(g_res[t], g_res[0]);
  return g_res[0];
}
static size_t run_rank(RankSelect& rs, int nt, int v, uint32_t a, size_t i) {
  return run_one(false, rs, nt, v, a, i);
}
static size_t run_select(RankSelect& rs, int nt, int v, uint32_t a, size_t i) {
  return run_one(true, rs, nt, v, a, i);
}

// m queries in groups of nt lanes over 16 workgroups of 64 lanes; out holds
// m * nt answers
template <int NT>
static void launch_groups(bool sel, RankSelect& rs, int v, const size_t* d_idx, size_t m,
                          size_t* d_out) {
  if (sel) {
    if (v) selectGroups<NT, 1><<<16, 64>>>(rs, 0, d_idx, m, d_out);
    else selectGroups<NT, 0><<<16, 64>>>(rs, 0, d_idx, m, d_out);
  } else {
    if (v) rankGroups<NT, 1><<<16, 64>>>(rs, 0, d_idx, m, d_out);
    else rankGroups<NT, 0><<<16, 64>>>(rs, 0, d_idx, m, d_out);
  }
}
static void run_groups(bool sel, RankSelect& rs, int nt, int v, const size_t* d_idx, size_t m,
                       size_t* d_out) {
  switch (nt) {
    case 1: launch_groups<1>(sel, rs, v, d_idx, m, d_out); break;
    case 2: launch_groups<2>(sel, rs, v, d_idx, m, d_out); break;
    case 4: launch_groups<4>(sel, rs, v, d_idx, m, d_out); break;
    case 8: launch_groups<8>(sel, rs, v, d_idx, m, d_out); break;
    case 16: launch_groups<16>(sel, rs, v, d_idx, m, d_out); break;
    case 32: launch_groups<32>(sel, rs, v, d_idx, m, d_out); break;
    default: launch_groups<64>(sel, rs, v, d_idx, m, d_out); break;
  }
  kernelCheck();
}

static void test_bit_array() {
  // test_bit_array.cu:178-230
  BitArray ba(std::vector<size_t>{66000}, false);
  writeWordKernel<<<1, 1>>>(ba, 0, 0, ~0u);
  kernelCheck();
  twoWordsKernel<<<1, 1>>>(ba, 0, 0, g_res);
  kernelCheck();
  EXPECT_EQ(*g_res, size_t(0xFFFFFFFF));
  writeWordKernel<<<1, 1>>>(ba, 0, 0, 0);
  writeWordKernel<<<1, 1>>>(ba, 0, 1, ~0u);
  kernelCheck();
  twoWordsKernel<<<1, 1>>>(ba, 0, 0, g_res);
  kernelCheck();
  EXPECT_EQ(*g_res, size_t(0xFFFFFFFF00000000ull));
  partialWordKernel<<<1, 1>>>(ba, 0xAAAAAAAAu, 4, g_res);
  kernelCheck();
  EXPECT_EQ(*g_res, size_t(0b1010));
  partialWordKernel<<<1, 1>>>(ba, 0xAAAAAAAAu, 9, g_res);
  kernelCheck();
  EXPECT_EQ(*g_res, size_t(0b010101010));
  std::vector<size_t> sizes{2, 4, 8, 64, 128, 1024};
  BitArray b2(sizes, true);
  for (size_t i = 0; i < sizes.size(); ++i) {
    EXPECT_EQ(b2.size(i), sizes[i]);
    accessKernel<<<1, 1>>>(b2, i, sizes[i] - 1, g_res);
    kernelCheck();
    EXPECT_EQ(*g_res, size_t(1));
  }
  std::printf("bit_array: ok so far (%d failures)\n", g_fail);
}

static void test_rank_select() {
  const size_t L1 = RSConfig::L1_BIT_SIZE, L2 = RSConfig::L2_BIT_SIZE;
  {  // index sizes (test_rank_select.cu:191-253)
    BitArray ba(std::vector<size_t>{1, L2 - 1, L2 + 1, 2 * L2 - 1, L1 - 1, L1 + 1, 2 * L1 - 1},
                false);
    RankSelect rs(std::move(ba), 0);
    size_t l1[] = {1, 1, 1, 1, 1, 2, 2}, l2[] = {1, 1, 2, 2, 128, 129, 256},
           last[] = {1, 1, 2, 2, 128, 1, 128};
    for (uint32_t i = 0; i < 7; ++i) {
      numL1Kernel<<<1, 1>>>(rs, i, g_res); kernelCheck(); EXPECT_EQ(*g_res, l1[i]);
      numL2Kernel<<<1, 1>>>(rs, i, g_res); kernelCheck(); EXPECT_EQ(*g_res, l2[i]);
      lastL2Kernel<<<1, 1>>>(rs, i, g_res); kernelCheck(); EXPECT_EQ(*g_res, last[i]);
    }
  }
  for (bool val : {true, false}) {  // operations on all-ones / all-zeros
    std::vector<size_t> sizes{L2 + 1, 2 * L2 - 1, L1 - 1, L1 + 1, 2 * L1 - 1, 2 * L1 + 1, 3 * L1};
    BitArray ba(sizes, val);
    RankSelect rs(std::move(ba), 0);
    std::mt19937 gen(7);
    for (uint32_t i = 0; i < sizes.size(); ++i) {
      size_t m = rs.bit_array_.size(i);
      EXPECT_EQ(m, sizes[i]);
      std::uniform_int_distribution<size_t> dis(1, m - 1);
      for (int nt : kGroupSizes) {
        for (int q = 0; q < 10; ++q) {
          size_t pos = dis(gen);
          EXPECT_EQ(run_rank(rs, nt, val, i, pos), pos);
          EXPECT_EQ(run_select(rs, nt, val, i, pos), pos - 1);
        }
        EXPECT_EQ(run_rank(rs, nt, val, i, 0), size_t(0));
        EXPECT_EQ(run_rank(rs, nt, val, i, m), m);
        EXPECT_EQ(run_select(rs, nt, val, i, 1), size_t(0));
        EXPECT_EQ(run_select(rs, nt, val, i, m), m - 1);
        EXPECT_EQ(run_rank(rs, nt, !val, i, m), size_t(0));
        EXPECT_EQ(run_select(rs, nt, !val, i, 1), m);
      }
      totalKernel<<<1, 1>>>(rs, i, val, g_res); kernelCheck(); EXPECT_EQ(*g_res, m);
      for (size_t j = 1; j <= m / RSConfig::SELECT_SAMPLE_RATE; ++j) {
        sampleKernel<<<1, 1>>>(rs, i, j - 1, val, g_res);
        kernelCheck();
        EXPECT_EQ(*g_res, j * RSConfig::SELECT_SAMPLE_RATE - 1);
      }
      size_t nl1 = (m + L1 - 1) / L1;
      for (size_t j = 0; j < nl1; ++j) {
        l1Kernel<<<1, 1>>>(rs, i, j, g_res); kernelCheck(); EXPECT_EQ(*g_res, val ? j * L1 : 0);
      }
      if ((m + L2 - 1) / L2 > 3) {  // L2 entry 3 exists
        l2Kernel<<<1, 1>>>(rs, i, 3, g_res); kernelCheck(); EXPECT_EQ(*g_res, val ? 3 * L2 : 0);
      }
    }
  }
  // random words vs prefix counts (RankSelectOperationsRandom), at density
  // 1/2 and 1/16 (select brackets of ~16 and ~128 blocks), every group size,
  // each query answered by its own group among many in a launch, on the
  // directory, then with the query lines, then with the select lines too
  for (int sparse = 0; sparse < 2; ++sparse) {
    std::mt19937 gen(11 + sparse);
    const size_t m = sparse ? 2000003 : 300001;
    std::vector<uint32_t> words((m + 31) / 32);
    for (auto& w : words) w = sparse ? gen() & gen() & gen() & gen() : gen();
    BitArray ba(std::vector<size_t>{m}, false);
    ba.writeWordsFromHost(0, 0, words.data(), words.size());
    RankSelect rs(std::move(ba), 0);
    std::vector<size_t> pref(m + 1, 0);
    for (size_t i = 0; i < m; ++i) pref[i + 1] = pref[i] + ((words[i / 32] >> (i % 32)) & 1);
    const size_t ones = pref[m], zeros = m - ones;
    // rank positions (ends included) and select ranks of both Values
    const size_t nq = 3000;
    std::vector<size_t> rpos(nq), sel1(nq), sel0(nq);
    std::uniform_int_distribution<size_t> dis(0, m);
    for (size_t j = 0; j < nq; ++j) {
      rpos[j] = j == 0 ? 0 : j == 1 ? m : j == 2 ? m - 1 : dis(gen);
      sel1[j] = j == 0 ? 1 : j == 1 ? ones : j == 2 ? ones + 1 : 1 + dis(gen) % ones;
      sel0[j] = j == 0 ? 1 : j == 1 ? zeros : j == 2 ? zeros + 1 : 1 + dis(gen) % zeros;
    }
    // expected answers: rank by prefix counts, select by a sweep over the bits
    std::vector<size_t> pos1(ones), pos0(zeros);
    for (size_t i = 0, a1 = 0, a0 = 0; i < m; ++i) {
      if ((words[i / 32] >> (i % 32)) & 1) pos1[a1++] = i;
      else pos0[a0++] = i;
    }
    auto sel_exp = [&](const std::vector<size_t>& pos, size_t k) {
      return k >= 1 && k <= pos.size() ? pos[k - 1] : m;
    };
    size_t *d_idx, *d_out;
    gpuErrchk(hipMalloc(&d_idx, nq * sizeof(size_t)));
    gpuErrchk(hipMalloc(&d_out, nq * 64 * sizeof(size_t)));
    std::vector<size_t> out(nq * 64);
    for (int layout = 0; layout < 3; ++layout) {
      if (layout == 1) utils::throwOnError(ecl_rank_select_build_query_lines(rs.handle()));
      if (layout == 2) utils::throwOnError(ecl_rank_select_build_select_lines(rs.handle()));
      for (int nt : kGroupSizes) {
        for (int op = 0; op < 4; ++op) {  // rank1, rank0, select1, select0
          const bool sel = op >= 2;
          const int v = op == 0 || op == 2;
          const std::vector<size_t>& in = !sel ? rpos : v ? sel1 : sel0;
          gpuErrchk(hipMemcpy(d_idx, in.data(), nq * sizeof(size_t), hipMemcpyHostToDevice));
          run_groups(sel, rs, nt, v, d_idx, nq, d_out);
          gpuErrchk(hipMemcpy(out.data(), d_out, nq * nt * sizeof(size_t),
                              hipMemcpyDeviceToHost));
          int bad = 0;
          for (size_t j = 0; j < nq; ++j) {
            const size_t e = !sel ? (v ? pref[in[j]] : in[j] - pref[in[j]])
                                  : sel_exp(v ? pos1 : pos0, in[j]);
            for (int t = 0; t < nt; ++t) bad += out[j * nt + t] != e;
          }
          if (bad)
            std::printf("rank/select groups: density 1/%d layout %d nt %d op %d: %d wrong\n",
                        sparse ? 16 : 2, layout, nt, op, bad);
          EXPECT_EQ(bad, 0);
        }
      }
      // single-group kernels agree on the same arrays
      for (int nt : {1, 4, 64}) {
        for (int q = 0; q < 20; ++q) {
          const size_t i = rpos[q], k = sel1[q], z = sel0[q];
          EXPECT_EQ(run_rank(rs, nt, 1, 0, i), pref[i]);
          EXPECT_EQ(run_rank(rs, nt, 0, 0, i), i - pref[i]);
          EXPECT_EQ(run_select(rs, nt, 1, 0, k), sel_exp(pos1, k));
          EXPECT_EQ(run_select(rs, nt, 0, 0, z), sel_exp(pos0, z));
        }
      }
    }
    gpuErrchk(hipFree(d_idx));
    gpuErrchk(hipFree(d_out));
  }
  std::printf("rank_select: ok so far (%d failures)\n", g_fail);
}

template <typename T, int TPQ>
static void launch_user(WaveletTree<T>& wt, int op, const size_t* d_idx,
                        const RankSelectQuery<T>* d_q, size_t m, size_t* d_out) {
  if (op == 0) userAccess<T, TPQ><<<8, 64>>>(wt, d_idx, m, d_out);
  else if (op == 1) userRank<T, TPQ><<<8, 64>>>(wt, d_q, m, d_out);
  else userSelect<T, TPQ><<<8, 64>>>(wt, d_q, m, d_out);
}
template <typename T>
static void run_user(WaveletTree<T>& wt, int tpq, int op, const size_t* d_idx,
                     const RankSelectQuery<T>* d_q, size_t m, size_t* d_out) {
  switch (tpq) {
    case 1: launch_user<T, 1>(wt, op, d_idx, d_q, m, d_out); break;
    case 2: launch_user<T, 2>(wt, op, d_idx, d_q, m, d_out); break;
    case 4: launch_user<T, 4>(wt, op, d_idx, d_q, m, d_out); break;
    case 8: launch_user<T, 8>(wt, op, d_idx, d_q, m, d_out); break;
    case 16: launch_user<T, 16>(wt, op, d_idx, d_q, m, d_out); break;
    case 32: launch_user<T, 32>(wt, op, d_idx, d_q, m, d_out); break;
    default: launch_user<T, 64>(wt, op, d_idx, d_q, m, d_out); break;
  }
  kernelCheck();
}

template <typename T>
static void test_tree_type() {
  {  // structure sigma=5 (test_wavelet_tree.cu:520-565)
    std::vector<T> data{4, 1, 1, 2, 1, 1, 3, 4, 4, 0, 0, 2, 3, 2, 4, 2, 4, 1, 3, 2};
    WaveletTree<T> wt(data.data(), data.size(), std::vector<T>{0, 1, 2, 3, 4}, 0);
    const char* exp[3] = {"10000001100000101000", "001001001111011", "111100101010010"};
    for (int l = 0; l < 3; ++l) {
      for (size_t i = 0; exp[l][i]; ++i) {
        baAccess<T><<<1, 1>>>(wt.rank_select_.bit_array_, l, i, g_res);
        kernelCheck();
        EXPECT_EQ(*g_res, size_t(exp[l][i] - '0'));
      }
    }
  }
  {  // access / rank / select on i % 10 (test_wavelet_tree.cu:567-709)
    std::vector<T> data(100);
    for (size_t i = 0; i < 100; ++i) data[i] = i % 10;
    std::vector<T> alphabet(10);
    std::iota(alphabet.begin(), alphabet.end(), 0);
    WaveletTree<T> wt(data.data(), data.size(), std::move(alphabet), 0);
    std::vector<size_t> idx(100);
    std::iota(idx.begin(), idx.end(), 0);
    auto res = wt.access(idx.data(), idx.size());
    for (size_t i = 0; i < 100; ++i) EXPECT_EQ(size_t(res[i]), size_t(data[i]));
    std::vector<RankSelectQuery<T>> q;
    for (size_t j = 0; j < 100; ++j) q.push_back({j, data[j]});
    auto rr = wt.rank(q.data(), q.size());
    for (size_t j = 0; j < 100; ++j) EXPECT_EQ(rr[j], j / 10);
    q.clear();
    for (size_t j = 0; j < 100; ++j) q.push_back({j / 10 + 1, data[j]});
    auto sr = wt.select(q.data(), q.size());
    for (size_t j = 0; j < 100; ++j) EXPECT_EQ(sr[j], j);
    std::vector<RankSelectQuery<T>> q11{{11, 0}};
    EXPECT_EQ(wt.select(q11.data(), 1)[0], size_t(100));
    EXPECT_EQ(wt.getTotalAppearances(3), size_t(10));
  }
  {  // random sorted non-minimal alphabet, device API from user kernels
    std::mt19937_64 gen(sizeof(T) * 13);
    std::vector<T> alphabet;
    while (alphabet.size() < 37) {
      T v = static_cast<T>(gen());
      if (std::find(alphabet.begin(), alphabet.end(), v) == alphabet.end()) alphabet.push_back(v);
    }
    std::sort(alphabet.begin(), alphabet.end());
    const size_t n = 200000;
    std::vector<T> data(n);
    for (auto& x : data) x = alphabet[gen() % alphabet.size()];
    auto a_copy = alphabet;
    WaveletTree<T> wt(data.data(), n, std::move(a_copy), 0);
    EXPECT_EQ(wt.isMinAlphabet(), false);
    EXPECT_EQ(wt.getAlphabetSize(), alphabet.size());
    std::vector<size_t> idx(4000);
    for (auto& i : idx) i = gen() % n;
    auto res = wt.access(idx.data(), idx.size());
    for (size_t i = 0; i < idx.size(); ++i) EXPECT_EQ(size_t(res[i]), size_t(data[idx[i]]));
    // host rank rewrites symbols to minimal indices (wavelet_tree.cuh:1426-1436)
    std::vector<RankSelectQuery<T>> q;
    for (int j = 0; j < 2000; ++j) q.push_back({gen() % n, alphabet[gen() % alphabet.size()]});
    std::vector<size_t> exp;
    for (auto& x : q) exp.push_back(std::count(data.begin(), data.begin() + x.index_, x.symbol_));
    auto rr = wt.rank(q.data(), q.size());
    for (size_t j = 0; j < q.size(); ++j) EXPECT_EQ(rr[j], exp[j]);
    for (auto& x : q) EXPECT_EQ(x.symbol_ < alphabet.size(), true);
    // device API from user kernels: minimal-alphabet symbols and indices,
    // every ThreadsPerQuery, on the directory, then with the query lines,
    // then with the select lines too
    std::vector<size_t> aexp(idx.size());
    for (size_t i = 0; i < idx.size(); ++i)
      aexp[i] = std::lower_bound(alphabet.begin(), alphabet.end(), data[idx[i]]) - alphabet.begin();
    // select the k-th occurrence, plus one past the last of a symbol
    std::vector<RankSelectQuery<T>> sq;
    std::vector<size_t> sexp;
    for (int j = 0; j < 1000; ++j) {
      size_t p = gen() % n;
      T c = data[p];
      size_t k = std::count(data.begin(), data.begin() + p + 1, c);
      sq.push_back({k, static_cast<T>(std::lower_bound(alphabet.begin(), alphabet.end(), c) - alphabet.begin())});
      sexp.push_back(p);
    }
    sq.push_back({size_t(std::count(data.begin(), data.end(), alphabet[3])) + 1, T(3)});
    sexp.push_back(n);
    // rank at both ends
    q.push_back({0, T(5)});
    exp.push_back(0);
    q.push_back({n, T(5)});
    exp.push_back(std::count(data.begin(), data.end(), alphabet[5]));
    const size_t mq = std::max({idx.size(), q.size(), sq.size()});
    size_t *d_idx, *d_out;
    RankSelectQuery<T>* d_q;
    gpuErrchk(hipMalloc(&d_idx, idx.size() * 8));
    gpuErrchk(hipMalloc(&d_out, mq * 64 * 8));
    gpuErrchk(hipMalloc(&d_q, mq * sizeof(RankSelectQuery<T>)));
    gpuErrchk(hipMemcpy(d_idx, idx.data(), idx.size() * 8, hipMemcpyHostToDevice));
    std::vector<size_t> out(mq * 64);
    for (int layout = 0; layout < 3; ++layout) {
      if (layout == 1) utils::throwOnError(ecl_wavelet_tree_build_query_lines(wt.handle()));
      if (layout == 2) utils::throwOnError(ecl_wavelet_tree_build_select_lines(wt.handle()));
      for (int tpq : kGroupSizes) {
        for (int op = 0; op < 3; ++op) {  // access, rank, select
          const size_t cnt = op == 0 ? idx.size() : op == 1 ? q.size() : sq.size();
          if (op == 1)
            gpuErrchk(hipMemcpy(d_q, q.data(), cnt * sizeof(RankSelectQuery<T>), hipMemcpyHostToDevice));
          if (op == 2)
            gpuErrchk(hipMemcpy(d_q, sq.data(), cnt * sizeof(RankSelectQuery<T>), hipMemcpyHostToDevice));
          run_user<T>(wt, tpq, op, d_idx, d_q, cnt, d_out);
          gpuErrchk(hipMemcpy(out.data(), d_out, cnt * tpq * 8, hipMemcpyDeviceToHost));
          const std::vector<size_t>& e = op == 0 ? aexp : op == 1 ? exp : sexp;
          int bad = 0;
          for (size_t j = 0; j < cnt; ++j)
            for (int t = 0; t < tpq; ++t) bad += out[j * tpq + t] != e[j];
          if (bad)
            std::printf("device API <%zu>: layout %d TPQ %d op %d: %d wrong\n", sizeof(T),
                        layout, tpq, op, bad);
          EXPECT_EQ(bad, 0);
        }
      }
    }
    gpuErrchk(hipFree(d_idx));
    gpuErrchk(hipFree(d_out));
    gpuErrchk(hipFree(d_q));
  }
  std::printf("wavelet_tree<%zu>: ok so far (%d failures)\n", sizeof(T), g_fail);
}

// The device API over the quad and hex layouts: power-of-two alphabets (16,
// 64 and 256 symbols; not minimal except 256 of uint8), the same user
// kernels for every ThreadsPerQuery over the binary levels, then with the
// quad lines (access and rank, ThreadsPerQuery >= 2), the quad select lines
// (select) and the hex lines (access and rank, 16 and 256 symbols), built
// through the C ABI after the tree was passed to earlier launches.
template <typename T>
static void test_device_api_layouts() {
  for (size_t sigma : {size_t(16), size_t(64), size_t(256)}) {
    std::mt19937_64 gen(sizeof(T) * 101 + sigma);
    std::vector<T> alphabet;
    if (sizeof(T) == 1 && sigma == 256) {
      for (size_t c = 0; c < 256; ++c) alphabet.push_back(static_cast<T>(c));
    } else {
      while (alphabet.size() < sigma) {
        T v = static_cast<T>(gen());
        if (std::find(alphabet.begin(), alphabet.end(), v) == alphabet.end()) alphabet.push_back(v);
      }
      std::sort(alphabet.begin(), alphabet.end());
    }
    const size_t n = 150001;
    std::vector<T> data(n);
    for (auto& x : data) {  // skewed: symbol j with weight ~ 1 / (j + 1)
      const size_t r = gen() % (sigma * 4);
      x = alphabet[r < sigma ? r : gen() % (1 + r % sigma)];
    }
    auto a_copy = alphabet;
    WaveletTree<T> wt(data.data(), n, std::move(a_copy), 0);
    auto minimal = [&](T c) { return size_t(std::lower_bound(alphabet.begin(), alphabet.end(), c) - alphabet.begin()); };
    std::vector<size_t> idx(3000), aexp(idx.size());
    for (size_t i = 0; i < idx.size(); ++i) {
      idx[i] = i < 2 ? (i ? n - 1 : 0) : gen() % n;
      aexp[i] = minimal(data[idx[i]]);
    }
    std::vector<RankSelectQuery<T>> q, sq;
    std::vector<size_t> exp, sexp;
    for (int j = 0; j < 2000; ++j) {
      const size_t i = j == 0 ? n : gen() % (n + 1);
      const size_t c = gen() % sigma;
      q.push_back({i, static_cast<T>(c)});
      exp.push_back(std::count(data.begin(), data.begin() + i, alphabet[c]));
    }
    for (int j = 0; j < 1500; ++j) {
      const size_t p = gen() % n;
      const size_t k = std::count(data.begin(), data.begin() + p + 1, data[p]);
      sq.push_back({k, static_cast<T>(minimal(data[p]))});
      sexp.push_back(p);
    }
    for (size_t c : {size_t(0), sigma - 1}) {  // one past the last occurrence
      sq.push_back({size_t(std::count(data.begin(), data.end(), alphabet[c])) + 1, static_cast<T>(c)});
      sexp.push_back(n);
    }
    const size_t mq = std::max({idx.size(), q.size(), sq.size()});
    size_t *d_idx, *d_out;
    RankSelectQuery<T>* d_q;
    gpuErrchk(hipMalloc(&d_idx, idx.size() * 8));
    gpuErrchk(hipMalloc(&d_out, mq * 64 * 8));
    gpuErrchk(hipMalloc(&d_q, mq * sizeof(RankSelectQuery<T>)));
    gpuErrchk(hipMemcpy(d_idx, idx.data(), idx.size() * 8, hipMemcpyHostToDevice));
    std::vector<size_t> out(mq * 64);
    ecl_wavelet_tree_info info;
    utils::throwOnError(ecl_wavelet_tree_info_get(wt.handle(), &info));
    const uint32_t L = info.num_levels;
    for (int layout = 0; layout < 4; ++layout) {
      if (layout == 1) utils::throwOnError(ecl_wavelet_tree_build_quad_lines(wt.handle()));
      if (layout == 2) utils::throwOnError(ecl_wavelet_tree_build_quad_select_lines(wt.handle()));
      if (layout == 3) {
        if (L % 4 != 0) break;
        utils::throwOnError(ecl_wavelet_tree_build_hex_lines(wt.handle()));
      }
      for (int tpq : kGroupSizes) {
        for (int op = 0; op < 3; ++op) {
          const size_t cnt = op == 0 ? idx.size() : op == 1 ? q.size() : sq.size();
          if (op == 1)
            gpuErrchk(hipMemcpy(d_q, q.data(), cnt * sizeof(RankSelectQuery<T>), hipMemcpyHostToDevice));
          if (op == 2)
            gpuErrchk(hipMemcpy(d_q, sq.data(), cnt * sizeof(RankSelectQuery<T>), hipMemcpyHostToDevice));
          run_user<T>(wt, tpq, op, d_idx, d_q, cnt, d_out);
          gpuErrchk(hipMemcpy(out.data(), d_out, cnt * tpq * 8, hipMemcpyDeviceToHost));
          const std::vector<size_t>& e = op == 0 ? aexp : op == 1 ? exp : sexp;
          int bad = 0;
          for (size_t j = 0; j < cnt; ++j)
            for (int t = 0; t < tpq; ++t) bad += out[j * tpq + t] != e[j];
          if (bad)
            std::printf("device API layouts <%zu> sigma %zu: layout %d TPQ %d op %d: %d wrong\n",
                        sizeof(T), sigma, layout, tpq, op, bad);
          EXPECT_EQ(bad, 0);
        }
      }
    }
    gpuErrchk(hipFree(d_idx));
    gpuErrchk(hipFree(d_out));
    gpuErrchk(hipFree(d_q));
  }
  std::printf("device API layouts <%zu>: ok so far (%d failures)\n", sizeof(T), g_fail);
}

static void test_codes_and_nodes() {
  // createMinimalCodes KATs (test_wavelet_tree.cu:174-217)
  std::vector<uint8_t> a(72);
  std::iota(a.begin(), a.end(), 0);
  auto codes = WaveletTree<uint8_t>::createMinimalCodes(a);
  EXPECT_EQ(codes.size(), size_t(8));
  for (int i = 64; i < 72; ++i) {
    EXPECT_EQ(size_t(codes[i - 64].code_), size_t(64 + 8 * (i - 64)));
    EXPECT_EQ(size_t(codes[i - 64].len_), size_t(4));
  }
  a.resize(63);
  codes = WaveletTree<uint8_t>::createMinimalCodes(a);
  EXPECT_EQ(size_t(codes[0].code_), size_t(62));
  EXPECT_EQ(size_t(codes[0].len_), size_t(5));
  a.resize(75);
  std::iota(a.begin(), a.end(), 0);
  codes = WaveletTree<uint8_t>::createMinimalCodes(a);
  for (int i = 64; i < 72; ++i) {
    EXPECT_EQ(size_t(codes[i - 64].code_), size_t(64 + 4 * (i - 64)));
    EXPECT_EQ(size_t(codes[i - 64].len_), size_t(5));
  }
  EXPECT_EQ(size_t(codes[8].code_), size_t(96));
  EXPECT_EQ(size_t(codes[9].code_), size_t(104));
  EXPECT_EQ(size_t(codes[10].code_), size_t(112));
  EXPECT_EQ(size_t(codes[10].len_), size_t(3));
  a.resize(64);
  EXPECT_EQ(WaveletTree<uint8_t>::createMinimalCodes(a).size(), size_t(0));
  // node positions (getNodePosRandom): starts of a sigma=200 tree are
  // numbered 0, 1, 2, ... per level
  std::vector<uint16_t> alpha(200);
  std::iota(alpha.begin(), alpha.end(), 0);
  auto a2 = alpha;
  WaveletTree<uint16_t> wt(alpha.data(), alpha.size(), std::move(a2), 0);
  std::vector<uint64_t> ranks(wt.handle() ? 512 : 0), starts(512);
  std::vector<uint8_t> lv(512);
  ecl_wavelet_tree_info info;
  utils::throwOnError(ecl_wavelet_tree_info_get(wt.handle(), &info));
  utils::throwOnError(ecl_wavelet_tree_node_ranks(wt.handle(), ranks.data(), starts.data(), lv.data()));
  size_t counter = 0;
  for (size_t j = 0; j < info.num_ranks; ++j) {
    if (j > 0 && lv[j] != lv[j - 1]) counter = 0;
    nodePosKernel<uint16_t, false><<<1, 1>>>(wt, static_cast<uint16_t>(starts[j]), lv[j], g_res);
    kernelCheck();
    EXPECT_EQ(*g_res, counter++);
  }
  std::printf("codes/nodes: ok so far (%d failures)\n", g_fail);
}

static void test_oom() {
  // WaveletTreeConstructor (test_wavelet_tree.cu:97-172): too much data
  // raises the reference's exact message
  size_t fr = 0, tot = 0;
  gpuErrchk(hipMemGetInfo(&fr, &tot));
  const size_t n = tot * 4;  // never fits; the text is not read before admission
  std::string msg;
  try {
    uint8_t dummy = 0;
    WaveletTree<uint8_t> wt(&dummy, n, std::vector<uint8_t>{0, 1, 2, 3}, 0);
  } catch (std::runtime_error const& e) {
    msg = e.what();
  }
  EXPECT_EQ(msg == "Not enough memory available for the wavelet tree.", true);
  std::printf("oom: '%s'\n", msg.c_str());
}

static void example_dna() {
  // example/example.cu: DNA over {0,1,2,3}
  std::vector<uint8_t> dna(1000);
  std::mt19937 gen(1);
  for (auto& x : dna) x = gen() % 4;
  WaveletTree<uint8_t> wt(dna.data(), dna.size(), std::vector<uint8_t>{0, 1, 2, 3}, 0);
  std::vector<size_t> idx(dna.size());
  std::iota(idx.begin(), idx.end(), 0);
  auto acc = wt.access(idx.data(), idx.size());
  for (size_t i = 0; i < dna.size(); ++i) EXPECT_EQ(acc[i], dna[i]);
  std::vector<RankSelectQuery<uint8_t>> rq{{500, 0}, {500, 1}, {500, 2}, {500, 3}};
  auto rr = wt.rank(rq.data(), rq.size());
  for (int c = 0; c < 4; ++c)
    EXPECT_EQ(rr[c], size_t(std::count(dna.begin(), dna.begin() + 500, c)));
  std::vector<RankSelectQuery<uint8_t>> sq;
  for (uint8_t c = 0; c < 4; ++c) sq.push_back({1, c});
  for (uint8_t c = 0; c < 4; ++c) sq.push_back({wt.getTotalAppearances(c), c});
  auto sr = wt.select(sq.data(), sq.size());
  for (int c = 0; c < 4; ++c) {
    EXPECT_EQ(sr[c], size_t(std::find(dna.begin(), dna.end(), c) - dna.begin()));
    EXPECT_EQ(sr[4 + c], size_t(dna.rend() - std::find(dna.rbegin(), dna.rend(), c) - 1));
  }
  std::printf("example: ok so far (%d failures)\n", g_fail);
}

static void test_save_load() {
  // persistence: the loaded tree answers like the saved one
  std::vector<uint16_t> text(200000);
  std::mt19937 gen(7);
  for (auto& x : text) x = static_cast<uint16_t>(gen() % 1000);
  std::vector<uint16_t> alpha(1000);
  std::iota(alpha.begin(), alpha.end(), 0);
  WaveletTree<uint16_t> wt(text.data(), text.size(), std::move(alpha), 0);
  char path[] = "/tmp/ecl_dropin_treeXXXXXX";
  const int fd = mkstemp(path);
  EXPECT_EQ(fd >= 0, true);
  if (fd >= 0) close(fd);
  wt.save(path);
  WaveletTree<uint16_t> lt(WaveletTree<uint16_t>::FromFile{}, path, 0);
  std::remove(path);
  std::vector<RankSelectQuery<uint16_t>> q;
  for (size_t i = 0; i < 3000; ++i)
    q.push_back({static_cast<size_t>(gen() % text.size()), static_cast<uint16_t>(gen() % 1000)});
  auto q2 = q;
  auto a = wt.rank(q.data(), q.size());
  std::vector<size_t> ra(a.begin(), a.end());
  auto b = lt.rank(q2.data(), q2.size());
  for (size_t i = 0; i < q.size(); ++i) EXPECT_EQ(ra[i], b[i]);
  std::vector<size_t> idx(2000);
  for (auto& i : idx) i = gen() % text.size();
  auto acc = lt.access(idx.data(), idx.size());
  for (size_t i = 0; i < idx.size(); ++i) EXPECT_EQ(acc[i], text[idx[i]]);
  std::string msg;
  try {
    WaveletTree<uint8_t> bad(WaveletTree<uint8_t>::FromFile{}, "/nonexistent/tree", 0);
  } catch (std::runtime_error const& e) {
    msg = e.what();
  }
  EXPECT_EQ(msg.empty(), false);
  std::printf("save/load: ok so far (%d failures)\n", g_fail);
}

// getPrevPowTwo (largest power of two strictly below i, 0 for 0 and 1) and
// ceilLog2 (ceilLog2(1) = 0) for every i < min(1024, max(T)), in a kernel
// and, for ceilLog2, on the host, as the reference's UtilsTest does.
template <typename T>
static void test_utils_type() {
  const uint32_t limit =
      static_cast<uint32_t>(std::min<uint64_t>(1024, std::numeric_limits<T>::max()));
  uint64_t *prev = nullptr, *clog = nullptr;
  gpuErrchk(hipMallocManaged(&prev, 1024 * sizeof(uint64_t)));
  gpuErrchk(hipMallocManaged(&clog, 1024 * sizeof(uint64_t)));
  utilsKernel<T><<<1, 256>>>(limit, prev, clog);
  kernelCheck();
  uint64_t expected = 0;
  for (uint32_t i = 0; i < limit; ++i) {
    EXPECT_EQ(prev[i], expected);
    if (utils::isPowTwo(static_cast<T>(i))) expected = i;
  }
  expected = 0;
  for (uint32_t i = 1; i < limit; ++i) {
    EXPECT_EQ(static_cast<uint64_t>(utils::ceilLog2<T>(static_cast<T>(i))), expected);
    EXPECT_EQ(clog[i], expected);
    if (utils::isPowTwo(static_cast<T>(i))) ++expected;
  }
  gpuErrchk(hipFree(prev));
  gpuErrchk(hipFree(clog));
  std::printf("utils<%zu>: ok so far (%d failures)\n", sizeof(T), g_fail);
}

// readDataFromFile + convertDataToMinAlphabet (test_benchmark_utils.cuh:
// 588-679) feeding a tree: a raw u16 file over a sparse alphabet, a prefix
// of it converted to the minimal alphabet and built with the alphabet
// 0 .. sigma-1 answers like the direct scan of the converted text.
static void test_text_file() {
  char path[] = "/tmp/ecl_dropin_textXXXXXX";
  const int fd = mkstemp(path);
  EXPECT_EQ(fd >= 0, true);
  std::mt19937_64 gen(5);
  std::vector<uint16_t> raw(200000);
  for (auto& x : raw) x = static_cast<uint16_t>(1000 + 7 * (gen() % 50));
  EXPECT_EQ(write(fd, raw.data(), raw.size() * 2), static_cast<ssize_t>(raw.size() * 2));
  close(fd);
  auto data = utils::readDataFromFile<uint16_t>(path, 150000);
  EXPECT_EQ(data.size(), size_t(150000));
  EXPECT_EQ(data[149999], raw[149999]);
  std::string msg;
  try {
    (void)utils::readDataFromFile<uint16_t>(path, 200001);
  } catch (std::runtime_error const& e) {
    msg = e.what();
  }
  EXPECT_EQ(msg == "Data size is larger than file size", true);
  const size_t sigma = utils::convertDataToMinAlphabet(data.data(), data.size());
  EXPECT_EQ(sigma, size_t(50));
  std::vector<uint16_t> alphabet(sigma);
  std::iota(alphabet.begin(), alphabet.end(), 0);
  WaveletTree<uint16_t> wt(data.data(), data.size(), std::move(alphabet), 0);
  std::vector<RankSelectQuery<uint16_t>> q(500);
  for (auto& x : q) x = {gen() % data.size(), static_cast<uint16_t>(gen() % sigma)};
  std::vector<size_t> exp(q.size());
  for (size_t i = 0; i < q.size(); ++i)
    exp[i] = std::count(data.begin(), data.begin() + q[i].index_, q[i].symbol_);
  auto r = wt.rank(q.data(), q.size());
  for (size_t i = 0; i < q.size(); ++i) EXPECT_EQ(r[i], exp[i]);
  unlink(path);
  std::printf("text file: ok so far (%d failures)\n", g_fail);
}

int main() {
  gpuErrchk(hipMallocManaged(&g_res, 64 * sizeof(size_t)));
  try {
    test_utils_type<uint8_t>();
    test_utils_type<uint16_t>();
    test_utils_type<uint32_t>();
    test_utils_type<uint64_t>();
    test_bit_array();
    test_rank_select();
    test_tree_type<uint8_t>();
    test_tree_type<uint16_t>();
    test_tree_type<uint32_t>();
    test_tree_type<uint64_t>();
    test_device_api_layouts<uint8_t>();
    test_device_api_layouts<uint16_t>();
    test_device_api_layouts<uint32_t>();
    test_device_api_layouts<uint64_t>();
    test_codes_and_nodes();
    test_oom();
    example_dna();
    test_save_load();
    test_text_file();
  } catch (std::exception const& e) {
    std::printf("EXCEPTION: %s\n", e.what());
    return 2;
  }
  gpuErrchk(hipFree(g_res));
  std::printf("%d checks, %d failures\n", g_checks, g_fail);
  if (g_fail == 0) std::printf("ALL OK\n");
  return g_fail == 0 ? 0 : 1;
}
