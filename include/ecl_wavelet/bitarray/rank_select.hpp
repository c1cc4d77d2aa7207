// ecl_wavelet/bitarray/rank_select.hpp -- drop-in for bitarray/
// rank_select.cuh (mfranzreb/ECL_wavelet) on HIP / gfx950.
//
// The directory has the reference's layout and values (RSConfig below); its
// select samples of one array share a buffer of floor(size/4096) slots, ones
// from the front and zeros from the back, which getSelectSample hides.
// rank/select take the reference's template arguments; NumThreads may be
// any power of two up to the wave width (64): a group is NumThreads
// consecutive lanes of one wave, all active and calling with the same
// arguments; each lane reads its share of the block (or line), the group
// combines the counts over shuffles, and every lane returns the answer.
#pragma once

#include "ecl_wavelet/bitarray/bit_array.hpp"

namespace ecl {

// RSConfig (rank_select.cuh:44-64)
struct RSConfig {
  static constexpr size_t L2_BIT_SIZE = 512;
  static constexpr size_t NUM_L2_PER_L1 = 128;
  static constexpr size_t L1_BIT_SIZE = NUM_L2_PER_L1 * L2_BIT_SIZE;
  static constexpr size_t L2_WORD_SIZE = L2_BIT_SIZE / 32;
  static constexpr size_t L1_WORD_SIZE = L1_BIT_SIZE / 32;
  static constexpr size_t SELECT_SAMPLE_RATE = 4096;
  using L1_TYPE = uint64_t;
  using L2_TYPE = uint16_t;
};

// RankResult (rank_select.cuh:71-74)
struct RankResult {
  size_t rank;
  bool bit;
};

template <typename T> class WaveletTree;

class RankSelect {
 public:
  BitArray bit_array_;

  RankSelect() noexcept = default;

  // RankSelect(BitArray&&, GPU_index) (rank_select.cu:50-304)
  __host__ RankSelect(BitArray&& bit_array, uint32_t const GPU_index) {
    ecl_bit_array h = bit_array.handle_;
    utils::throwOnError(ecl_rank_select_create(&h, static_cast<int>(GPU_index), &handle_));
    bit_array.is_copy_ = true;  // the library object now belongs to the RankSelect
    is_copy_ = false;
    bind();
  }
  // copy = shallow view (rank_select.cu:306-321)
  __host__ RankSelect(RankSelect const& o) noexcept
      : bit_array_(o.bit_array_), handle_(o.handle_), d_(o.d_), levels_(o.levels_),
        is_copy_(true) {}
  __host__ RankSelect& operator=(RankSelect&& o) noexcept {
    release();
    bit_array_ = std::move(o.bit_array_);
    handle_ = o.handle_; d_ = o.d_; levels_ = o.levels_; is_copy_ = o.is_copy_;
    o.is_copy_ = true;
    return *this;
  }
  __host__ ~RankSelect() { release(); }

  // rank<NumThreads, GetBit, Value> (rank_select.cuh:125-208): number of
  // Value bits in [0, index) of array `array_index`; index in [0, size].
  template <int NumThreads, bool GetBit, int Value>
  __device__ auto rank(uint8_t const array_index, size_t const index, size_t const /*offset*/) {
    static_assert(Value == 0 || Value == 1, "Value must be 0 or 1.");
    static_assert(NumThreads > 0 && NumThreads <= WS && (WS % NumThreads) == 0,
                  "NumThreads must divide the wave width");
    bool bit = false;
    const size_t r = detail::level_rank_coop<Value, NumThreads>(levels_[array_index], index,
                                                                GetBit ? &bit : nullptr);
    if constexpr (GetBit) {
      return RankResult{r, bit};
    } else {
      return r;
    }
  }

  // select<Value, NumThreads, UseSamples> (rank_select.cuh:224-523):
  // position of the i-th (1-based) Value, size of the array when absent.
  template <uint32_t Value, int NumThreads, bool UseSamples = true>
  __device__ size_t select(uint32_t const array_index, size_t i, size_t const /*BA_offset*/) {
    static_assert(Value == 0 || Value == 1, "Value must be 0 or 1.");
    static_assert(NumThreads > 0 && NumThreads <= WS && (WS % NumThreads) == 0,
                  "NumThreads must divide the wave width");
    return detail::level_select_coop<Value, NumThreads>(levels_[array_index], i);
  }

  // ---- directory accessors (rank_select.cuh:530-701)
  __device__ size_t getNumL1Blocks(uint32_t const a) const { return d_.num_l1_blocks[a]; }
  __device__ size_t getNumL2Blocks(uint32_t const a) const {
    if (a == bit_array_.numArrays() - 1u) return d_.total_num_l2_blocks - d_.l2_offsets[a];
    return d_.l2_offsets[a + 1] - d_.l2_offsets[a];
  }
  __device__ size_t getNumLastL2Blocks(uint32_t const a) const { return d_.num_last_l2_blocks[a]; }
  template <uint32_t Value>
  __device__ size_t getTotalNumVals(uint32_t const a) const {
    return Value ? d_.total_ones[a] : bit_array_.size(a) - d_.total_ones[a];
  }
  template <uint32_t Value>
  __device__ size_t getSelectSample(uint32_t const a, size_t const index) const {
    return Value ? d_.samples[d_.sample_offsets[a] + index]
                 : d_.samples[d_.sample_offsets[a] + d_.sample_caps[a] - 1 - index];
  }
  __device__ RSConfig::L1_TYPE getL1Entry(uint32_t const a, size_t const index) const {
    return d_.l1[d_.l1_offsets[a] + index];
  }
  __device__ size_t getL2Entry(uint32_t const a, size_t const index) const {
    return d_.l2[d_.l2_offsets[a] + index];
  }
  __device__ void writeL2Index(uint32_t const a, size_t const index, RSConfig::L2_TYPE const v) {
    d_.l2[d_.l2_offsets[a] + index] = v;
  }
  __device__ void writeL1Index(uint32_t const a, size_t const index, RSConfig::L1_TYPE const v) {
    d_.l1[d_.l1_offsets[a] + index] = v;
  }
  __device__ void writeTotalNumOnes(uint32_t const a, size_t const v) { d_.total_ones[a] = v; }
  template <uint32_t Value>
  __device__ void writeSelectSample(uint32_t const a, size_t const index, size_t const v) {
    if (Value)
      d_.samples[d_.sample_offsets[a] + index] = v;
    else
      d_.samples[d_.sample_offsets[a] + d_.sample_caps[a] - 1 - index] = v;
  }
  __device__ void writeNumLastL2Blocks(uint32_t const a, uint16_t const v) {
    d_.num_last_l2_blocks[a] = v;
  }
  __host__ RSConfig::L1_TYPE* getL1EntryPointer(uint32_t const a, size_t const index) const {
    uint64_t off = 0;
    ecl_rank_select_get_l1(handle_, a, 0, 0, nullptr);  // validates a
    ecl_rank_select_info inf;
    for (uint32_t k = 0; k < a; ++k) {
      utils::throwOnError(ecl_rank_select_info_get(handle_, k, &inf));
      off += inf.num_l1_blocks;
    }
    return d_.l1 + off + index;
  }

  // RankSelect::getNeededGPUMemory (rank_select.cu:382-409)
  __host__ static size_t getNeededGPUMemory(size_t const size, uint8_t const num_arrays) {
    return ecl_rank_select_needed_gpu_memory(size, num_arrays);
  }

  __host__ ecl_rank_select handle() const { return handle_; }

 private:
  template <typename T> friend class WaveletTree;
  __host__ void release() {
    if (!is_copy_ && handle_ != nullptr) ecl_rank_select_destroy(handle_);
    handle_ = nullptr;
    is_copy_ = true;
  }
  __host__ void bind() {
    utils::throwOnError(ecl_rank_select_device_pointers(handle_, &d_));
    levels_ = static_cast<const detail::LevelView*>(d_.levels);
    ecl_bit_array ba = nullptr;
    utils::throwOnError(ecl_rank_select_bit_array(handle_, &ba));
    bit_array_.adopt(ba);
  }
  // view a RankSelect owned by a WaveletTree
  __host__ void adopt(ecl_rank_select h) {
    release();
    handle_ = h;
    is_copy_ = true;
    bind();
  }

  ecl_rank_select handle_ = nullptr;
  ecl_rank_select_device d_{};
  const detail::LevelView* levels_ = nullptr;
  bool is_copy_ = true;
};

}  // namespace ecl
