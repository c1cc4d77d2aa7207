// capi.hip -- the C ABI of libecl_wavelet_hip.so: object lifetimes, memory
// admission, build orchestration and the batch query entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "host.hpp"

using ecl::detail::LevelView;
using ecl::detail::TreeView;
using namespace ecl::lib;
namespace detail = ecl::detail;

// ===========================================================================
// errors, buffers, profiling
// ===========================================================================
namespace ecl {
namespace lib {

static thread_local std::string g_last_error;
std::string& last_error() { return g_last_error; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(ECL_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- per-device block cache -----------------------------------------------
// hipMalloc and hipFree of multi-GB blocks take from milliseconds to most of
// a second while the runtime maps and unmaps the memory, and the library
// repeats the same sizes (a tree, its directory and its query and select
// lines when trees of one shape are rebuilt; a forest's arena; a server's
// trees).  Freed blocks of at least kPoolMinBytes are therefore kept per
// device and handed out again for a request of the same size class (the
// smallest cached block of at least the request and at most 1/8 plus one
// grain more).  A release synchronizes the block's device first, as hipFree
// does, so no kernel still uses a block when it is handed out again.  The
// cache keeps at most kPoolCapFraction of the device's memory (oldest
// blocks are freed first); when an allocation fails, the whole cache is
// freed and the allocation tried once more.
namespace {
constexpr size_t kPoolGrain = size_t(2) << 20;
constexpr int kPoolDevices = 64;
constexpr double kPoolCapFraction = 0.4;
struct PoolBlock { void* ptr; size_t cap; uint64_t seq; };
struct BlockPool {
  std::mutex mu;
  std::vector<PoolBlock> blocks;
  size_t cached = 0;
  size_t limit = 0;  // 0 until the device's memory size is known
  uint64_t seq = 0;
};
// never destroyed: freeing device memory during static destruction may run
// after the HIP runtime has shut down
BlockPool* const g_pool = new BlockPool[kPoolDevices];

struct DeviceScope {  // makes `dev` current for a scope
  int prev = -1;
  explicit DeviceScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

void pool_free_all(BlockPool& P) {  // P.mu held, device current
  for (const PoolBlock& b : P.blocks) (void)hipFree(b.ptr);
  P.blocks.clear();
  P.cached = 0;
}
}  // namespace

size_t pooled_bytes(int device) {
  if (device < 0 || device >= kPoolDevices) return 0;
  std::lock_guard<std::mutex> lk(g_pool[device].mu);
  return g_pool[device].cached;
}
void release_block_pool(int device) {
  if (device < 0 || device >= kPoolDevices) return;
  DeviceScope ds(device);
  std::lock_guard<std::mutex> lk(g_pool[device].mu);
  pool_free_all(g_pool[device]);
}

DevBuf::DevBuf(size_t b) : bytes(b) {
  if (b == 0) return;
  ECL_HIP(hipGetDevice(&dev));
  if (b >= kPoolMinBytes && dev >= 0 && dev < kPoolDevices) {
    const size_t want = (b + kPoolGrain - 1) / kPoolGrain * kPoolGrain;
    BlockPool& P = g_pool[dev];
    std::lock_guard<std::mutex> lk(P.mu);
    size_t best = SIZE_MAX;
    for (size_t i = 0; i < P.blocks.size(); ++i) {
      const size_t c = P.blocks[i].cap;
      if (c >= want && c <= want + want / 8 + kPoolGrain &&
          (best == SIZE_MAX || c < P.blocks[best].cap))
        best = i;
    }
    if (best != SIZE_MAX) {
      ptr = P.blocks[best].ptr;
      cap = P.blocks[best].cap;
      P.cached -= cap;
      P.blocks.erase(P.blocks.begin() + static_cast<std::ptrdiff_t>(best));
      return;
    }
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess && !P.blocks.empty()) {
      (void)hipGetLastError();
      pool_free_all(P);
      e = hipMalloc(&ptr, want);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ptr = nullptr;
      throw Error(ECL_ERROR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    cap = want;
    return;
  }
  hipError_t e = hipMalloc(&ptr, b);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    throw Error(ECL_ERROR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  cap = b;
}
DevBuf::~DevBuf() { reset(); }
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    reset();
    ptr = o.ptr; bytes = o.bytes; cap = o.cap; dev = o.dev;
    o.ptr = nullptr; o.bytes = 0; o.cap = 0; o.dev = -1;
  }
  return *this;
}
void DevBuf::reset() {
  if (ptr) {
    if (cap >= kPoolMinBytes && dev >= 0 && dev < kPoolDevices) {
      DeviceScope ds(dev);
      (void)hipDeviceSynchronize();  // hipFree's implicit wait
      (void)hipGetLastError();
      BlockPool& P = g_pool[dev];
      std::lock_guard<std::mutex> lk(P.mu);
      if (P.limit == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess)
          P.limit = static_cast<size_t>(static_cast<double>(tot) * kPoolCapFraction);
        else
          (void)hipGetLastError();
      }
      P.blocks.push_back({ptr, cap, P.seq++});
      P.cached += cap;
      while (P.cached > P.limit && !P.blocks.empty()) {
        size_t old = 0;
        for (size_t i = 1; i < P.blocks.size(); ++i)
          if (P.blocks[i].seq < P.blocks[old].seq) old = i;
        (void)hipFree(P.blocks[old].ptr);
        P.cached -= P.blocks[old].cap;
        P.blocks.erase(P.blocks.begin() + static_cast<std::ptrdiff_t>(old));
      }
    } else {
      (void)hipFree(ptr);
    }
  }
  ptr = nullptr;
  bytes = 0;
  cap = 0;
  dev = -1;
}

// ---- uploads ----------------------------------------------------------------
void Uploader::clear() {
  b_.nseg = 0;
  b_.off[0] = 0;
}
void Uploader::flush() {
  if (b_.nseg) launch_upload(b_, s_);
  clear();
}
void Uploader::add_bytes(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  if (bytes > (64u << 10)) {
    flush();
    ECL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return;
  }
  const uint8_t* from = static_cast<const uint8_t*>(src);
  uint8_t* to = static_cast<uint8_t*>(dst);
  while (bytes) {
    if (b_.nseg == kUploadSegs || b_.off[b_.nseg] == kUploadBytes) flush();
    const uint32_t at = b_.off[b_.nseg];
    const size_t take = std::min<size_t>(bytes, kUploadBytes - at);
    std::memcpy(b_.data + at, from, take);
    b_.dst[b_.nseg] = reinterpret_cast<uint64_t>(to);
    b_.off[++b_.nseg] = at + static_cast<uint32_t>(take);
    from += take;
    to += take;
    bytes -= take;
  }
}

// ---- host copy threads -----------------------------------------------------------
namespace {
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool();  // never destroyed (threads run to exit)
    return *pool;
  }
  void copy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kPiece = 1u << 20;  // copies of 2 MiB and more are split
    const size_t parts = std::min<size_t>(workers_ + 1, bytes / kPiece);
    if (parts <= 1) {
      std::memcpy(dst, src, bytes);
      return;
    }
    struct Batch {
      std::mutex m;
      std::condition_variable cv;
      size_t left;
    } batch;
    batch.left = parts - 1;
    const size_t per = (bytes / parts + 63) & ~size_t(63);
    auto* d = static_cast<uint8_t*>(dst);
    auto* sp = static_cast<const uint8_t*>(src);
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (size_t i = 1; i < parts; ++i) {
        const size_t a = i * per, b = std::min(bytes, a + per);
        tasks_.push_back([=, &batch] {
          if (a < b) std::memcpy(d + a, sp + a, b - a);
          std::lock_guard<std::mutex> bl(batch.m);
          if (--batch.left == 0) batch.cv.notify_one();
        });
      }
    }
    cv_.notify_all();
    std::memcpy(d, sp, std::min(per, bytes));
    std::unique_lock<std::mutex> bl(batch.m);
    batch.cv.wait(bl, [&] { return batch.left == 0; });
  }

 private:
  CopyPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    workers_ = std::min<unsigned>(7, hw > 1 ? hw / 2 : 0);
    for (unsigned i = 0; i < workers_; ++i) std::thread([this] { run(); }).detach();
  }
  void run() {
    for (;;) {
      std::function<void()> task;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !tasks_.empty(); });
        task = std::move(tasks_.front());
        tasks_.pop_front();
      }
      task();
    }
  }
  unsigned workers_ = 0;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::function<void()>> tasks_;
};

struct Staging {
  std::mutex mu;
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
};
constexpr size_t kStageBytes = 64u << 20;
Staging* const g_staging = new Staging();  // never destroyed
}  // namespace

void parallel_memcpy(void* dst, const void* src, size_t bytes) {
  CopyPool::get().copy(dst, src, bytes);
}

void upload_staged(void* d_dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < 2 * kStageBytes) {
    h2d(d_dst, static_cast<const uint8_t*>(src), bytes);
    return;
  }
  Staging& st = *g_staging;
  std::lock_guard<std::mutex> lk(st.mu);
  for (int b = 0; b < 2; ++b) {
    if (!st.pin[b]) ECL_HIP(hipHostMalloc(&st.pin[b], kStageBytes, hipHostMallocDefault));
    if (!st.done[b]) ECL_HIP(hipEventCreateWithFlags(&st.done[b], hipEventDisableTiming));
  }
  auto* d = static_cast<uint8_t*>(d_dst);
  auto* sp = static_cast<const uint8_t*>(src);
  bool used[2] = {false, false};
  for (size_t off = 0, c = 0; off < bytes; off += kStageBytes, ++c) {
    const int b = static_cast<int>(c & 1);
    const size_t len = std::min(kStageBytes, bytes - off);
    if (used[b]) ECL_HIP(hipEventSynchronize(st.done[b]));  // its last transfer is done
    parallel_memcpy(st.pin[b], sp + off, len);
    ECL_HIP(hipMemcpyAsync(d + off, st.pin[b], len, hipMemcpyHostToDevice, s));
    ECL_HIP(hipEventRecord(st.done[b], s));
    used[b] = true;
  }
  ECL_HIP(hipStreamSynchronize(s));
}

// ---- per-device scratch cache ------------------------------------------------
namespace {
struct ScratchCache {
  std::mutex mu;
  DevBuf buf;
};
constexpr int kMaxDevices = 64;
// never destroyed: freeing device memory during static destruction may
// run after the HIP runtime has shut down
ScratchCache* const g_scratch = new ScratchCache[kMaxDevices];
}  // namespace

Scratch::Scratch(int device, hipStream_t s, size_t bytes) : device_(device), s_(s) {
  bytes = std::max<size_t>(units(bytes), 256);
  if (device >= 0 && device < kMaxDevices && g_scratch[device].mu.try_lock()) {
    ScratchCache& c = g_scratch[device];
    try {
      if (c.buf.bytes < bytes) {
        c.buf.reset();
        c.buf = DevBuf(bytes);
      }
    } catch (...) {
      c.mu.unlock();
      throw;
    }
    cached_ = true;
    base_ = c.buf.as<uint8_t>();
    cap_ = c.buf.bytes;
  } else {
    own_ = DevBuf(bytes);
    base_ = own_.as<uint8_t>();
    cap_ = bytes;
  }
}
Scratch::~Scratch() {
  // kernels that use the buffer finish before another build may take it
  (void)hipStreamSynchronize(s_);
  (void)hipGetLastError();
  if (cached_) g_scratch[device_].mu.unlock();
}
size_t cached_scratch_bytes(int device) {
  if (device < 0 || device >= kMaxDevices) return 0;
  ScratchCache& c = g_scratch[device];
  if (!c.mu.try_lock()) return 0;
  const size_t b = c.buf.bytes;
  c.mu.unlock();
  return b;
}
void release_scratch_cache(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch[device].mu);
  g_scratch[device].buf.reset();
}

static const char* kKernelNames[K_COUNT] = {
    "histogram", "wt_pass", "rs_build", "wt_access", "wt_rank", "wt_select",
    "wt_mixed", "rs_rank", "rs_select", "generate", "fill", "tile_hist", "tile_scan",
    "forest_query"};
const char* kernel_name(int id) { return id >= 0 && id < K_COUNT ? kKernelNames[id] : "?"; }

struct ProfRecord { int id; hipEvent_t e0, e1; };
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRecord> g_prof_pending;
static uint64_t g_prof_count[K_COUNT];
static double g_prof_ms[K_COUNT];

ProfScope::ProfScope(int i, hipStream_t s) : id(i), stream(s) {
  if (!g_prof_on) return;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    e0 = e1 = nullptr;
    return;
  }
  (void)hipEventRecord(e0, s);
}
ProfScope::~ProfScope() {
  if (!e0) return;
  (void)hipEventRecord(e1, stream);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_pending.push_back({id, e0, e1});
}

static void prof_drain() {
  std::vector<ProfRecord> pend;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    pend.swap(g_prof_pending);
  }
  for (auto& r : pend) {
    float ms = 0.f;
    if (hipEventSynchronize(r.e1) == hipSuccess &&
        hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      g_prof_count[r.id]++;
      g_prof_ms[r.id] += ms;
    }
    (void)hipEventDestroy(r.e0);
    (void)hipEventDestroy(r.e1);
  }
}

}  // namespace lib
}  // namespace ecl

namespace ecl {
namespace lib {

void set_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    throw Error(ECL_ERROR_NO_DEVICE, "no HIP device available");
  }
  require(device >= 0 && device < n, "invalid device index");
  ECL_HIP(hipSetDevice(device));
}

// per-device cap on the free memory the admission checks assume (0: none)
constexpr int kCapDevices = 64;
std::atomic<uint64_t> g_mem_cap[kCapDevices];

// free device memory as the library sees it: what the runtime reports plus
// the blocks the library's cache holds for reuse, at most the device's cap
size_t free_device_memory() {
  size_t fr = 0, tot = 0;
  ECL_HIP(hipMemGetInfo(&fr, &tot));
  int dev = -1;
  ECL_HIP(hipGetDevice(&dev));
  const size_t f = fr + pooled_bytes(dev);
  const uint64_t cap = dev >= 0 && dev < kCapDevices ? g_mem_cap[dev].load() : 0;
  return cap ? std::min<size_t>(f, cap) : f;
}

// MemAvailable of /proc/meminfo in bytes (getAvailableMemoryLinux,
// utils.cu:122-139); 0 when it cannot be read
uint64_t available_host_memory() {
  std::FILE* f = std::fopen("/proc/meminfo", "r");
  if (!f) return 0;
  char line[256];
  uint64_t kb = 0;
  while (std::fgets(line, sizeof line, f)) {
    unsigned long long v = 0;
    if (std::sscanf(line, "MemAvailable: %llu kB", &v) == 1) {
      kb = v;
      break;
    }
  }
  std::fclose(f);
  return kb * 1024;
}

}  // namespace lib
}  // namespace ecl

static hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// ===========================================================================
// objects
// ===========================================================================
struct ecl_bit_array_t {
  int device = 0;
  std::vector<uint64_t> sizes;    // bits per array
  std::vector<uint64_t> offsets;  // word offset per array (multiple of 32)
  uint64_t total_words = 0;
  DevBuf data, d_sizes, d_offsets;
  uint32_t* words() const { return data.as<uint32_t>(); }
};

struct ecl_rank_select_t {
  int device = 0;
  ecl_bit_array_t* ba = nullptr;  // owned
  std::vector<uint64_t> num_l1, num_l2, num_last_l2, l1_off, l2_off, s_off, cap, ones;
  DevBuf l1, l2, samples, total_ones;
  DevBuf lines;      // optional query lines and select hints of every array
  DevBuf sel_lines;  // optional select lines of every array
  DevBuf d_l1_off, d_l2_off, d_num_l1, d_num_last_l2, d_s_off, d_cap, d_levels;
  std::vector<LevelView> levels;
  uint64_t total_l2 = 0;
  ~ecl_rank_select_t() { delete ba; }
};

// Host-issued batches: two default-priority streams, pinned staging and
// device buffers per stream, grown on demand and kept with the handle.
struct HostPipe {
  hipStream_t stream[2] = {nullptr, nullptr};
  void* pin_in[2] = {nullptr, nullptr};
  void* pin_out[2] = {nullptr, nullptr};
  DevBuf d_in[2], d_out[2];
  size_t cap_in = 0, cap_out = 0;
  std::mutex mu;
  ~HostPipe() {
    for (int i = 0; i < 2; ++i) {
      if (stream[i]) (void)hipStreamDestroy(stream[i]);
      if (pin_in[i]) (void)hipHostFree(pin_in[i]);
      if (pin_out[i]) (void)hipHostFree(pin_out[i]);
    }
  }
  void reserve(size_t in_bytes, size_t out_bytes) {
    for (int i = 0; i < 2; ++i)
      if (!stream[i]) ECL_HIP(hipStreamCreate(&stream[i]));
    if (in_bytes > cap_in) {
      for (int i = 0; i < 2; ++i) {
        if (pin_in[i]) ECL_HIP(hipHostFree(pin_in[i]));
        pin_in[i] = nullptr;
        ECL_HIP(hipHostMalloc(&pin_in[i], in_bytes, hipHostMallocDefault));
        d_in[i] = DevBuf(in_bytes);
      }
      cap_in = in_bytes;
    }
    if (out_bytes > cap_out) {
      for (int i = 0; i < 2; ++i) {
        if (pin_out[i]) ECL_HIP(hipHostFree(pin_out[i]));
        pin_out[i] = nullptr;
        ECL_HIP(hipHostMalloc(&pin_out[i], out_bytes, hipHostMallocDefault));
        d_out[i] = DevBuf(out_bytes);
      }
      cap_out = out_bytes;
    }
  }
};

struct ecl_wavelet_tree_t {
  int device = 0;
  int esize = 1;
  uint64_t n = 0, sigma = 0;
  uint32_t L = 0;
  bool min_alphabet = true, pow2 = true;
  uint64_t codes_start = 0;
  std::vector<uint64_t> alphabet;       // original symbols (as u64)
  std::vector<uint64_t> counts;         // exclusive prefix C
  std::vector<uint64_t> codes;          // codes of symbols codes_start..
  std::vector<uint8_t> code_lens;
  std::vector<uint64_t> ranks, node_start, nodes_at_level;
  std::vector<uint8_t> node_level;
  uint64_t nodes_until_last_level = 0;
  ecl_rank_select_t* rs = nullptr;      // owned
  DevBuf d_counts, d_ranks, d_nal, d_codes, d_code_lens, d_alphabet, d_lut;
  int map_kind = 0;
  TreeView view{};
  uint32_t lanes = 0;  // lanes per access/rank query (0: the library-wide setting)
  uint64_t build_host_bytes = 0;  // build buffers placed in pinned host memory
  DevBuf quad_lines, quad_base, quad_tab;  // optional quad lines, bases, node tables
  DevBuf quad_sel;                          // optional quad select lines
  QuadView quad{};
  DevBuf hex_lines, hex_tab;                // optional hex lines, bases and superblocks
  HexView hex{};
  DevBuf d_layouts;  // device copies of quad and hex (once asked for; kept current)
  std::unique_ptr<HostPipe> pipe{new HostPipe()};
  ~ecl_wavelet_tree_t() { delete rs; }
};

// ===========================================================================
// BitArray
// ===========================================================================
// Allocates the arrays; `up`, when given, carries the size/offset tables
// (else they are copied before returning).
static ecl_bit_array_t* make_bit_array(const std::vector<uint64_t>& sizes, int init,
                                       int device, Uploader* up = nullptr) {
  require(sizes.size() < 255, "a BitArray holds at most 254 arrays");
  auto* ba = new ecl_bit_array_t();
  try {
    ba->device = device;
    ba->sizes = sizes;
    ba->offsets.resize(sizes.size());
    uint64_t total = 0;
    for (size_t i = 0; i < sizes.size(); ++i) {
      ba->offsets[i] = total;
      const uint64_t w = (sizes[i] + 31) / 32;
      total += (w + 31) & ~31ull;  // 128 B per array start
    }
    total += total % 2;  // even word count: u64 loads never leave the buffer
    if (total == 0) total = 2;
    ba->total_words = total;
    ba->data = DevBuf(total * sizeof(uint32_t));
    ba->d_sizes = DevBuf(std::max<size_t>(1, sizes.size()) * sizeof(uint64_t));
    ba->d_offsets = DevBuf(std::max<size_t>(1, sizes.size()) * sizeof(uint64_t));
    if (up != nullptr) {
      up->add(ba->d_sizes.ptr, ba->sizes.data(), sizes.size());
      up->add(ba->d_offsets.ptr, ba->offsets.data(), sizes.size());
    } else {
      h2d(ba->d_sizes.ptr, ba->sizes.data(), sizes.size());
      h2d(ba->d_offsets.ptr, ba->offsets.data(), sizes.size());
    }
    if (init >= 0)
      launch_fill_u32(ba->words(), total, init ? 0xFFFFFFFFu : 0u, nullptr);
  } catch (...) {
    delete ba;
    throw;
  }
  return ba;
}

// ===========================================================================
// RankSelect
// ===========================================================================
// Directory allocation for every array of `ba` (rank_select.cu:57-143
// sizes); the contents are written by rs_build_kernel or, for a wavelet
// tree, by the level passes.
static ecl_rank_select_t* alloc_rank_select(ecl_bit_array_t* ba, hipStream_t s) {
  auto* rs = new ecl_rank_select_t();
  try {
    rs->device = ba->device;
    const size_t na = ba->sizes.size();
    rs->num_l1.resize(na); rs->num_l2.resize(na); rs->num_last_l2.resize(na);
    rs->l1_off.resize(na); rs->l2_off.resize(na); rs->s_off.resize(na);
    rs->cap.resize(na); rs->ones.assign(na, 0);
    uint64_t t1 = 0, t2 = 0, ts = 0;
    for (size_t i = 0; i < na; ++i) {
      const uint64_t m = ba->sizes[i];
      rs->num_l1[i] = (m + detail::kL1Bits - 1) / detail::kL1Bits;
      uint64_t last = ((m % detail::kL1Bits) + detail::kL2Bits - 1) / detail::kL2Bits;
      if (last == 0 && m) last = detail::kL2PerL1;
      rs->num_last_l2[i] = m ? last : 0;
      rs->num_l2[i] = m ? (rs->num_l1[i] - 1) * detail::kL2PerL1 + last : 0;
      rs->cap[i] = m / detail::kSampleRate;
      rs->l1_off[i] = t1; rs->l2_off[i] = t2; rs->s_off[i] = ts;
      t1 += rs->num_l1[i]; t2 += rs->num_l2[i]; ts += rs->cap[i];
    }
    rs->total_l2 = t2;
    rs->l1 = DevBuf(std::max<uint64_t>(t1, 1) * sizeof(uint64_t));
    rs->l2 = DevBuf(std::max<uint64_t>(t2, 1) * sizeof(uint16_t));
    rs->samples = DevBuf(std::max<uint64_t>(ts, 1) * sizeof(uint64_t));
    rs->total_ones = DevBuf(std::max<size_t>(na, 1) * sizeof(uint64_t));
    launch_fill_u64(rs->total_ones.as<uint64_t>(), std::max<size_t>(na, 1), 0, s);
  } catch (...) {
    delete rs;
    throw;
  }
  return rs;
}

// Totals back to the host (the directory kernels must have finished),
// reference-shaped device directory and the LevelView table through `up`;
// takes ownership of `ba`.
static void finalize_rank_select(ecl_rank_select_t* rs, ecl_bit_array_t* ba, Uploader& up) {
  const size_t na = ba->sizes.size();
  d2h(rs->ones.data(), rs->total_ones.ptr, na);
  auto put = [&](DevBuf& b, const std::vector<uint64_t>& v) {
    b = DevBuf(std::max<size_t>(v.size(), 1) * sizeof(uint64_t));
    up.add(b.ptr, v.data(), v.size());
  };
  put(rs->d_l1_off, rs->l1_off);
  put(rs->d_l2_off, rs->l2_off);
  put(rs->d_num_l1, rs->num_l1);
  put(rs->d_s_off, rs->s_off);
  put(rs->d_cap, rs->cap);
  std::vector<uint16_t> last16(na);
  for (size_t i = 0; i < na; ++i) last16[i] = static_cast<uint16_t>(rs->num_last_l2[i]);
  rs->d_num_last_l2 = DevBuf(std::max<size_t>(na, 1) * sizeof(uint16_t));
  up.add(rs->d_num_last_l2.ptr, last16.data(), na);
  rs->levels.resize(na);
  for (size_t i = 0; i < na; ++i) {
    LevelView& v = rs->levels[i];
    v.words = ba->words() + ba->offsets[i];
    v.l1 = rs->l1.as<uint64_t>() + rs->l1_off[i];
    v.l2 = rs->l2.as<uint16_t>() + rs->l2_off[i];
    v.samples = rs->samples.as<uint64_t>() + rs->s_off[i];
    v.size = ba->sizes[i];
    v.ones = rs->ones[i];
    v.num_l1 = rs->num_l1[i];
    v.sample_cap = rs->cap[i];
  }
  rs->d_levels = DevBuf(std::max<size_t>(na, 1) * sizeof(LevelView));
  up.add(rs->d_levels.ptr, rs->levels.data(), na);
  rs->ba = ba;
}

// Scratch of build_directories for `ba`: the array table and 2 u32 per L1
// block (upper bound for arrays of the given sizes).
static size_t directory_scratch_bytes(const std::vector<uint64_t>& sizes) {
  uint64_t blocks = 0;
  for (uint64_t m : sizes) blocks += (m + detail::kL1Bits - 1) / detail::kL1Bits;
  return Scratch::units(std::max<size_t>(sizes.size(), 1) * sizeof(RsArrayDesc)) +
         Scratch::units(std::max<uint64_t>(blocks, 1) * 8);
}

// Directory of every non-empty array of `ba` into `rs`: the array table goes
// up through `up`, then the directory kernels are queued on `s` (nothing
// waits for them here).
static void build_directories(ecl_rank_select_t* rs, ecl_bit_array_t* ba, hipStream_t s,
                              Scratch& scr, Uploader& up) {
  const size_t na = ba->sizes.size();
  std::vector<RsArrayDesc> desc;
  uint64_t blocks = 0, chunks = 0;
  for (size_t i = 0; i < na; ++i) {
    if (ba->sizes[i] == 0) continue;
    RsArrayDesc d{};
    d.words64 = reinterpret_cast<const uint64_t*>(ba->words() + ba->offsets[i]);
    d.m = ba->sizes[i];
    d.l1 = rs->l1.as<uint64_t>() + rs->l1_off[i];
    d.l2 = rs->l2.as<uint16_t>() + rs->l2_off[i];
    d.num_l1 = rs->num_l1[i];
    d.num_l2 = rs->num_l2[i];
    d.samples = rs->samples.as<uint64_t>() + rs->s_off[i];
    d.cap = rs->cap[i];
    d.total_ones = rs->total_ones.as<uint64_t>() + i;
    d.first_tile = blocks;
    d.first_chunk = chunks;
    blocks += d.num_l1;
    chunks += rs_scan_chunks(d.num_l1);
    desc.push_back(d);
  }
  if (desc.empty()) return;
  RsArrayDesc* d_desc = scr.carve<RsArrayDesc>(na);
  uint32_t* d_blk = scr.carve<uint32_t>(2 * blocks);
  up.add(d_desc, desc.data(), desc.size());
  up.flush();
  launch_rs_build(d_desc, desc.size(), blocks, chunks, d_blk, s);
}

// RankSelect(BitArray&&) (rank_select.cu:50-304): one rs_build pass per array.
static ecl_rank_select_t* make_rank_select(ecl_bit_array_t* ba, hipStream_t s) {
  ecl_rank_select_t* rs = alloc_rank_select(ba, s);
  try {
    Scratch scr(ba->device, s, directory_scratch_bytes(ba->sizes));
    Uploader up(s);
    build_directories(rs, ba, s, scr, up);
    ECL_HIP(hipStreamSynchronize(s));
    finalize_rank_select(rs, ba, up);
    up.flush();
    ECL_HIP(hipStreamSynchronize(s));
  } catch (...) {
    rs->ba = nullptr;
    delete rs;
    throw;
  }
  return rs;
}

// ===========================================================================
// WaveletTree: host-side shape (codes, nodes, node tables)
// ===========================================================================
namespace ecl {
namespace lib {
namespace {

// Minimal codes as the leaves of the left-complete tree: a node of s >= 2
// symbols splits after prev_pow2(s) symbols (the shape
// createMinimalCodes, wavelet_tree.cuh:2100-2159, produces).
void assign_codes(uint64_t lo, uint64_t hi, uint32_t depth, uint64_t prefix, uint32_t L,
                  std::vector<uint64_t>& code, std::vector<uint8_t>& len,
                  std::vector<std::vector<Node>>& levels) {
  // iterative over the left spine to bound recursion depth by L
  while (true) {
    if (lo == hi) {
      code[lo] = L > depth ? (prefix << (L - depth)) : prefix;
      len[lo] = static_cast<uint8_t>(depth);
      return;
    }
    levels[depth].push_back({lo, hi});
    const uint64_t mid = lo + detail::prev_pow2(hi - lo + 1);
    assign_codes(lo, mid - 1, depth + 1, prefix << 1, L, code, len, levels);
    lo = mid;
    depth += 1;
    prefix = (prefix << 1) | 1;
  }
}

// first-round count of coded symbols (wavelet_tree.cuh:2122-2133)
uint64_t codes_start_of(uint64_t sigma, uint32_t L) {
  if (detail::is_pow2(sigma)) return sigma;
  uint64_t num = sigma;
  for (uint32_t i = L - 1; i > 0; --i) {
    const uint64_t p = 1ull << i;
    if (num <= p) break;
    num -= p;
  }
  return sigma - num;
}

}  // namespace

TreeShape tree_shape(uint64_t sigma) {
  TreeShape sh;
  sh.sigma = sigma;
  sh.L = detail::ceil_log2(sigma);
  sh.pow2 = detail::is_pow2(sigma);
  sh.code.resize(sigma);
  sh.len.resize(sigma);
  sh.nodes.resize(sh.L + 1);
  assign_codes(0, sigma - 1, 0, 0, sh.L, sh.code, sh.len, sh.nodes);
  sh.codes_start = codes_start_of(sigma, sh.L);
  return sh;
}

// node ranks: zeros of level l before each node start (precomputeRanksKernel,
// wavelet_tree.cuh:332-352, over the node starts of :2501-2532)
TreeMeta tree_meta(const TreeShape& sh, const std::vector<uint64_t>& sym_count, uint64_t n) {
  TreeMeta m;
  const uint64_t sigma = sh.sigma;
  const uint32_t L = sh.L;
  m.counts.resize(sigma);
  uint64_t acc = 0;
  for (uint64_t c = 0; c < sigma; ++c) { m.counts[c] = acc; acc += sym_count[c]; }
  if (acc != n) throw Error(ECL_ERROR_INTERNAL, "histogram does not sum to n");
  auto C = [&](uint64_t c) { return c >= sigma ? n : m.counts[c]; };
  m.nodes_at_level.assign(L > 1 ? L - 1 : 0, 0);
  for (uint32_t l = 0; l < L; ++l) {
    uint64_t zb = 0;  // zeros of level l before the node
    for (const Node& nd : sh.nodes[l]) {
      const uint64_t mid = nd.lo + detail::prev_pow2(nd.hi - nd.lo + 1);
      if (l >= 1 && nd.lo >= 1) {
        m.ranks.push_back(zb);
        m.node_start.push_back(nd.lo);
        m.node_level.push_back(static_cast<uint8_t>(l));
        if (l + 1 < L) m.nodes_at_level[l]++;
      }
      zb += C(mid) - C(nd.lo);
    }
  }
  for (uint64_t x : m.nodes_at_level) m.nodes_until_last_level += x;
  return m;
}

std::vector<uint64_t> level_sizes(const TreeShape& sh, const std::vector<uint64_t>& sym_count) {
  std::vector<uint64_t> out(sh.L, 0);
  for (uint32_t l = 0; l < sh.L; ++l)
    for (uint64_t c = 0; c < sh.sigma; ++c)
      if (sh.len[c] > l) out[l] += sym_count[c];
  return out;
}

}  // namespace lib
}  // namespace ecl

namespace {

template <typename T>
std::vector<uint64_t> to_u64(const void* a, size_t n) {
  std::vector<uint64_t> v(n);
  const T* p = static_cast<const T*>(a);
  for (size_t i = 0; i < n; ++i) v[i] = p[i];
  return v;
}
std::vector<uint64_t> elems_to_u64(const void* a, size_t n, int esize) {
  switch (esize) {
    case 1: return to_u64<uint8_t>(a, n);
    case 2: return to_u64<uint16_t>(a, n);
    case 4: return to_u64<uint32_t>(a, n);
    default: return to_u64<uint64_t>(a, n);
  }
}
void u64_to_elems(const std::vector<uint64_t>& v, void* out, int esize) {
  for (size_t i = 0; i < v.size(); ++i) {
    switch (esize) {
      case 1: static_cast<uint8_t*>(out)[i] = static_cast<uint8_t>(v[i]); break;
      case 2: static_cast<uint16_t*>(out)[i] = static_cast<uint16_t>(v[i]); break;
      case 4: static_cast<uint32_t*>(out)[i] = static_cast<uint32_t>(v[i]); break;
      default: static_cast<uint64_t*>(out)[i] = v[i]; break;
    }
  }
}

// Construction plan: codes are stored cw bytes wide and the levels are
// built in passes of at most 8 levels (kernels_tree.hip).
struct BuildPlan {
  int cw = 1;
  uint32_t npasses = 1;
  std::vector<uint32_t> k, l0;  // levels per pass and first level
};
BuildPlan plan_of(uint32_t L) {
  BuildPlan b;
  b.cw = L <= 8 ? 1 : L <= 16 ? 2 : L <= 32 ? 4 : 8;
  b.npasses = (L + 7) / 8;
  b.k.resize(b.npasses);
  b.l0.resize(b.npasses);
  uint32_t acc = 0;
  for (uint32_t q = 0; q < b.npasses; ++q) {
    b.k[q] = L / b.npasses + (q < L % b.npasses ? 1u : 0u);
    b.l0[q] = acc;
    acc += b.k[q];
  }
  return b;
}

// Memory a tree of n symbols needs beyond its text (upper bound, bytes), in
// two parts.  resident: what the build keeps on the device in any case --
// bits + L2/L1/samples per level, the tile histograms and their scan,
// metadata and the reference's 10M-query headroom (wavelet_tree.cuh:
// 644-648).  staging: a host text's copy and the code buffers (the mapped
// codes and the level-order buffers of multi-pass trees), each of which goes
// to pinned host memory when the device cannot hold it after the resident
// part (the reference's unified-memory fallback, wavelet_tree.cuh:655-689).
struct TreeMemory {
  uint64_t resident = 0;
  uint64_t text = 0;     // the device copy of a host text
  uint64_t x = 0, u = 0, v = 0;  // mapped codes, level-order buffers
  uint64_t staging() const { return text + x + u + v; }
};
TreeMemory tree_memory(uint64_t n, uint64_t sigma, uint32_t L, int esize, bool host_text,
                       bool identity) {
  const BuildPlan b = plan_of(L);
  const uint64_t per_level = ((n + 1023) / 1024) * 128 + 256 /* bits, padded */ +
                             ((n + 511) / 512) * 2 + ((n + 65535) / 65536) * 8 +
                             (n / 4096) * 8 + 64;
  const uint64_t tiles = n / pass_tile_codes(b.cw, 4) + 2 * sigma + 2;
  const uint64_t scan = tiles * 256 * 12 + (tiles / 256 + 1) * 256 * 8;
  TreeMemory m;
  m.resident = per_level * L + scan + sigma * 64 + (n / 4096 + 4096) * 8 +
               10'000'000ull * 16;
  // level order with padded nodes (the build's pad_cap)
  const uint64_t pad = (n + (2 * sigma + 2) * (16 / b.cw)) * b.cw;
  m.text = host_text ? n * esize : 0;
  m.x = identity ? 0 : (b.npasses > 2 ? pad : n * b.cw);
  m.u = b.npasses > 1 ? pad : 0;
  m.v = b.npasses > 2 && identity ? pad : 0;
  return m;
}
uint64_t tree_needed_bytes(uint64_t n, uint64_t sigma, uint32_t L, int esize, bool host_text) {
  return tree_memory(n, sigma, L, esize, host_text, false).resident +
         tree_memory(n, sigma, L, esize, host_text, false).staging();
}

// A build buffer on the device, or in pinned host memory that the kernels
// read and write over the host link when the device cannot hold it.
struct StagingBuf {
  DevBuf dev;
  void* host = nullptr;
  void* ptr = nullptr;
  StagingBuf() = default;
  StagingBuf(const StagingBuf&) = delete;
  StagingBuf& operator=(const StagingBuf&) = delete;
  ~StagingBuf() { reset(); }
  void alloc(uint64_t bytes, bool on_host) {
    reset();
    if (bytes == 0) return;
    if (on_host) {
      ECL_HIP(hipHostMalloc(&host, bytes, hipHostMallocDefault));
      ptr = host;
    } else {
      dev = DevBuf(bytes);
      ptr = dev.ptr;
    }
  }
  void reset() {
    dev.reset();
    if (host) (void)hipHostFree(host);
    host = ptr = nullptr;
  }
};

}  // namespace

static const char* kOomTree = "Not enough memory available for the wavelet tree.";

// Fields that follow from the alphabet (wt->alphabet set, sorted): size,
// levels, minimal alphabet and symbol map (wavelet_tree.cuh:590-638), codes
// (createMinimalCodes, :2100-2159); returns the tree shape.
static TreeShape set_alphabet(ecl_wavelet_tree_t* wt) {
  require(wt->alphabet.size() >= 2, "the alphabet needs at least 2 symbols");
  const uint64_t sigma = wt->sigma = wt->alphabet.size();
  wt->L = detail::ceil_log2(sigma);
  require(wt->L <= 64, "alphabet too large");
  wt->pow2 = detail::is_pow2(sigma);
  wt->min_alphabet = true;
  for (uint64_t i = 0; i < sigma; ++i)
    if (wt->alphabet[i] != i) { wt->min_alphabet = false; break; }
  wt->map_kind = wt->min_alphabet ? 0 : (wt->esize <= 2 ? 1 : 2);
  TreeShape shape = tree_shape(sigma);
  wt->codes_start = shape.codes_start;
  wt->codes.clear();
  wt->code_lens.clear();
  for (uint64_t c = wt->codes_start; c < sigma; ++c) {
    wt->codes.push_back(shape.code[c]);
    wt->code_lens.push_back(shape.len[c]);
  }
  return shape;
}

// Symbol -> index table (8/16-bit symbols), the alphabet and the codes on
// the device, through `up` (flushed here).
static void upload_symbol_maps(ecl_wavelet_tree_t* wt, Uploader& up) {
  const uint64_t sigma = wt->sigma;
  const int esize = wt->esize;
  if (wt->map_kind == 1) {
    const uint64_t range = esize == 1 ? 256 : 65536;
    std::vector<uint32_t> lut(range);
    uint64_t j = 0;
    for (uint64_t v = 0; v < range; ++v) {  // lower_bound
      while (j < sigma && wt->alphabet[j] < v) ++j;
      lut[v] = static_cast<uint32_t>(std::min<uint64_t>(j, sigma));
    }
    wt->d_lut = DevBuf(range * 4);
    up.add(wt->d_lut.ptr, lut.data(), range);
  }
  wt->d_alphabet = DevBuf(sigma * esize);
  {
    std::vector<uint8_t> raw(sigma * esize);
    u64_to_elems(wt->alphabet, raw.data(), esize);
    up.add(wt->d_alphabet.ptr, raw.data(), sigma * esize);
  }
  if (!wt->pow2) {
    wt->d_codes = DevBuf(std::max<size_t>(wt->codes.size(), 1) * 8);
    up.add(wt->d_codes.ptr, wt->codes.data(), wt->codes.size());
    wt->d_code_lens = DevBuf(std::max<size_t>(wt->code_lens.size(), 1));
    up.add(wt->d_code_lens.ptr, wt->code_lens.data(), wt->code_lens.size());
  }
  up.flush();
}

// Host metadata from the symbol counts (C, node ranks, nodes per level),
// the directory totals and level views, the device copies of the metadata
// and the tree view; the directory kernels have finished.  Takes the bit
// array and the directory.
static void finish_tree(ecl_wavelet_tree_t* wt, const TreeShape& shape,
                        const std::vector<uint64_t>& sym_count,
                        std::unique_ptr<ecl_bit_array_t>& ba,
                        std::unique_ptr<ecl_rank_select_t>& rs, Uploader& up, hipStream_t s) {
  const uint64_t sigma = wt->sigma;
  {
    TreeMeta meta = tree_meta(shape, sym_count, wt->n);
    wt->counts = std::move(meta.counts);
    wt->ranks = std::move(meta.ranks);
    wt->node_start = std::move(meta.node_start);
    wt->node_level = std::move(meta.node_level);
    wt->nodes_at_level = std::move(meta.nodes_at_level);
    wt->nodes_until_last_level = meta.nodes_until_last_level;
  }
  finalize_rank_select(rs.get(), ba.get(), up);
  ba.release();
  wt->rs = rs.release();
  // device metadata (on the device when the build returns)
  wt->d_counts = DevBuf(sigma * 8);
  up.add(wt->d_counts.ptr, wt->counts.data(), sigma);
  wt->d_ranks = DevBuf(std::max<size_t>(wt->ranks.size(), 1) * 8);
  up.add(wt->d_ranks.ptr, wt->ranks.data(), wt->ranks.size());
  wt->d_nal = DevBuf(std::max<size_t>(wt->nodes_at_level.size(), 1) * 8);
  up.add(wt->d_nal.ptr, wt->nodes_at_level.data(), wt->nodes_at_level.size());
  up.flush();
  ECL_HIP(hipStreamSynchronize(s));
  TreeView& v = wt->view;
  v.levels = wt->rs->d_levels.as<LevelView>();
  v.counts = wt->d_counts.as<uint64_t>();
  v.ranks = wt->d_ranks.as<uint64_t>();
  v.nodes_at_level = wt->d_nal.as<uint64_t>();
  v.codes = wt->d_codes.as<uint64_t>();
  v.code_lens = wt->d_code_lens.as<uint8_t>();
  v.n = wt->n;
  v.sigma = sigma;
  v.codes_start = wt->codes_start;
  v.nodes_until_last_level = wt->nodes_until_last_level;
  v.num_ranks = wt->ranks.size();
  v.num_levels = wt->L;
  v.pow2 = wt->pow2 ? 1 : 0;
}

static ecl_wavelet_tree_t* build_tree(int esize, const void* text, bool text_on_device,
                                      uint64_t n, const void* alphabet, size_t asize,
                                      int device) {
  require(n > 0, "data_size must be positive");
  require(esize == 1 || esize == 2 || esize == 4 || esize == 8, "invalid element type");
  require(text != nullptr, "text is NULL");
  set_device(device);
  hipStream_t s = nullptr;
  auto* wt = new ecl_wavelet_tree_t();
  try {
    wt->device = device;
    wt->esize = esize;
    wt->n = n;
    // ---- alphabet (wavelet_tree.cuh:590-621)
    if (asize > 0) {
      require(asize >= 2, "the alphabet needs at least 2 symbols");
      wt->alphabet = elems_to_u64(alphabet, asize, esize);
      require(std::is_sorted(wt->alphabet.begin(), wt->alphabet.end()),
              "the alphabet must be sorted");
      require(std::adjacent_find(wt->alphabet.begin(), wt->alphabet.end()) ==
                  wt->alphabet.end(), "the alphabet must not repeat symbols");
    } else if (esize <= 2) {
      const uint64_t range = esize == 1 ? 256 : 65536;
      std::vector<uint32_t> flags(range);
      if (text_on_device) {
        DevBuf d_flags(range * 4);
        launch_fill_u32(d_flags.as<uint32_t>(), range, 0, s);
        launch_presence(text, esize, n, d_flags.as<uint32_t>(), s);
        ECL_HIP(hipStreamSynchronize(s));
        d2h(flags.data(), d_flags.ptr, range);
      } else {
        for (uint64_t i = 0; i < n; ++i)
          flags[esize == 1 ? static_cast<const uint8_t*>(text)[i]
                           : static_cast<const uint16_t*>(text)[i]] = 1;
      }
      for (uint64_t v = 0; v < range; ++v)
        if (flags[v]) wt->alphabet.push_back(v);
    } else {
      std::vector<uint64_t> all;
      if (text_on_device) {
        std::vector<uint8_t> raw(n * esize);
        d2h(raw.data(), text, n * esize);
        all = elems_to_u64(raw.data(), n, esize);
      } else {
        all = elems_to_u64(text, n, esize);
      }
      std::sort(all.begin(), all.end());
      all.erase(std::unique(all.begin(), all.end()), all.end());
      wt->alphabet = std::move(all);
    }
    const TreeShape shape = set_alphabet(wt);
    const uint64_t sigma = wt->sigma;
    const uint32_t L = wt->L;
    const std::vector<uint64_t>& code = shape.code;
    const std::vector<uint8_t>& len = shape.len;
    const BuildPlan plan = plan_of(L);
    const int cw = plan.cw;
    const uint32_t rows = 4;
    const uint64_t T = pass_tile_codes(cw, rows);   // codes per tile
    const uint64_t E = 16 / cw;               // codes per 16 bytes

    // the passes read 16-byte vectors: a text that is its own code and is
    // aligned is read in place, anything else is mapped to codes first (a
    // host text's copy is aligned)
    const bool aligned =
        !text_on_device || (reinterpret_cast<uintptr_t>(text) & 15) == 0;
    const bool identity = wt->min_alphabet && wt->pow2 && esize == cw && aligned;

    // ---- memory admission (wavelet_tree.cuh:644-689): the resident part
    // must fit the device (the build's scratch reuses the device's scratch
    // cache when it holds one); then the text copy and the code buffers, in
    // that order, each stay on the device while they fit and otherwise go
    // to pinned host memory, which must fit the host's available memory
    const TreeMemory mem = tree_memory(n, sigma, L, esize, !text_on_device, identity);
    const uint64_t avail = free_device_memory() + cached_scratch_bytes(device);
    if (mem.resident > avail) throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
    uint64_t room = avail - mem.resident, host_bytes = 0;
    auto on_host = [&](uint64_t bytes) {
      if (bytes <= room) {
        room -= bytes;
        return false;
      }
      host_bytes += bytes;
      return true;
    };
    const bool text_host = on_host(mem.text);
    const bool x_host = on_host(mem.x);
    const bool u_host = on_host(mem.u);
    const bool v_host = on_host(mem.v);
    if (host_bytes > available_host_memory())
      throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
    wt->build_host_bytes = host_bytes;

    // ---- the text on the device (or in pinned host memory)
    StagingBuf own_text;
    const void* d_text = text;
    if (!text_on_device) {
      own_text.alloc(n * esize, text_host);
      if (text_host) parallel_memcpy(own_text.ptr, text, n * esize);
      else upload_staged(own_text.ptr, text, n * esize, s);
      d_text = own_text.ptr;
    }

    // Every launch of the build is queued before the host waits for the
    // device: small tables go up as kernel arguments (Uploader), and the one
    // wait comes after the directory kernels.  Multi-pass trees and
    // single-pass trees over a non-power-of-2 alphabet read their digit
    // counts back after the first scan.
    Uploader up(s);

    upload_symbol_maps(wt, up);
    up.flush();

    // ---- symbol counts: a multi-pass tree over a non-power-of-2 alphabet
    // needs its level sizes (hence the counts) before the first pass;
    // otherwise they follow from the digit counts of the passes
    std::vector<uint64_t> sym_count;
    if (plan.npasses > 1 && !wt->pow2) {
      DevBuf d_hist(sigma * 8);
      launch_fill_u64(d_hist.as<uint64_t>(), sigma, 0, s);
      HistParams hp{};
      hp.text = d_text; hp.n = n; hp.esize = esize; hp.sigma = sigma;
      hp.map_kind = wt->map_kind;
      hp.lut = wt->d_lut.as<uint32_t>();
      hp.alphabet = wt->d_alphabet.ptr;
      hp.codes_start = wt->codes_start;
      hp.pow2 = 0;
      hp.codes = wt->d_codes.as<uint64_t>();
      hp.codes_out = nullptr;
      hp.hist = d_hist.as<uint64_t>();
      launch_histogram(hp, s);
      sym_count.resize(sigma);
      ECL_HIP(hipStreamSynchronize(s));
      d2h(sym_count.data(), d_hist.ptr, sigma);
    }

    // ---- code and level-order buffers; scratch: tile histograms and their
    // scan for the largest pass, error flag, first pass digit starts and
    // thresholds, directory tables
    StagingBuf buf_x, buf_u, buf_v;
    uint64_t max_tiles = (n + T - 1) / T;
    for (uint32_t q = 1; q < plan.npasses; ++q)
      max_tiles = std::max<uint64_t>(max_tiles, n / T + (1ull << plan.l0[q]) + 1);
    const uint32_t kmax = plan.k[0];
    const uint64_t pad_cap = n + (2 * sigma + 2) * E;  // level order with padded nodes
    if (!identity) buf_x.alloc((plan.npasses > 2 ? pad_cap : n) * cw, x_host);
    if (plan.npasses > 1) buf_u.alloc(pad_cap * cw, u_host);
    if (plan.npasses > 2 && identity) buf_v.alloc(pad_cap * cw, v_host);
    if ((!identity && (plan.npasses > 2 ? pad_cap : n) * cw != mem.x) ||
        (plan.npasses > 1 ? pad_cap * cw : 0) != mem.u ||
        (plan.npasses > 2 && identity ? pad_cap * cw : 0) != mem.v)
      throw Error(ECL_ERROR_INTERNAL, "code buffer sizes differ from the memory estimate");
    const uint32_t k0 = plan.k[0], nb0 = 1u << k0;
    const uint64_t tiles0 = (n + T - 1) / T;
    Scratch scr(device, s,
                Scratch::units((max_tiles << kmax) * 4) +
                    Scratch::units(((max_tiles + 1) << kmax) * 8) +
                    Scratch::units(scan_partial_entries(max_tiles, kmax) * 8) +
                    Scratch::units(4) + Scratch::units((nb0 + 1) * 8) +
                    Scratch::units((k0 + 1) * 2) +
                    directory_scratch_bytes(std::vector<uint64_t>(L, n)));
    uint32_t* d_hist_t = scr.carve<uint32_t>(max_tiles << kmax);
    uint64_t* d_excl = scr.carve<uint64_t>((max_tiles + 1) << kmax);
    uint64_t* d_partial = scr.carve<uint64_t>(scan_partial_entries(max_tiles, kmax));
    uint32_t* d_err = scr.carve<uint32_t>(1);
    uint64_t* d_pc0 = scr.carve<uint64_t>(nb0 + 1);
    uint16_t* d_thr0 = scr.carve<uint16_t>(k0 + 1);

    PassParams pp{};
    pp.map_kind = wt->map_kind;
    pp.lut = wt->d_lut.as<uint32_t>();
    pp.alphabet = wt->d_alphabet.ptr;
    pp.sigma = sigma;
    pp.codes_start = wt->codes_start;
    pp.pow2 = wt->pow2 ? 1 : 0;
    pp.code_table = wt->d_codes.as<uint64_t>();
    pp.cw = cw;
    pp.tile_rows = rows;
    pp.tile_codes = static_cast<uint32_t>(T);
    launch_fill_u32(d_err, 1, 0, s);
    pp.err = d_err;
    pp.hist = d_hist_t;
    pp.excl = d_excl;
    pp.partial = d_partial;

    // ---- pass 0 digit counts (the whole text is one node)
    pp.src = d_text;
    pp.src_esize = esize;
    pp.codes_out = identity ? nullptr : buf_x.ptr;
    pp.codes = identity ? d_text : buf_x.ptr;
    pp.k = k0;
    pp.dshift = L - k0;
    pp.n = n;
    pp.num_tiles = tiles0;
    // A single pass over a power-of-2 alphabet needs no counts on the host
    // before the end: every level holds n bits, the digit starts are
    // scanned on the device and the thresholds do not depend on counts.
    // Its bit arrays are not zeroed: the pass stores every word once except
    // those its workgroups share, which are zeroed just before it.
    const bool single = plan.npasses == 1;
    const bool sizes_known = single && wt->pow2;
    std::vector<uint64_t> level_size(L, n);
    std::unique_ptr<ecl_bit_array_t> ba_guard;
    if (sizes_known) ba_guard.reset(make_bit_array(level_size, -1, device, &up));
    launch_tile_hist(pp, s);
    launch_tile_scan(pp, s);
    // bin counts of the current pass: [node][digit] codes of the pass input
    std::vector<uint64_t> binc(nb0);
    if (sizes_known) {
      launch_digit_starts(d_excl + tiles0 * nb0, nb0, d_pc0, s);
    } else {
      ECL_HIP(hipStreamSynchronize(s));
      d2h(binc.data(), d_excl + tiles0 * nb0, nb0);
      if (single) {  // digits are whole codes
        sym_count.resize(sigma);
        for (uint64_t c = 0; c < sigma; ++c) sym_count[c] = binc[code[c]];
      }
    }

    // ---- level sizes and the bit arrays (zeroed: run edges are OR-ed in)
    if (!sym_count.empty()) {
      uint64_t tot = 0;
      for (uint64_t c = 0; c < sigma; ++c) tot += sym_count[c];
      if (tot != n) throw Error(ECL_ERROR_INTERNAL, "histogram does not sum to n");
      level_size = level_sizes(shape, sym_count);
    }
    if (!ba_guard) ba_guard.reset(make_bit_array(level_size, 0, device, &up));
    ecl_bit_array_t* ba = ba_guard.get();
    ecl_rank_select_t* rs = alloc_rank_select(ba, s);
    std::unique_ptr<ecl_rank_select_t> rs_guard(rs);

    // ---- the passes
    std::vector<uint64_t> node_size{n};   // level-l0 node sizes of the current pass
    std::vector<uint64_t> node_pstart{0}; // their padded starts in the pass input
    DevBuf d_tstart, d_tlen, d_tnode, d_tfirst, d_rows, d_pc, d_thr, d_obase;
    const void* in_codes = pp.codes;
    for (uint32_t q = 0; q < plan.npasses; ++q) {
      const uint32_t k = plan.k[q], l0 = plan.l0[q], nb = 1u << k;
      const uint64_t nodes = 1ull << l0;
      const bool last = q + 1 == plan.npasses;
      pp.k = k;
      pp.dshift = L - l0 - k;
      if (q > 0) {
        // node-aligned tiles over the padded level-l0 order
        std::vector<uint64_t> tstart;
        std::vector<uint32_t> tlen, tnode, tfirst(nodes + 1);
        for (uint64_t P = 0; P < nodes; ++P) {
          tfirst[P] = static_cast<uint32_t>(tstart.size());
          for (uint64_t o = 0; o < node_size[P]; o += T) {
            tstart.push_back(node_pstart[P] + o);
            tlen.push_back(static_cast<uint32_t>(std::min<uint64_t>(T, node_size[P] - o)));
            tnode.push_back(static_cast<uint32_t>(P));
          }
        }
        tfirst[nodes] = static_cast<uint32_t>(tstart.size());
        const uint64_t ntiles = tstart.size();
        if (ntiles > max_tiles) throw Error(ECL_ERROR_INTERNAL, "tile estimate exceeded");
        d_tstart = DevBuf(std::max<uint64_t>(ntiles, 1) * 8);
        d_tlen = DevBuf(std::max<uint64_t>(ntiles, 1) * 4);
        d_tnode = DevBuf(std::max<uint64_t>(ntiles, 1) * 4);
        d_tfirst = DevBuf((nodes + 1) * 4);
        h2d(d_tstart.ptr, tstart.data(), ntiles);
        h2d(d_tlen.ptr, tlen.data(), ntiles);
        h2d(d_tnode.ptr, tnode.data(), ntiles);
        h2d(d_tfirst.ptr, tfirst.data(), nodes + 1);
        pp.src = in_codes;
        pp.src_esize = cw;
        pp.codes_out = nullptr;
        pp.codes = in_codes;
        pp.map_kind = 0;
        pp.num_tiles = ntiles;
        pp.tile_start = d_tstart.as<uint64_t>();
        pp.tile_len = d_tlen.as<uint32_t>();
        pp.tile_node = d_tnode.as<uint32_t>();
        pp.node_first_tile = d_tfirst.as<uint32_t>();
        launch_tile_hist(pp, s);
        launch_tile_scan(pp, s);
        // digit counts per node: differences of the scan rows at node starts
        d_rows = DevBuf((nodes + 1) * nb * 8);
        launch_gather_rows(d_excl, d_tfirst.as<uint32_t>(), nodes + 1, nb,
                           d_rows.as<uint64_t>(), s);
        std::vector<uint64_t> rows((nodes + 1) * nb);
        ECL_HIP(hipStreamSynchronize(s));
        d2h(rows.data(), d_rows.ptr, rows.size());
        binc.assign(nodes * nb, 0);
        for (uint64_t P = 0; P < nodes; ++P)
          for (uint32_t d = 0; d < nb; ++d)
            binc[P * nb + d] = rows[(P + 1) * nb + d] - rows[P * nb + d];
      }
      // first absent digit of every node at levels l0 .. l0+k (group aligned)
      std::vector<uint16_t> thr(nodes * (k + 1), static_cast<uint16_t>(nb));
      if (!wt->pow2) {
        std::vector<int32_t> maxd(nodes * (k + 1), -1);
        for (uint64_t c = 0; c < sigma; ++c) {
          if (len[c] <= l0) continue;
          const uint64_t P = detail::shr64(code[c], L - l0);
          const int32_t d = static_cast<int32_t>((code[c] >> (L - l0 - k)) & (nb - 1));
          for (uint32_t i = 0; i <= k; ++i)
            if (len[c] > l0 + i) maxd[P * (k + 1) + i] = std::max(maxd[P * (k + 1) + i], d);
        }
        for (uint64_t x = 0; x < nodes * (k + 1); ++x) {
          const uint32_t i = static_cast<uint32_t>(x % (k + 1));
          thr[x] = maxd[x] < 0 ? 0
                               : static_cast<uint16_t>(((static_cast<uint32_t>(maxd[x]) >> (k - i)) + 1)
                                                       << (k - i));
        }
      }
      if (single) {
        up.add(d_thr0, thr.data(), thr.size());
        pp.thr = d_thr0;
      } else {
        d_thr = DevBuf(thr.size() * 2);
        h2d(d_thr.ptr, thr.data(), thr.size());
        pp.thr = d_thr.as<uint16_t>();
      }
      // global start of (node, digit): codes of earlier nodes + earlier digits
      if (sizes_known) {
        pp.pc = d_pc0;  // scanned on the device
      } else {
        std::vector<uint64_t> pc(nodes * (nb + 1));
        uint64_t ns = 0;
        for (uint64_t P = 0; P < nodes; ++P) {
          uint64_t run = ns;
          for (uint32_t d = 0; d < nb; ++d) {
            pc[P * (nb + 1) + d] = run;
            run += binc[P * nb + d];
          }
          pc[P * (nb + 1) + nb] = run;
          ns += node_size[P];
        }
        if (single) {
          up.add(d_pc0, pc.data(), pc.size());
          pp.pc = d_pc0;
        } else {
          d_pc = DevBuf(pc.size() * 8);
          h2d(d_pc.ptr, pc.data(), pc.size());
          pp.pc = d_pc.as<uint64_t>();
        }
      }
      // next pass: node sizes and padded starts of level l0+k
      std::vector<uint64_t> next_size, next_pstart;
      void* out_buf = nullptr;
      if (!last) {
        next_size.assign(nodes * nb, 0);
        next_pstart.assign(nodes * nb, 0);
        uint64_t run = 0;
        for (uint64_t P = 0; P < nodes; ++P)
          for (uint32_t d = 0; d < nb; ++d) {
            const uint64_t Q = P * nb + d;
            next_size[Q] = d < thr[P * (k + 1) + k] ? binc[Q] : 0;
            next_pstart[Q] = run;
            run += (next_size[Q] + E - 1) / E * E;
          }
        if (run > pad_cap) throw Error(ECL_ERROR_INTERNAL, "level buffer too small");
        d_obase = DevBuf(next_pstart.size() * 8);
        h2d(d_obase.ptr, next_pstart.data(), next_pstart.size());
        // outputs alternate between two buffers that differ from the input
        out_buf = (q % 2 == 0) ? buf_u.ptr : (buf_x.ptr ? buf_x.ptr : buf_v.ptr);
        pp.obase = d_obase.as<uint64_t>();
      }
      pp.dst = out_buf;
      pp.dst_cap = out_buf ? pad_cap : 0;
      pp.words = ba->words();
      for (uint32_t i = 0; i < k; ++i) {
        pp.level_word[i] = ba->offsets[l0 + i];
        pp.level_nwords[i] = (level_size[l0 + i] + 31) / 32;
      }
      up.flush();
      if (sizes_known) {
        pp.tiles_per_wg = pass_span(pp.num_tiles);
        launch_span_words(pp, s);
      }
      launch_pass(pp, s);
      if (!last) {
        in_codes = out_buf;
        node_size.swap(next_size);
        node_pstart.swap(next_pstart);
      }
    }

    // ---- rank/select directory of every level, queued behind the passes;
    // then the one wait of the build
    build_directories(rs, ba, s, scr, up);
    ECL_HIP(hipStreamSynchronize(s));
    {
      uint32_t err = 0;
      d2h(&err, d_err, 1);
      if (err) throw Error(ECL_ERROR_INTERNAL, "construction pass wrote outside its arrays");
    }
    if (sizes_known) d2h(binc.data(), d_excl + tiles0 * nb0, nb0);
    if (sym_count.empty()) {  // power-of-2 alphabet: the last pass's digits end the codes
      const uint32_t kl = plan.k.back(), nbl = 1u << kl;
      sym_count.resize(sigma);
      uint64_t tot = 0;
      for (uint64_t c = 0; c < sigma; ++c) {
        sym_count[c] = binc[(code[c] >> kl) * nbl + (code[c] & (nbl - 1))];
        tot += sym_count[c];
      }
      if (tot != n) throw Error(ECL_ERROR_INTERNAL, "histogram does not sum to n");
    }
    own_text.reset();
    buf_x.reset();
    buf_u.reset();
    buf_v.reset();

    finish_tree(wt, shape, sym_count, ba_guard, rs_guard, up, s);
  } catch (...) {
    delete wt;
    throw;
  }
  return wt;
}

// Query lines for every array of a directory (an opt-in copy of the bits,
// 128/120 of their size, that lets a rank read one line) and the select
// hints (64 B per 4096 ones or zeros, that let a select read its hint and
// one block); the level views point to them afterwards.
static void build_query_lines(ecl_rank_select_t* rs) {
  if (rs->lines.ptr != nullptr) return;
  const size_t na = rs->levels.size();
  std::vector<uint64_t> off(na), nl(na), hoff(na), nh1(na), nh0(na);
  uint64_t total = 0, htotal = 0;
  for (size_t i = 0; i < na; ++i) {
    const LevelView& v = rs->levels[i];
    nl[i] = (v.size + detail::kLineBits - 1) / detail::kLineBits;
    off[i] = total;
    total += nl[i];
    // one select hint per sample bracket of each Value (none for an empty array)
    nh1[i] = v.size ? v.ones / detail::kSampleRate + 1 : 0;
    nh0[i] = v.size ? (v.size - v.ones) / detail::kSampleRate + 1 : 0;
    hoff[i] = htotal;
    htotal += nh1[i] + nh0[i];
  }
  // lines [total x 128 B] | select hints [htotal x 64 B]; scratch for the
  // hints' fit counts (2 u32 per hint workgroup of the largest array)
  const uint64_t bytes = total * 128 + htotal * 64;
  uint64_t fit_words = 0;
  for (size_t i = 0; i < na; ++i)
    fit_words = std::max<uint64_t>(fit_words, 2 * select_hint_blocks(nh1[i], nh0[i]));
  if (bytes + fit_words * 4 + (64u << 20) > free_device_memory())
    throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
  DevBuf lines(std::max<uint64_t>(bytes, 1));
  DevBuf fit(std::max<uint64_t>(fit_words, 1) * 4);
  uint4* L = lines.as<uint4>();
  uint4* H = L + total * 8;
  // the views point to the lines and hints first; each array's hint kernels
  // then set its hint_values in the device view
  std::vector<LevelView> lv = rs->levels;
  for (size_t i = 0; i < na; ++i) {
    lv[i].lines = nl[i] ? L + off[i] * 8 : nullptr;
    lv[i].sel_hints = nh1[i] + nh0[i] ? H + hoff[i] * 4 : nullptr;
    lv[i].hint_values = 0;
  }
  hipStream_t s = nullptr;
  ECL_HIP(hipMemcpyAsync(rs->d_levels.ptr, lv.data(), na * sizeof(LevelView),
                         hipMemcpyHostToDevice, s));
  for (size_t i = 0; i < na; ++i) {
    launch_query_lines(rs->levels[i], L + off[i] * 8, nl[i], s);
    // the fit scratch is reused array after array (same stream, in order)
    launch_select_hints(rs->levels[i], H + hoff[i] * 4, nh1[i], nh0[i], fit.as<uint32_t>(),
                        rs->d_levels.as<LevelView>() + i, s);
  }
  // the host copies of the views get the flags the device computed
  std::vector<LevelView> dv(na);
  if (na) ECL_HIP(hipMemcpyAsync(dv.data(), rs->d_levels.ptr, na * sizeof(LevelView),
                                 hipMemcpyDeviceToHost, s));
  ECL_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < na; ++i) lv[i].hint_values = dv[i].hint_values;
  rs->levels = std::move(lv);
  rs->lines = std::move(lines);
}

// Quad lines of a tree (QuadView): for quad level q, from the symbol
// counts, per node N of level 2q (4^q nodes, N = the codes' top 2q bits) and
// digit d: cnt[N][d] = the symbols with top 2q + 2 bits 4N + d; S_N = C of
// the node's first symbol; CD[N][d] = the d's before the node (exclusive
// prefix of cnt over N); base[4N + d] = C of child 4N + d minus CD[N][d].
// The lines themselves come from the binary levels 2q and 2q + 1 on the
// device.
// the device copies of the tree's quad and hex views, when they exist
static void sync_layout_views(ecl_wavelet_tree_t* wt) {
  if (wt->d_layouts.ptr == nullptr) return;
  h2d(wt->d_layouts.ptr, &wt->quad, 1);
  h2d(static_cast<uint8_t*>(wt->d_layouts.ptr) + sizeof(QuadView), &wt->hex, 1);
}

static void build_quad_lines(ecl_wavelet_tree_t* wt) {
  if (wt->quad.levels != 0) return;
  const uint32_t L = wt->L;
  if (!wt->pow2 || L % 2 != 0 || L < 2 || L > 2 * kMaxQuadLevels)
    throw Error(ECL_ERROR_INVALID_ARGUMENT,
                "quad lines need a power-of-two alphabet with an even number of levels "
                "(at most 16)");
  const uint32_t Q = L / 2;
  const uint64_t n = wt->n, sigma = wt->sigma;
  // C[c] for c in [0, sigma], C[sigma] = n
  std::vector<uint64_t> C(sigma + 1);
  for (uint64_t c = 0; c < sigma; ++c) C[c] = wt->counts[c];
  C[sigma] = n;
  // host tables: bases of all levels; per level S and CD for the build
  std::vector<uint64_t> base, build_tab;
  std::vector<uint64_t> off(Q), s_off(Q), cd_off(Q), nodes(Q);
  for (uint32_t q = 0; q < Q; ++q) {
    const uint64_t nn = 1ull << (2 * q);  // nodes of level 2q
    const uint32_t shift = L - 2 * q - 2;  // child 4N + d spans symbols [(4N+d) << shift, ...)
    nodes[q] = nn;
    off[q] = base.size();
    s_off[q] = build_tab.size();
    for (uint64_t N = 0; N <= nn; ++N) build_tab.push_back(C[std::min<uint64_t>(N << (shift + 2), sigma)]);
    cd_off[q] = build_tab.size();
    uint64_t acc[4] = {0, 0, 0, 0};
    for (uint64_t N = 0; N <= nn; ++N) {
      for (uint32_t d = 0; d < 4; ++d) build_tab.push_back(acc[d]);  // CD[N][d]
      if (N == nn) break;
      for (uint32_t d = 0; d < 4; ++d) {
        const uint64_t lo = (4 * N + d) << shift, hi = (4 * N + d + 1) << shift;
        const uint64_t cnt = C[hi] - C[lo];
        base.push_back(C[lo] - acc[d]);  // base[4N + d]
        acc[d] += cnt;
      }
    }
  }
  // lines: n / 448 + 1 per level
  const uint64_t nlines = n / kQuadDigits + 1;
  const uint64_t bytes = Q * nlines * 128;
  const uint64_t tab_bytes = (base.size() + build_tab.size()) * 8;
  if (bytes + tab_bytes + (64u << 20) > free_device_memory())
    throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
  DevBuf lines(bytes);
  DevBuf dbase(base.size() * 8);
  DevBuf dtab(build_tab.size() * 8);
  DevBuf wg_info(quad_build_workgroups(nlines) * kQuadWgInfo * 8);  // reused level by level
  // the 4-position deposit table (QuadBuildParams::lut)
  std::vector<uint8_t> lut(4096);
  for (uint32_t x = 0; x < 4096; ++x) {
    const uint32_t a = x & 15, z = (x >> 4) & 15, o = x >> 8;
    uint32_t iz = 0, io = 0, v = 0;
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t ai = (a >> i) & 1, bi = ai ? (o >> io++) & 1 : (z >> iz++) & 1;
      v |= ((ai << 1) | bi) << (2 * i);
    }
    lut[x] = static_cast<uint8_t>(v);
  }
  DevBuf dlut(4096);
  h2d(dlut.ptr, lut.data(), 4096);
  h2d(dbase.ptr, base.data(), base.size());
  h2d(dtab.ptr, build_tab.data(), build_tab.size());
  hipStream_t s = nullptr;
  QuadView qv{};
  for (uint32_t q = 0; q < Q; ++q) {
    QuadBuildParams qp{};
    qp.a = wt->rs->levels[2 * q];
    qp.b = wt->rs->levels[2 * q + 1];
    qp.S = dtab.as<uint64_t>() + s_off[q];
    qp.CD = dtab.as<uint64_t>() + cd_off[q];
    qp.nodes = nodes[q];
    qp.n = n;
    qp.nlines = nlines;
    qp.out = lines.as<uint4>() + q * nlines * 8;
    qp.wg_info = wg_info.as<uint64_t>();
    qp.lut = dlut.as<uint8_t>();
    launch_quad_lines(qp, s);
    qv.lines[q] = qp.out;
    qv.off[q] = off[q];
  }
  ECL_HIP(hipStreamSynchronize(s));
  qv.base = dbase.as<uint64_t>();
  qv.nbase = base.size();
  qv.levels = Q;
  qv.tab = dtab.as<uint64_t>();
  qv.ntab = build_tab.size();
  for (uint32_t q = 0; q < Q; ++q) {
    qv.s_off[q] = s_off[q];
    qv.cd_off[q] = cd_off[q];
  }
  wt->quad_lines = std::move(lines);
  wt->quad_base = std::move(dbase);
  wt->quad_tab = std::move(dtab);
  wt->quad = qv;
  sync_layout_views(wt);
}

// Quad select lines (QuadView): for each quad level and digit d with t
// occurrences, a line every r = max(1, floor(340 t / n)) of them (448
// digits hold the next r with a margin of about 2.4 standard deviations at
// density 1/4), built from the level's quad lines.  Builds the quad lines
// first when the tree has none.
static void build_quad_select_lines(ecl_wavelet_tree_t* wt) {
  build_quad_lines(wt);
  QuadView& qv = wt->quad;
  if (qv.has_sel) return;
  const uint32_t Q = qv.levels;
  const uint64_t n = wt->n;
  const uint64_t nql = n / kQuadDigits + 1;
  // totals of every level's digits: the last row of its CD table
  std::vector<uint64_t> tab(qv.ntab);
  d2h(tab.data(), qv.tab, qv.ntab);
  uint64_t total_lines = 0;
  uint64_t first[kMaxQuadLevels][4] = {};
  uint64_t cnt[kMaxQuadLevels][4] = {};
  uint32_t rate[kMaxQuadLevels][4] = {};
  uint64_t level_first[kMaxQuadLevels] = {};
  for (uint32_t q = 0; q < Q; ++q) {
    const uint64_t nodes = 1ull << (2 * q);
    level_first[q] = total_lines;
    uint64_t run = 0;
    for (uint32_t d = 0; d < 4; ++d) {
      const uint64_t t = tab[qv.cd_off[q] + 4 * nodes + d];
      cnt[q][d] = t;
      const uint64_t r = t ? std::max<uint64_t>(1, kQuadSelSpan * t / n) : 0;
      rate[q][d] = static_cast<uint32_t>(r);
      first[q][d] = run;
      run += r ? t / r + 1 : 0;
    }
    total_lines += run;
  }
  if (total_lines * 128 + (64u << 20) > free_device_memory())
    throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
  DevBuf lines(std::max<uint64_t>(total_lines, 1) * 128);
  hipStream_t s = nullptr;
  for (uint32_t q = 0; q < Q; ++q) {
    uint4* level_lines = lines.as<uint4>() + level_first[q] * 8;
    QuadSelBuildParams sp{};
    sp.qlines = qv.lines[q];
    sp.nql = nql;
    for (uint32_t d = 0; d < 4; ++d) {
      sp.total[d] = cnt[q][d];
      sp.r[d] = rate[q][d];
      sp.magic[d] = rate[q][d] > 1 ? ~0ull / rate[q][d] + 1 : 0;
      sp.nlines[d] = rate[q][d] ? cnt[q][d] / rate[q][d] + 1 : 0;
      sp.out[d] = level_lines + first[q][d] * 8;
    }
    launch_quad_select_lines(sp, s);
    qv.sel_lines[q] = level_lines;
    for (uint32_t d = 0; d < 4; ++d) {
      qv.sel_first[q][d] = first[q][d];
      qv.sel_rate[q][d] = rate[q][d];
      qv.sel_magic[q][d] = rate[q][d] > 1 ? ~0ull / rate[q][d] + 1 : 0;
    }
  }
  ECL_HIP(hipStreamSynchronize(s));
  wt->quad_sel = std::move(lines);
  qv.has_sel = 1;
  sync_layout_views(wt);
}

// Hex lines of a tree (HexView): for hex level h, from the symbol counts,
// per node N of level 4h (16^h nodes, N = the codes' top 4h bits) and hex
// digit x: cnt[N][x] = the symbols with top 4h + 4 bits 16N + x; CHD[N][x] =
// the x's before the node (exclusive prefix of cnt over N); base[16N + x] =
// C of child 16N + x minus CHD[N][x].  The lines come from the quad lines of
// quad levels 2h and 2h + 1 on the device, which build first; the kernel
// also writes the superblock counts.
static void build_hex_lines(ecl_wavelet_tree_t* wt) {
  if (wt->hex.levels != 0) return;
  const uint32_t L = wt->L;
  if (!wt->pow2 || L % 4 != 0 || L < 4 || L > 4 * kMaxHexLevels)
    throw Error(ECL_ERROR_INVALID_ARGUMENT,
                "hex lines need a power-of-two alphabet whose level count is a multiple of 4 "
                "(at most 16)");
  build_quad_lines(wt);
  const uint32_t H = L / 4;
  const uint64_t n = wt->n, sigma = wt->sigma;
  std::vector<uint64_t> C(sigma + 1);
  for (uint64_t c = 0; c < sigma; ++c) C[c] = wt->counts[c];
  C[sigma] = n;
  const uint64_t nlines = n / kHexDigits + 1;
  const uint64_t nsb = (nlines >> kHexSbShift) + 1;
  // tab (kept): bases, then superblocks; btab (build only): CHD, node starts
  std::vector<uint64_t> tab, btab;
  uint64_t base_off[kMaxHexLevels] = {}, sb_off[kMaxHexLevels] = {};
  uint64_t chd_off[kMaxHexLevels] = {}, s_off[kMaxHexLevels] = {}, nodes[kMaxHexLevels] = {};
  for (uint32_t h = 0; h < H; ++h) {
    const uint64_t nn = 1ull << (4 * h);
    const uint32_t shift = L - 4 * h - 4;  // child 16N + x spans symbols [(16N+x) << shift, ...)
    nodes[h] = nn;
    base_off[h] = tab.size();
    chd_off[h] = btab.size();
    uint64_t acc[16] = {};
    for (uint64_t N = 0; N <= nn; ++N) {
      for (uint32_t x = 0; x < 16; ++x) btab.push_back(acc[x]);  // CHD[N][x]
      if (N == nn) break;
      for (uint32_t x = 0; x < 16; ++x) {
        const uint64_t lo = (16 * N + x) << shift, hi = (16 * N + x + 1) << shift;
        tab.push_back(C[lo] - acc[x]);  // base[16N + x]
        acc[x] += C[hi] - C[lo];
      }
    }
    s_off[h] = btab.size();
    for (uint64_t N = 0; N <= nn; ++N) btab.push_back(C[std::min<uint64_t>(N << (shift + 4), sigma)]);
  }
  for (uint32_t h = 0; h < H; ++h) {
    sb_off[h] = tab.size();
    tab.resize(tab.size() + 16 * nsb, 0);  // written by the build kernel
  }
  const uint64_t bytes = H * nlines * 128;
  if (bytes + (tab.size() + btab.size()) * 8 + (64u << 20) > free_device_memory())
    throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
  DevBuf lines(bytes);
  DevBuf dtab(tab.size() * 8);
  DevBuf dbtab(btab.size() * 8);
  h2d(dtab.ptr, tab.data(), tab.size());
  h2d(dbtab.ptr, btab.data(), btab.size());
  const QuadView& qv = wt->quad;
  hipStream_t s = nullptr;
  HexView hv{};
  for (uint32_t h = 0; h < H; ++h) {
    HexBuildParams hp{};
    hp.qa = qv.lines[2 * h];
    hp.qb = qv.lines[2 * h + 1];
    hp.qbase = qv.base + qv.off[2 * h];
    hp.qcd = qv.tab + qv.cd_off[2 * h + 1];
    hp.chd = dbtab.as<uint64_t>() + chd_off[h];
    hp.S = dbtab.as<uint64_t>() + s_off[h];
    hp.nodes = nodes[h];
    hp.n = n;
    hp.nql = n / kQuadDigits + 1;
    hp.nlines = nlines;
    hp.out = lines.as<uint4>() + h * nlines * 8;
    hp.sb = dtab.as<uint64_t>() + sb_off[h];
    launch_hex_lines(hp, s);
    hv.lines[h] = hp.out;
    hv.base_off[h] = base_off[h];
    hv.sb_off[h] = sb_off[h];
  }
  ECL_HIP(hipStreamSynchronize(s));
  hv.tab = dtab.as<uint64_t>();
  hv.ntab = tab.size();
  hv.levels = H;
  wt->hex_lines = std::move(lines);
  wt->hex_tab = std::move(dtab);
  wt->hex = hv;
  sync_layout_views(wt);
}

// Select lines for every array of a directory (opt-in): for each Value a
// line every r Values, r = max(1, floor(800 x density)) so that the next r
// Values lie in the line's 960 bits; the level views point to them
// afterwards.  About 0.32 bytes per bit of every array.
static void build_select_lines(ecl_rank_select_t* rs) {
  if (rs->sel_lines.ptr != nullptr) return;
  const size_t na = rs->levels.size();
  std::vector<uint64_t> off(na), n1(na), n0(na);
  std::vector<uint32_t> r1(na), r0(na);
  uint64_t total = 0;
  for (size_t i = 0; i < na; ++i) {
    const LevelView& v = rs->levels[i];
    const uint64_t zeros = v.size - v.ones;
    r1[i] = v.size ? select_line_rate(v.ones, v.size) : 0;
    r0[i] = v.size ? select_line_rate(zeros, v.size) : 0;
    n1[i] = r1[i] ? v.ones / r1[i] + 1 : 0;
    n0[i] = r0[i] ? zeros / r0[i] + 1 : 0;
    off[i] = total;
    total += n1[i] + n0[i];
  }
  if (total * 128 + (64u << 20) > free_device_memory())
    throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
  DevBuf lines(std::max<uint64_t>(total, 1) * 128);
  uint4* L = lines.as<uint4>();
  hipStream_t s = nullptr;
  for (size_t i = 0; i < na; ++i)
    launch_select_lines(rs->levels[i], L + off[i] * 8, n1[i], n0[i], r1[i], r0[i], s);
  ECL_HIP(hipStreamSynchronize(s));
  std::vector<LevelView> lv = rs->levels;
  for (size_t i = 0; i < na; ++i) {
    if (n1[i] + n0[i] == 0) continue;
    lv[i].sel_lines = L + off[i] * 8;
    lv[i].sel_zero_line = n1[i];
    lv[i].sel_rate[1] = r1[i];
    lv[i].sel_rate[0] = r0[i];
    lv[i].sel_magic[1] = select_line_magic(r1[i]);
    lv[i].sel_magic[0] = select_line_magic(r0[i]);
  }
  h2d(rs->d_levels.ptr, lv.data(), na);
  rs->levels = std::move(lv);
  rs->sel_lines = std::move(lines);
}

// ===========================================================================
// persistence: header, alphabet, symbol counts, level bits (the directory and
// the metadata are rebuilt on load)
// ===========================================================================
namespace {
struct TreeFileHeader {
  char magic[8];
  uint32_t version;
  uint32_t esize;
  uint64_t n;
  uint64_t sigma;
  uint32_t num_levels;
  uint32_t reserved;
};
constexpr char kTreeMagic[8] = {'E', 'C', 'L', 'W', 'T', 'R', 'E', 'E'};
constexpr uint32_t kTreeFileVersion = 1;
constexpr size_t kFileChunk = 64u << 20;

struct File {
  std::FILE* f = nullptr;
  File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() { if (f) std::fclose(f); }
};
void write_all(std::FILE* f, const void* p, size_t bytes) {
  if (bytes && std::fwrite(p, 1, bytes, f) != bytes)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "writing the tree file failed");
}
void read_all(std::FILE* f, void* p, size_t bytes) {
  if (bytes && std::fread(p, 1, bytes, f) != bytes)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "the tree file is truncated");
}
}  // namespace

namespace ecl {
namespace lib {
void read_tree_file_header(const char* path, int* esize, uint64_t* n, uint64_t* sigma) {
  require(path != nullptr, "path is NULL");
  File file(path, "rb");
  if (!file.f) throw Error(ECL_ERROR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  TreeFileHeader h{};
  read_all(file.f, &h, sizeof(h));
  if (std::memcmp(h.magic, kTreeMagic, 8) != 0 || h.version != kTreeFileVersion)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "not a wavelet-tree file of this version");
  *esize = static_cast<int>(h.esize);
  *n = h.n;
  *sigma = h.sigma;
}
}  // namespace lib
}  // namespace ecl

static void save_tree(const ecl_wavelet_tree_t* wt, const char* path) {
  require(path != nullptr, "path is NULL");
  set_device(wt->device);
  File file(path, "wb");
  if (!file.f) throw Error(ECL_ERROR_INVALID_ARGUMENT, std::string("cannot create ") + path);
  TreeFileHeader h{};
  std::memcpy(h.magic, kTreeMagic, 8);
  h.version = kTreeFileVersion;
  h.esize = static_cast<uint32_t>(wt->esize);
  h.n = wt->n;
  h.sigma = wt->sigma;
  h.num_levels = wt->L;
  write_all(file.f, &h, sizeof(h));
  std::vector<uint8_t> raw(wt->sigma * wt->esize);
  u64_to_elems(wt->alphabet, raw.data(), wt->esize);
  write_all(file.f, raw.data(), raw.size());
  std::vector<uint64_t> cnt(wt->sigma);
  for (uint64_t c = 0; c < wt->sigma; ++c)
    cnt[c] = (c + 1 < wt->sigma ? wt->counts[c + 1] : wt->n) - wt->counts[c];
  write_all(file.f, cnt.data(), cnt.size() * 8);
  const ecl_bit_array_t* ba = wt->rs->ba;
  std::vector<uint32_t> buf(kFileChunk / 4);
  for (uint32_t l = 0; l < wt->L; ++l) {
    const uint64_t m = ba->sizes[l], nw = (m + 31) / 32;
    write_all(file.f, &m, 8);
    for (uint64_t w = 0; w < nw; w += buf.size()) {
      const uint64_t c = std::min<uint64_t>(buf.size(), nw - w);
      d2h(buf.data(), ba->words() + ba->offsets[l] + w, c);
      write_all(file.f, buf.data(), c * 4);
    }
  }
  if (std::fflush(file.f) != 0) throw Error(ECL_ERROR_INVALID_ARGUMENT, "writing the tree file failed");
}

static ecl_wavelet_tree_t* load_tree(const char* path, int device) {
  require(path != nullptr, "path is NULL");
  File file(path, "rb");
  if (!file.f) throw Error(ECL_ERROR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  TreeFileHeader h{};
  read_all(file.f, &h, sizeof(h));
  if (std::memcmp(h.magic, kTreeMagic, 8) != 0 || h.version != kTreeFileVersion)
    throw Error(ECL_ERROR_INVALID_ARGUMENT, "not a wavelet-tree file of this version");
  require(h.esize == 1 || h.esize == 2 || h.esize == 4 || h.esize == 8, "invalid element size");
  require(h.n > 0 && h.sigma >= 2 && h.sigma <= h.n + (1ull << 32), "invalid tree sizes");
  set_device(device);
  hipStream_t s = nullptr;
  auto* wt = new ecl_wavelet_tree_t();
  try {
    wt->device = device;
    wt->esize = static_cast<int>(h.esize);
    wt->n = h.n;
    std::vector<uint8_t> raw(h.sigma * h.esize);
    read_all(file.f, raw.data(), raw.size());
    wt->alphabet = elems_to_u64(raw.data(), h.sigma, wt->esize);
    require(std::is_sorted(wt->alphabet.begin(), wt->alphabet.end()) &&
                std::adjacent_find(wt->alphabet.begin(), wt->alphabet.end()) ==
                    wt->alphabet.end(),
            "the file's alphabet is not sorted and distinct");
    const TreeShape shape = set_alphabet(wt);
    require(wt->L == h.num_levels, "the file's level count does not match its alphabet");
    std::vector<uint64_t> sym_count(h.sigma);
    read_all(file.f, sym_count.data(), h.sigma * 8);
    uint64_t tot = 0;
    for (uint64_t c : sym_count) tot += c;
    require(tot == h.n, "the file's symbol counts do not sum to its size");
    const std::vector<uint64_t> level_size = level_sizes(shape, sym_count);
    if (tree_needed_bytes(h.n, h.sigma, wt->L, wt->esize, false) >
        free_device_memory() + cached_scratch_bytes(device))
      throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, kOomTree);
    Uploader up(s);
    upload_symbol_maps(wt, up);
    std::unique_ptr<ecl_bit_array_t> ba(make_bit_array(level_size, -1, device, &up));
    up.flush();
    // Level bits are checked against the counts while they stream in: the
    // internal nodes of level l tile its array in symbol order (node lo..hi
    // at positions [C(lo), C(hi + 1))), and a node's ones are the codes of
    // its right child, C(hi + 1) - C(mid).  Queries rely on that: a descent
    // stays inside its node only if every node holds exactly these ones.
    std::vector<uint64_t> prefix(h.sigma + 1, 0);
    for (uint64_t c = 0; c < h.sigma; ++c) prefix[c + 1] = prefix[c] + sym_count[c];
    std::vector<uint32_t> buf(kFileChunk / 4);
    for (uint32_t l = 0; l < wt->L; ++l) {
      uint64_t m = 0;
      read_all(file.f, &m, 8);
      require(m == level_size[l], "the file's level sizes do not match its symbol counts");
      const std::vector<Node>& nodes = shape.nodes[l];
      std::vector<uint64_t> bound;  // node starts, then m
      for (const Node& nd : nodes) bound.push_back(prefix[nd.lo]);
      bound.push_back(m);
      require(bound.front() == 0, "the file's level sizes do not match its symbol counts");
      std::vector<uint64_t> ones_at(bound.size(), 0);  // ones before each bound
      size_t bi = 0;
      uint64_t ones = 0;
      const uint64_t nw = (m + 31) / 32;
      for (uint64_t w = 0; w < nw; w += buf.size()) {
        const uint64_t c = std::min<uint64_t>(buf.size(), nw - w);
        read_all(file.f, buf.data(), c * 4);
        // bits past m in the last word are not part of the array
        if (w + c == nw && (m & 31)) buf[c - 1] &= (1u << (m & 31)) - 1u;
        for (uint64_t k = 0; k < c; ++k) {
          const uint64_t word_bit = (w + k) * 32;
          while (bi < bound.size() && bound[bi] < word_bit + 32) {
            const uint32_t b = static_cast<uint32_t>(bound[bi] - word_bit);
            ones_at[bi++] = ones + __builtin_popcount(b ? buf[k] & ((1u << b) - 1u) : 0u);
          }
          ones += __builtin_popcount(buf[k]);
        }
        h2d(ba->words() + ba->offsets[l] + w, buf.data(), c);
      }
      while (bi < bound.size()) ones_at[bi++] = ones;  // bound == m at a word edge
      for (size_t j = 0; j < nodes.size(); ++j) {
        const Node& nd = nodes[j];
        const uint64_t mid = nd.lo + detail::prev_pow2(nd.hi - nd.lo + 1);
        require(bound[j + 1] == prefix[nd.hi + 1] &&
                    ones_at[j + 1] - ones_at[j] == prefix[nd.hi + 1] - prefix[mid],
                "the file's level bits do not match its symbol counts");
      }
    }
    if (std::fgetc(file.f) != EOF)
      throw Error(ECL_ERROR_INVALID_ARGUMENT, "the tree file has trailing bytes");
    std::unique_ptr<ecl_rank_select_t> rs(alloc_rank_select(ba.get(), s));
    {
      Scratch scr(device, s, directory_scratch_bytes(level_size));
      build_directories(rs.get(), ba.get(), s, scr, up);
      ECL_HIP(hipStreamSynchronize(s));
    }
    finish_tree(wt, shape, sym_count, ba, rs, up, s);
  } catch (...) {
    delete wt;
    throw;
  }
  return wt;
}

// ===========================================================================
// query helpers
// ===========================================================================
static TreeQueryParams query_params(const ecl_wavelet_tree_t* wt, uint32_t flags) {
  TreeQueryParams p{};
  p.view = wt->view;
  p.esize = wt->esize;
  p.symbols_minimal = (flags & ECL_SYMBOLS_MINIMAL) ? 1 : 0;
  p.map_kind = wt->map_kind;
  p.lut = wt->d_lut.as<uint32_t>();
  p.alphabet = wt->d_alphabet.ptr;
  p.alphabet_out = (!wt->min_alphabet && !(flags & ECL_RESULTS_MINIMAL)) ? wt->d_alphabet.ptr
                                                                         : nullptr;
  p.has_lines = wt->rs->lines.ptr != nullptr ? 1 : 0;
  p.has_sel_lines = wt->rs->sel_lines.ptr != nullptr ? 1 : 0;
  p.lanes = wt->lanes;
  p.quad = wt->quad;
  p.hex = wt->hex;
  return p;
}

namespace ecl {
namespace lib {
TreeQueryParams tree_query_params(ecl_wavelet_tree wt) {
  return query_params(wt, ECL_SYMBOLS_MINIMAL | ECL_RESULTS_MINIMAL);
}
}  // namespace lib
}  // namespace ecl

static const uint64_t kHostChunk = 1ull << 24;

static void check_query_memory(uint64_t bytes, const char* kind) {
  if (bytes > free_device_memory())
    throw Error(ECL_ERROR_OUT_OF_MEMORY_QUERIES,
                std::string("Not enough GPU memory available for the ") + kind + " queries.");
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* ecl_last_error(void) { return g_last_error.c_str(); }
int ecl_abi_version(void) { return ECL_WAVELET_ABI_VERSION; }

ecl_status ecl_device_count(int* count) {
  return guard([&] {
    require(count != nullptr, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
      (void)hipGetLastError();
      n = 0;
    }
    *count = n;
  });
}
ecl_status ecl_device_alloc(int device, size_t bytes, void** ptr) {
  return guard([&] {
    require(ptr != nullptr, "ptr is NULL");
    set_device(device);
    *ptr = nullptr;
    if (!bytes) return;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
      // the library's cached blocks go back to the runtime, then once more
      (void)hipGetLastError();
      release_block_pool(device);
      ECL_HIP(hipMalloc(ptr, bytes));
    }
  });
}
ecl_status ecl_device_free(int device, void* ptr) {
  return guard([&] {
    set_device(device);
    if (ptr) ECL_HIP(hipFree(ptr));
  });
}
ecl_status ecl_memcpy_h2d(int device, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    set_device(device);
    if (bytes) ECL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}
ecl_status ecl_memcpy_d2h(int device, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    set_device(device);
    if (bytes) ECL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}
ecl_status ecl_device_synchronize(int device) {
  return guard([&] {
    set_device(device);
    ECL_HIP(hipDeviceSynchronize());
  });
}
ecl_status ecl_set_build_span(uint32_t tiles_per_workgroup) {
  return guard([&] {
    require(tiles_per_workgroup <= 64, "at most 64 tiles per workgroup");
    set_pass_tiles_per_wg(tiles_per_workgroup);
  });
}
ecl_status ecl_set_query_lanes(uint32_t lanes) {
  return guard([&] {
    require(lanes <= 2, "lanes must be 0 (automatic), 1 or 2");
    set_query_lanes(lanes);
  });
}
ecl_status ecl_set_device_memory_cap(int device, uint64_t bytes) {
  return guard([&] {
    require(device >= 0 && device < kCapDevices, "invalid device");
    g_mem_cap[device].store(bytes);
  });
}
ecl_status ecl_wavelet_tree_build_memory(ecl_type type, size_t n, size_t sigma,
                                         int host_text, uint64_t* resident,
                                         uint64_t* staging) {
  return guard([&] {
    require(resident != nullptr && staging != nullptr, "NULL output");
    const int esize = static_cast<int>(type);
    require(esize == 1 || esize == 2 || esize == 4 || esize == 8, "invalid element type");
    require(sigma >= 2, "the alphabet needs at least 2 symbols");
    const uint32_t L = detail::ceil_log2(sigma);
    // an upper bound: the text is taken not to be its own code
    const TreeMemory m = tree_memory(n, sigma, L, esize, host_text != 0, false);
    *resident = m.resident;
    *staging = m.staging();
  });
}
ecl_status ecl_wavelet_tree_build_host_bytes(ecl_wavelet_tree wt, uint64_t* bytes) {
  return guard([&] {
    require(wt != nullptr && bytes != nullptr, "NULL argument");
    *bytes = wt->build_host_bytes;
  });
}
ecl_status ecl_release_cached_memory(int device) {
  return guard([&] {
    set_device(device);
    release_scratch_cache(device);
    release_block_pool(device);
  });
}
ecl_status ecl_device_mem_info(int device, size_t* fr, size_t* tot) {
  return guard([&] {
    require(fr != nullptr && tot != nullptr, "NULL output");
    set_device(device);
    ECL_HIP(hipMemGetInfo(fr, tot));
    *fr += pooled_bytes(device);
  });
}
ecl_status ecl_device_name(int device, char* buf, size_t len) {
  return guard([&] {
    set_device(device);
    hipDeviceProp_t prop;
    ECL_HIP(hipGetDeviceProperties(&prop, device));
    std::string nm = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    if (len) {
      std::strncpy(buf, nm.c_str(), len - 1);
      buf[len - 1] = 0;
    }
  });
}
ecl_status ecl_event_create(void** ev) {
  return guard([&] {
    hipEvent_t e;
    ECL_HIP(hipEventCreate(&e));
    *ev = e;
  });
}
ecl_status ecl_event_destroy(void* ev) {
  return guard([&] { ECL_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev))); });
}
ecl_status ecl_event_record(void* ev, void* stream) {
  return guard([&] { ECL_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), as_stream(stream))); });
}
ecl_status ecl_event_elapsed(void* a, void* b, float* ms) {
  return guard([&] {
    ECL_HIP(hipEventSynchronize(static_cast<hipEvent_t>(b)));
    ECL_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
  });
}
ecl_status ecl_profile_enable(int on) {
  return guard([&] {
    if (!on) prof_drain();
    g_prof_on = on != 0;
  });
}
ecl_status ecl_profile_reset(void) {
  return guard([&] {
    prof_drain();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (int i = 0; i < K_COUNT; ++i) { g_prof_count[i] = 0; g_prof_ms[i] = 0; }
  });
}
ecl_status ecl_profile_get(const char* kernel, uint64_t* launches, double* total_ms) {
  return guard([&] {
    prof_drain();
    require(kernel != nullptr, "kernel is NULL");
    for (int i = 0; i < K_COUNT; ++i) {
      if (std::strcmp(kernel, kKernelNames[i]) == 0) {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        *launches = g_prof_count[i];
        *total_ms = g_prof_ms[i];
        return;
      }
    }
    throw Error(ECL_ERROR_INVALID_ARGUMENT, std::string("unknown kernel ") + kernel);
  });
}
ecl_status ecl_generate_text(int device, ecl_type type, uint64_t seed, size_t n, uint64_t sigma,
                             const uint64_t* thresholds, void* d_out) {
  return guard([&] {
    require(sigma >= 1, "sigma must be positive");
    set_device(device);
    DevBuf thr;
    if (thresholds) {
      thr = DevBuf(sigma * 8);
      h2d(thr.ptr, thresholds, sigma);
    }
    launch_gen_text(d_out, static_cast<int>(type), n, seed, sigma, thr.as<uint64_t>(), nullptr);
    ECL_HIP(hipDeviceSynchronize());
  });
}

// ---- BitArray ---------------------------------------------------------------
ecl_status ecl_bit_array_create(const size_t* sizes, size_t num_arrays, int init_value,
                                int device, ecl_bit_array* out) {
  return guard([&] {
    require(out != nullptr, "out is NULL");
    require(num_arrays == 0 || sizes != nullptr, "sizes is NULL");
    set_device(device);
    std::vector<uint64_t> sz(sizes, sizes + num_arrays);
    *out = make_bit_array(sz, init_value, device);
    ECL_HIP(hipDeviceSynchronize());
  });
}
ecl_status ecl_bit_array_destroy(ecl_bit_array ba) {
  return guard([&] {
    if (ba) {
      set_device(ba->device);
      delete ba;
    }
  });
}
ecl_status ecl_bit_array_num_arrays(ecl_bit_array ba, size_t* out) {
  return guard([&] {
    require(ba && out, "NULL argument");
    *out = ba->sizes.size();
  });
}
ecl_status ecl_bit_array_size(ecl_bit_array ba, size_t i, size_t* bits) {
  return guard([&] {
    require(ba && bits && i < ba->sizes.size(), "invalid argument");
    *bits = ba->sizes[i];
  });
}
ecl_status ecl_bit_array_offset(ecl_bit_array ba, size_t i, size_t* off) {
  return guard([&] {
    require(ba && off && i < ba->sizes.size(), "invalid argument");
    *off = ba->offsets[i];
  });
}
ecl_status ecl_bit_array_device_pointers(ecl_bit_array ba, uint32_t** data, uint64_t** offsets,
                                         uint64_t** sizes) {
  return guard([&] {
    require(ba != nullptr, "ba is NULL");
    if (data) *data = ba->words();
    if (offsets) *offsets = ba->d_offsets.as<uint64_t>();
    if (sizes) *sizes = ba->d_sizes.as<uint64_t>();
  });
}
ecl_status ecl_bit_array_write_words(ecl_bit_array ba, size_t i, size_t first,
                                     const uint32_t* words, size_t num) {
  return guard([&] {
    require(ba && i < ba->sizes.size(), "invalid argument");
    const uint64_t nw = (ba->sizes[i] + 31) / 32;
    require(first + num <= nw, "word range outside the array");
    set_device(ba->device);
    h2d(ba->words() + ba->offsets[i] + first, words, num);
  });
}
ecl_status ecl_bit_array_read_words(ecl_bit_array ba, size_t i, size_t first, uint32_t* words,
                                    size_t num) {
  return guard([&] {
    require(ba && i < ba->sizes.size(), "invalid argument");
    const uint64_t nw = (ba->sizes[i] + 31) / 32;
    require(first + num <= nw, "word range outside the array");
    set_device(ba->device);
    d2h(words, ba->words() + ba->offsets[i] + first, num);
  });
}
ecl_status ecl_bit_array_fill_random(ecl_bit_array ba, size_t i, uint64_t seed,
                                     uint64_t threshold) {
  return guard([&] {
    require(ba && i < ba->sizes.size(), "invalid argument");
    set_device(ba->device);
    const uint64_t padded = (((ba->sizes[i] + 31) / 32 + 31) & ~31ull);
    launch_gen_bits(ba->words() + ba->offsets[i], ba->sizes[i], padded, seed, threshold,
                    nullptr);
    ECL_HIP(hipDeviceSynchronize());
  });
}
size_t ecl_bit_array_needed_gpu_memory(size_t size, uint8_t num_arrays) {
  // bit_array.cu:140-151: sizes, offsets and data padded per array
  return static_cast<size_t>(num_arrays) * (2 * sizeof(uint64_t) + (size + 7) / 8 + 128);
}

// ---- RankSelect ---------------------------------------------------------------
ecl_status ecl_rank_select_create(ecl_bit_array* ba, int device, ecl_rank_select* out) {
  return guard([&] {
    require(ba && *ba && out, "NULL argument");
    require((*ba)->device == device, "the bit array lives on another device");
    set_device(device);
    *out = make_rank_select(*ba, nullptr);
    *ba = nullptr;
  });
}
ecl_status ecl_rank_select_destroy(ecl_rank_select rs) {
  return guard([&] {
    if (rs) {
      set_device(rs->device);
      delete rs;
    }
  });
}
ecl_status ecl_rank_select_bit_array(ecl_rank_select rs, ecl_bit_array* out) {
  return guard([&] {
    require(rs && out, "NULL argument");
    *out = rs->ba;
  });
}
ecl_status ecl_rank_select_info_get(ecl_rank_select rs, size_t i, ecl_rank_select_info* o) {
  return guard([&] {
    require(rs && o && i < rs->ba->sizes.size(), "invalid argument");
    o->size_bits = rs->ba->sizes[i];
    o->offset_words = rs->ba->offsets[i];
    o->num_l1_blocks = rs->num_l1[i];
    o->num_l2_blocks = rs->num_l2[i];
    o->num_last_l2_blocks = rs->num_last_l2[i];
    o->total_ones = rs->ones[i];
    o->total_zeros = rs->ba->sizes[i] - rs->ones[i];
    o->num_samples_1 = rs->ones[i] / detail::kSampleRate;
    o->num_samples_0 = o->total_zeros / detail::kSampleRate;
  });
}
ecl_status ecl_rank_select_get_l1(ecl_rank_select rs, size_t i, size_t first, size_t count,
                                  uint64_t* out) {
  return guard([&] {
    require(rs && i < rs->ba->sizes.size(), "invalid argument");
    require(first + count <= rs->num_l1[i], "L1 range out of bounds");
    set_device(rs->device);
    d2h(out, rs->l1.as<uint64_t>() + rs->l1_off[i] + first, count);
  });
}
ecl_status ecl_rank_select_get_l2(ecl_rank_select rs, size_t i, size_t first, size_t count,
                                  uint16_t* out) {
  return guard([&] {
    require(rs && i < rs->ba->sizes.size(), "invalid argument");
    require(first + count <= rs->num_l2[i], "L2 range out of bounds");
    set_device(rs->device);
    d2h(out, rs->l2.as<uint16_t>() + rs->l2_off[i] + first, count);
  });
}
ecl_status ecl_rank_select_get_samples(ecl_rank_select rs, size_t i, int value, size_t first,
                                       size_t count, uint64_t* out) {
  return guard([&] {
    require(rs && i < rs->ba->sizes.size(), "invalid argument");
    const uint64_t ones = rs->ones[i], zeros = rs->ba->sizes[i] - ones;
    const uint64_t ns = (value ? ones : zeros) / detail::kSampleRate;
    require(first + count <= ns, "sample range out of bounds");
    set_device(rs->device);
    const uint64_t* base = rs->samples.as<uint64_t>() + rs->s_off[i];
    if (value) {
      d2h(out, base + first, count);
    } else {
      // zeros samples are stored from the back: S0[j] at cap - 1 - j
      std::vector<uint64_t> tmp(count);
      if (count) d2h(tmp.data(), base + rs->cap[i] - first - count, count);
      for (size_t j = 0; j < count; ++j) out[j] = tmp[count - 1 - j];
    }
  });
}
ecl_status ecl_rank_select_device_pointers(ecl_rank_select rs, ecl_rank_select_device* o) {
  return guard([&] {
    require(rs && o, "NULL argument");
    o->l1 = rs->l1.as<uint64_t>();
    o->l2 = rs->l2.as<uint16_t>();
    o->l1_offsets = rs->d_l1_off.as<uint64_t>();
    o->l2_offsets = rs->d_l2_off.as<uint64_t>();
    o->num_l1_blocks = rs->d_num_l1.as<uint64_t>();
    o->num_last_l2_blocks = rs->d_num_last_l2.as<uint16_t>();
    o->samples = rs->samples.as<uint64_t>();
    o->sample_offsets = rs->d_s_off.as<uint64_t>();
    o->sample_caps = rs->d_cap.as<uint64_t>();
    o->total_ones = rs->total_ones.as<uint64_t>();
    o->levels = rs->d_levels.ptr;
    o->total_num_l2_blocks = rs->total_l2;
  });
}
ecl_status ecl_rank_select_rank_device(ecl_rank_select rs, size_t i, int value,
                                       const uint64_t* d_index, size_t m, uint64_t* d_out,
                                       void* stream) {
  return guard([&] {
    require(rs && i < rs->ba->sizes.size(), "invalid argument");
    set_device(rs->device);
    launch_rs_rank(rs->d_levels.as<LevelView>() + i, value, d_index, m, d_out, 0, 0,
                   rs->lines.ptr != nullptr, as_stream(stream));
  });
}
ecl_status ecl_rank_select_select_device(ecl_rank_select rs, size_t i, int value,
                                         const uint64_t* d_k, size_t m, uint64_t* d_out,
                                         void* stream) {
  return guard([&] {
    require(rs && i < rs->ba->sizes.size(), "invalid argument");
    set_device(rs->device);
    launch_rs_select(rs->d_levels.as<LevelView>() + i, value, d_k, m, d_out, 0, 0,
                     as_stream(stream));
  });
}
static void rs_host(ecl_rank_select rs, size_t i, int value, const uint64_t* in, size_t m,
                    uint64_t* out, bool is_rank) {
  require(rs && i < rs->ba->sizes.size(), "invalid argument");
  set_device(rs->device);
  for (size_t off = 0; off < m; off += kHostChunk) {
    const size_t c = std::min<size_t>(kHostChunk, m - off);
    check_query_memory(c * 16, is_rank ? "rank" : "select");
    DevBuf din(c * 8), dout(c * 8);
    h2d(din.ptr, in + off, c);
    if (is_rank)
      launch_rs_rank(rs->d_levels.as<LevelView>() + i, value, din.as<uint64_t>(), c,
                     dout.as<uint64_t>(), 0, 0, rs->lines.ptr != nullptr, nullptr);
    else
      launch_rs_select(rs->d_levels.as<LevelView>() + i, value, din.as<uint64_t>(), c,
                       dout.as<uint64_t>(), 0, 0, nullptr);
    d2h(out + off, dout.ptr, c);
  }
}
ecl_status ecl_rank_select_rank(ecl_rank_select rs, size_t i, int value, const uint64_t* index,
                                size_t m, uint64_t* out) {
  return guard([&] { rs_host(rs, i, value, index, m, out, true); });
}
ecl_status ecl_rank_select_select(ecl_rank_select rs, size_t i, int value, const uint64_t* k,
                                  size_t m, uint64_t* out) {
  return guard([&] { rs_host(rs, i, value, k, m, out, false); });
}
ecl_status ecl_rank_select_rank_generated(ecl_rank_select rs, size_t i, int value,
                                          uint64_t seed, size_t m, uint64_t* d_out,
                                          void* stream) {
  return guard([&] {
    require(rs && i < rs->ba->sizes.size(), "invalid argument");
    set_device(rs->device);
    launch_rs_rank(rs->d_levels.as<LevelView>() + i, value, nullptr, m, d_out, 1, seed,
                   rs->lines.ptr != nullptr, as_stream(stream));
  });
}
ecl_status ecl_rank_select_select_generated(ecl_rank_select rs, size_t i, int value,
                                            uint64_t seed, size_t m, uint64_t* d_out,
                                            void* stream) {
  return guard([&] {
    require(rs && i < rs->ba->sizes.size(), "invalid argument");
    set_device(rs->device);
    launch_rs_select(rs->d_levels.as<LevelView>() + i, value, nullptr, m, d_out, 1, seed,
                     as_stream(stream));
  });
}
size_t ecl_rank_select_needed_gpu_memory(size_t size, uint8_t num_arrays) {
  // rank_select.cu:382-409, with the shared sample buffer
  const size_t l1 = (size + 65535) / 65536, l2 = (size + 511) / 512;
  return static_cast<size_t>(num_arrays) *
         (l1 * 8 + l2 * 2 + (size / 4096) * 8 + 7 * sizeof(uint64_t) + sizeof(LevelView));
}

// ---- WaveletTree --------------------------------------------------------------
ecl_status ecl_wavelet_tree_create(ecl_type type, const void* text, size_t n,
                                   const void* alphabet, size_t asize, int device,
                                   ecl_wavelet_tree* out) {
  return guard([&] {
    require(out != nullptr, "out is NULL");
    *out = build_tree(static_cast<int>(type), text, false, n, alphabet, asize, device);
  });
}
ecl_status ecl_wavelet_tree_create_device(ecl_type type, const void* d_text, size_t n,
                                          const void* alphabet, size_t asize, int device,
                                          ecl_wavelet_tree* out) {
  return guard([&] {
    require(out != nullptr, "out is NULL");
    *out = build_tree(static_cast<int>(type), d_text, true, n, alphabet, asize, device);
  });
}
ecl_status ecl_wavelet_tree_destroy(ecl_wavelet_tree wt) {
  return guard([&] {
    if (wt) {
      set_device(wt->device);
      delete wt;
    }
  });
}
ecl_status ecl_wavelet_tree_build_query_lines(ecl_wavelet_tree wt) {
  return guard([&] {
    require(wt != nullptr, "NULL tree");
    set_device(wt->device);
    build_query_lines(wt->rs);
  });
}
ecl_status ecl_wavelet_tree_build_quad_lines(ecl_wavelet_tree wt) {
  return guard([&] {
    require(wt != nullptr, "NULL tree");
    set_device(wt->device);
    build_quad_lines(wt);
  });
}
ecl_status ecl_wavelet_tree_layout_views(ecl_wavelet_tree wt, const void** quad_view,
                                         const void** hex_view) {
  return guard([&] {
    require(wt != nullptr && quad_view != nullptr && hex_view != nullptr, "NULL argument");
    set_device(wt->device);
    if (wt->d_layouts.ptr == nullptr) {
      wt->d_layouts = DevBuf(sizeof(QuadView) + sizeof(HexView));
      sync_layout_views(wt);
    }
    *quad_view = wt->d_layouts.ptr;
    *hex_view = static_cast<const uint8_t*>(wt->d_layouts.ptr) + sizeof(QuadView);
  });
}
ecl_status ecl_wavelet_tree_build_hex_lines(ecl_wavelet_tree wt) {
  return guard([&] {
    require(wt != nullptr, "NULL tree");
    set_device(wt->device);
    build_hex_lines(wt);
  });
}
ecl_status ecl_wavelet_tree_build_quad_select_lines(ecl_wavelet_tree wt) {
  return guard([&] {
    require(wt != nullptr, "NULL tree");
    set_device(wt->device);
    build_quad_select_lines(wt);
  });
}
ecl_status ecl_wavelet_tree_build_select_lines(ecl_wavelet_tree wt) {
  return guard([&] {
    require(wt, "NULL argument");
    set_device(wt->device);
    build_select_lines(wt->rs);
  });
}
ecl_status ecl_rank_select_build_select_lines(ecl_rank_select rs) {
  return guard([&] {
    require(rs, "NULL argument");
    set_device(rs->device);
    build_select_lines(rs);
  });
}
ecl_status ecl_rank_select_build_query_lines(ecl_rank_select rs) {
  return guard([&] {
    require(rs != nullptr, "NULL rank select");
    set_device(rs->device);
    build_query_lines(rs);
  });
}
ecl_status ecl_wavelet_tree_save(ecl_wavelet_tree wt, const char* path) {
  return guard([&] {
    require(wt != nullptr, "NULL tree");
    save_tree(wt, path);
  });
}
ecl_status ecl_wavelet_tree_load(const char* path, int device, ecl_wavelet_tree* out) {
  return guard([&] {
    require(out != nullptr, "out is NULL");
    *out = load_tree(path, device);
  });
}
ecl_status ecl_wavelet_tree_info_get(ecl_wavelet_tree wt, ecl_wavelet_tree_info* o) {
  return guard([&] {
    require(wt && o, "NULL argument");
    o->data_size = wt->n;
    o->alphabet_size = wt->sigma;
    o->num_levels = wt->L;
    o->is_min_alphabet = wt->min_alphabet ? 1 : 0;
    o->codes_start = wt->codes_start;
    o->num_ranks = wt->ranks.size();
    o->nodes_until_last_level = wt->nodes_until_last_level;
    o->element_size = static_cast<uint32_t>(wt->esize);
    o->is_pow_two = wt->pow2 ? 1 : 0;
  });
}
ecl_status ecl_wavelet_tree_alphabet(ecl_wavelet_tree wt, void* out) {
  return guard([&] {
    require(wt && out, "NULL argument");
    u64_to_elems(wt->alphabet, out, wt->esize);
  });
}
ecl_status ecl_wavelet_tree_counts(ecl_wavelet_tree wt, uint64_t* out) {
  return guard([&] {
    require(wt && out, "NULL argument");
    std::copy(wt->counts.begin(), wt->counts.end(), out);
  });
}
ecl_status ecl_wavelet_tree_codes(ecl_wavelet_tree wt, uint64_t* codes, uint8_t* lens) {
  return guard([&] {
    require(wt != nullptr, "NULL argument");
    if (codes) std::copy(wt->codes.begin(), wt->codes.end(), codes);
    if (lens) std::copy(wt->code_lens.begin(), wt->code_lens.end(), lens);
  });
}
ecl_status ecl_wavelet_tree_node_ranks(ecl_wavelet_tree wt, uint64_t* ranks,
                                       uint64_t* node_start, uint8_t* node_level) {
  return guard([&] {
    require(wt != nullptr, "NULL argument");
    if (ranks) std::copy(wt->ranks.begin(), wt->ranks.end(), ranks);
    if (node_start) std::copy(wt->node_start.begin(), wt->node_start.end(), node_start);
    if (node_level) std::copy(wt->node_level.begin(), wt->node_level.end(), node_level);
  });
}
ecl_status ecl_wavelet_tree_nodes_at_level(ecl_wavelet_tree wt, uint64_t* out) {
  return guard([&] {
    require(wt && out, "NULL argument");
    for (uint32_t l = 0; l + 1 < wt->L; ++l)
      out[l] = wt->pow2 ? (1ull << l) - 1 : wt->nodes_at_level[l];
  });
}
ecl_status ecl_wavelet_tree_rank_select(ecl_wavelet_tree wt, ecl_rank_select* out) {
  return guard([&] {
    require(wt && out, "NULL argument");
    *out = wt->rs;
  });
}
ecl_status ecl_wavelet_tree_device_view(ecl_wavelet_tree wt, void* view, size_t bytes) {
  return guard([&] {
    require(wt && view && bytes == sizeof(TreeView), "view size mismatch");
    std::memcpy(view, &wt->view, sizeof(TreeView));
  });
}
size_t ecl_wavelet_tree_needed_gpu_memory(size_t n, size_t sigma, uint8_t L, size_t) {
  return tree_needed_bytes(n, sigma, L, 8, false);
}

// Chunked host batch (the reference's graph of H2D -> kernel -> D2H per
// chunk, wavelet_tree.cuh:94-151, 1121-1211): chunk i uses stream i % 2; the
// host copies chunk i+1 into pinned memory while chunk i runs, and copies
// the results of chunk i-2 out before reusing its buffers.
}  // extern "C"

template <typename Launch>
static void pipelined(ecl_wavelet_tree wt, const void* in, size_t in_elem, void* out,
                      size_t out_elem, size_t m, const char* kind, Launch&& launch) {
  HostPipe& P = *wt->pipe;
  std::lock_guard<std::mutex> lk(P.mu);
  const size_t chunk = std::min<size_t>(m, kHostChunk / 4);
  const size_t nchunks = (m + chunk - 1) / chunk;
  check_query_memory(2 * chunk * (in_elem + out_elem), kind);
  P.reserve(chunk * in_elem, chunk * out_elem);
  size_t pending_off[2] = {0, 0}, pending_len[2] = {0, 0};
  auto drain = [&](int b) {
    if (!pending_len[b]) return;
    ECL_HIP(hipStreamSynchronize(P.stream[b]));
    parallel_memcpy(static_cast<uint8_t*>(out) + pending_off[b] * out_elem, P.pin_out[b],
                    pending_len[b] * out_elem);
    pending_len[b] = 0;
  };
  for (size_t c = 0; c < nchunks; ++c) {
    const int b = static_cast<int>(c & 1);
    const size_t off = c * chunk, len = std::min(chunk, m - off);
    drain(b);
    parallel_memcpy(P.pin_in[b], static_cast<const uint8_t*>(in) + off * in_elem, len * in_elem);
    ECL_HIP(hipMemcpyAsync(P.d_in[b].ptr, P.pin_in[b], len * in_elem, hipMemcpyHostToDevice,
                           P.stream[b]));
    launch(P.d_in[b].ptr, len, P.d_out[b].ptr, P.stream[b]);
    ECL_HIP(hipMemcpyAsync(P.pin_out[b], P.d_out[b].ptr, len * out_elem,
                           hipMemcpyDeviceToHost, P.stream[b]));
    pending_off[b] = off;
    pending_len[b] = len;
  }
  drain(0);
  drain(1);
}

extern "C" {

ecl_status ecl_wavelet_tree_access(ecl_wavelet_tree wt, const uint64_t* idx, size_t m,
                                   void* out, uint32_t flags) {
  return guard([&] {
    require(wt && (m == 0 || (idx && out)), "NULL argument");
    if (m == 0) return;
    set_device(wt->device);
    const TreeQueryParams p = query_params(wt, flags);
    pipelined(wt, idx, 8, out, wt->esize, m, "access",
              [&](void* din, size_t len, void* dout, hipStream_t st) {
                launch_tree_access(p, static_cast<uint64_t*>(din), len, dout, st);
              });
  });
}
static void tree_rs_host(ecl_wavelet_tree wt, const ecl_rank_select_query* q, size_t m,
                         uint64_t* out, uint32_t flags, bool is_rank) {
  require(wt && (m == 0 || (q && out)), "NULL argument");
  if (m == 0) return;
  set_device(wt->device);
  const TreeQueryParams p = query_params(wt, flags);
  pipelined(wt, q, 16, out, 8, m, is_rank ? "rank" : "select",
            [&](void* din, size_t len, void* dout, hipStream_t st) {
              if (is_rank)
                launch_tree_rank(p, din, len, static_cast<uint64_t*>(dout), st);
              else
                launch_tree_select(p, din, len, static_cast<uint64_t*>(dout), st);
            });
}
ecl_status ecl_wavelet_tree_rank(ecl_wavelet_tree wt, const ecl_rank_select_query* q,
                                 size_t m, uint64_t* out, uint32_t flags) {
  return guard([&] { tree_rs_host(wt, q, m, out, flags, true); });
}
ecl_status ecl_wavelet_tree_select(ecl_wavelet_tree wt, const ecl_rank_select_query* q,
                                   size_t m, uint64_t* out, uint32_t flags) {
  return guard([&] { tree_rs_host(wt, q, m, out, flags, false); });
}
ecl_status ecl_wavelet_tree_access_device(ecl_wavelet_tree wt, const uint64_t* d_idx, size_t m,
                                          void* d_out, uint32_t flags, void* stream) {
  return guard([&] {
    require(wt != nullptr, "NULL argument");
    set_device(wt->device);
    launch_tree_access(query_params(wt, flags), d_idx, m, d_out, as_stream(stream));
  });
}
ecl_status ecl_wavelet_tree_rank_device(ecl_wavelet_tree wt, const ecl_rank_select_query* d_q,
                                        size_t m, uint64_t* d_out, uint32_t flags,
                                        void* stream) {
  return guard([&] {
    require(wt != nullptr, "NULL argument");
    set_device(wt->device);
    launch_tree_rank(query_params(wt, flags), d_q, m, d_out, as_stream(stream));
  });
}
ecl_status ecl_wavelet_tree_select_device(ecl_wavelet_tree wt,
                                          const ecl_rank_select_query* d_q, size_t m,
                                          uint64_t* d_out, uint32_t flags, void* stream) {
  return guard([&] {
    require(wt != nullptr, "NULL argument");
    set_device(wt->device);
    launch_tree_select(query_params(wt, flags), d_q, m, d_out, as_stream(stream));
  });
}
ecl_status ecl_wavelet_tree_query_batch_device(ecl_wavelet_tree wt, const ecl_query* d_q,
                                               size_t m, uint64_t* d_out, void* stream) {
  return guard([&] {
    require(wt != nullptr, "NULL argument");
    set_device(wt->device);
    launch_tree_mixed(query_params(wt, ECL_SYMBOLS_MINIMAL | ECL_RESULTS_MINIMAL), d_q, m,
                      d_out, as_stream(stream));
  });
}
ecl_status ecl_wavelet_tree_generate_queries(ecl_wavelet_tree wt, int kind, uint64_t seed,
                                             size_t m, void* d_out, void* stream) {
  return guard([&] {
    require(wt != nullptr && kind >= 0 && kind <= 3, "invalid argument");
    set_device(wt->device);
    launch_tree_gen_queries(wt->view, nullptr, kind, seed, m, d_out, wt->esize,
                            as_stream(stream));
  });
}

ecl_status ecl_wavelet_tree_set_query_lanes(ecl_wavelet_tree wt, uint32_t lanes) {
  return guard([&] {
    require(wt != nullptr, "NULL tree");
    require(lanes <= 2, "lanes must be 0, 1 or 2");
    wt->lanes = lanes;
  });
}

// Times access and rank with one and with two lanes per query on m
// device-generated queries of each kind over the tree as it is now (with or
// without query lines), and keeps the faster setting for the tree.
ecl_status ecl_wavelet_tree_autotune(ecl_wavelet_tree wt, uint64_t m, uint64_t seed,
                                     uint32_t* lanes, float* ms) {
  return guard([&] {
    require(wt != nullptr, "NULL tree");
    require(m > 0, "m must be positive");
    set_device(wt->device);
    hipStream_t s = nullptr;
    DevBuf d_idx(m * 8), d_rq(m * 16), d_out(m * 8);
    launch_tree_gen_queries(wt->view, nullptr, 0, seed, m, d_idx.ptr, wt->esize, s);
    launch_tree_gen_queries(wt->view, nullptr, 1, seed + 1, m, d_rq.ptr, wt->esize, s);
    hipEvent_t e0, e1;
    ECL_HIP(hipEventCreate(&e0));
    ECL_HIP(hipEventCreate(&e1));
    float t[2] = {0.f, 0.f};
    try {
      for (uint32_t g = 1; g <= 2; ++g) {
        TreeQueryParams p = query_params(wt, ECL_SYMBOLS_MINIMAL | ECL_RESULTS_MINIMAL);
        p.lanes = g;
        for (int rep = 0; rep < 3; ++rep) {  // the first round warms up
          ECL_HIP(hipEventRecord(e0, s));
          launch_tree_access(p, d_idx.as<uint64_t>(), m, d_out.ptr, s);
          launch_tree_rank(p, d_rq.ptr, m, d_out.as<uint64_t>(), s);
          ECL_HIP(hipEventRecord(e1, s));
          ECL_HIP(hipEventSynchronize(e1));
          float x = 0.f;
          ECL_HIP(hipEventElapsedTime(&x, e0, e1));
          if (rep == 1 || (rep == 2 && x < t[g - 1])) t[g - 1] = x;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    wt->lanes = t[1] < t[0] ? 2 : 1;
    if (lanes) *lanes = wt->lanes;
    if (ms) {
      ms[0] = t[0];
      ms[1] = t[1];
    }
  });
}

}  // extern "C"
