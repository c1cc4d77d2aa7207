// server.hip -- a query server over device-resident trees with continuous
// batching.
//
// Requests wait in one FIFO queue.  A scheduler thread takes the free batch
// slot, fills it from the head of the queue (whole requests, and a part of
// the last one when it does not fit), and launches host-to-device copy,
// query kernel and device-to-host copy on the slot's own stream.  While
// that batch runs, the next slot is filled from whatever has arrived since:
// up to `max_inflight` batches overlap on the device.  A completion thread
// waits for the oldest batch, copies each request's results out and wakes
// its waiter.  With coalescing off, every request is its own batch and one
// batch runs at a time: the plain per-call form, kept for comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host.hpp"

using namespace ecl::lib;
using ecl::detail::TreeView;

namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kMaxSegments = 65536;  // requests per batch
// smallest batch the coalescing scheduler cuts when it splits a backlog
// over the free slots (1 MiB of queries)
constexpr uint64_t kMinBatch = 1ull << 16;

struct Request {
  uint64_t id = 0;
  uint32_t tree = 0;
  const ecl_query* q = nullptr;
  uint64_t m = 0;
  uint64_t* out = nullptr;
  uint64_t taken = 0;  // queries placed in batches
  uint64_t done = 0;   // queries with results back
  bool finished = false;
  bool detached = false;  // nobody waits: dropped when finished
  Clock::time_point t_submit, t_done;
};

struct Part {
  Request* r;
  uint64_t req_off, batch_off, count;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  ecl_query* pin_q = nullptr;     // queries of the batch
  uint64_t* pin_seg = nullptr;    // segment starts
  uint32_t* pin_tree = nullptr;   // segment trees
  uint64_t* pin_out = nullptr;
  DevBuf d_q, d_seg, d_tree, d_out;
  std::vector<Part> parts;
  uint64_t count = 0;
};

struct TreeEntry {
  ecl_wavelet_tree wt = nullptr;
  uint64_t bytes = 0;
  uint64_t inflight = 0;  // unfinished requests
  bool active = false, removing = false;
};

}  // namespace

struct ecl_server_t {
  int device = 0;
  ecl_server_config cfg{};
  DevBuf d_views;              // TreeView per tree slot
  DevBuf d_params;             // TreeQueryParams per tree slot (query_lines 3)
  std::vector<TreeEntry> trees;
  uint64_t hbm_used = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  uint32_t q_block = 256, q_grid = 2048;
  hipStream_t ctl = nullptr;   // view uploads

  std::mutex mu;
  std::mutex build_mu;         // one admission at a time (memory accounting)
  std::condition_variable cv_sched, cv_done, cv_wait;
  std::deque<Request*> queue;
  uint64_t queued = 0;         // queries in the queue not yet placed in a batch
  std::vector<Slot*> free_slots;
  std::deque<Slot*> inflight;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  uint64_t next_id = 1;
  bool stop = false, sched_exited = false;
  std::thread sched, completer;

  // statistics since the last reset
  std::vector<double> lat_ms;
  uint64_t st_queries = 0, st_batches = 0;
  bool st_any = false;
  Clock::time_point st_first, st_last;

  void scheduler_loop();
  void completer_loop();
  void release_tree_locked(uint32_t t);
};

void ecl_server_t::scheduler_loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    // on shutdown the queue still drains: stop only once it is empty
    cv_sched.wait(lk, [&] {
      return (stop && queue.empty()) ||
             (!queue.empty() && !free_slots.empty() && (cfg.coalesce || inflight.empty()));
    });
    if (queue.empty()) break;
    Slot* sl = free_slots.back();
    free_slots.pop_back();
    sl->parts.clear();
    // Batch size.  With coalescing, a backlog is split over the batch slots
    // (at least kMinBatch queries a batch), so that the first part of it
    // completes while the rest is staged: at 10^4 requests/s this gave a
    // p99 of 2.5 ms against 6.6 ms for batches that take everything queued.
    // The request at the head is not cut below its own size, though: a
    // lone request split over the slots has its parts' staging copies,
    // launches and copy-outs run one after another on the scheduler and
    // completion threads (a 10^6-query request took 1.43 ms as three
    // batches against 1.04 ms as one).  At most max_batch either way.  The
    // plain form takes one whole request (up to max_batch) at a time.
    uint64_t target = cfg.max_batch;
    if (cfg.coalesce) {
      const uint64_t share = (queued + cfg.max_inflight - 1) / cfg.max_inflight;
      const uint64_t head = queue.front()->m - queue.front()->taken;
      target = std::min<uint64_t>(cfg.max_batch, std::max({kMinBatch, share, head}));
    }
    uint64_t count = 0;
    while (!queue.empty() && count < target && sl->parts.size() < kMaxSegments) {
      Request* r = queue.front();
      const uint64_t c = std::min<uint64_t>(r->m - r->taken, target - count);
      sl->parts.push_back({r, r->taken, count, c});
      r->taken += c;
      count += c;
      if (r->taken == r->m) queue.pop_front();
      if (!cfg.coalesce) break;
    }
    queued -= count;
    sl->count = count;
    lk.unlock();
    // fill the pinned batch and launch copy-in, kernel, copy-out
    const uint32_t nseg = static_cast<uint32_t>(sl->parts.size());
    for (uint32_t j = 0; j < nseg; ++j) {
      const Part& pt = sl->parts[j];
      // large parts are copied by the library's host copy threads
      parallel_memcpy(sl->pin_q + pt.batch_off, pt.r->q + pt.req_off,
                      pt.count * sizeof(ecl_query));
      sl->pin_seg[j] = pt.batch_off;
      sl->pin_tree[j] = pt.r->tree;
    }
    bool ok = true;
    try {
      ECL_HIP(hipMemcpyAsync(sl->d_q.ptr, sl->pin_q, count * sizeof(ecl_query),
                             hipMemcpyHostToDevice, sl->stream));
      ECL_HIP(hipMemcpyAsync(sl->d_seg.ptr, sl->pin_seg, nseg * 8ull, hipMemcpyHostToDevice,
                             sl->stream));
      ECL_HIP(hipMemcpyAsync(sl->d_tree.ptr, sl->pin_tree, nseg * 4ull, hipMemcpyHostToDevice,
                             sl->stream));
      if (cfg.query_lines >= 3)
        launch_segment_query_quad(d_views.as<TreeView>(), d_params.as<TreeQueryParams>(),
                                  trees.size(), sl->d_q.as<ecl_query>(), count,
                                  sl->d_seg.as<uint64_t>(), sl->d_tree.as<uint32_t>(), nseg,
                                  sl->d_out.as<uint64_t>(), q_grid, sl->stream);
      else
        launch_segment_query(d_views.as<TreeView>(), trees.size(), sl->d_q.as<ecl_query>(),
                             count, sl->d_seg.as<uint64_t>(), sl->d_tree.as<uint32_t>(), nseg,
                             sl->d_out.as<uint64_t>(), q_block, q_grid, sl->stream);
      ECL_HIP(hipMemcpyAsync(sl->pin_out, sl->d_out.ptr, count * 8ull, hipMemcpyDeviceToHost,
                             sl->stream));
      ECL_HIP(hipEventRecord(sl->done, sl->stream));
    } catch (...) {
      ok = false;
    }
    lk.lock();
    if (!ok) {  // results stay zero; the requests still complete
      std::memset(sl->pin_out, 0, count * 8ull);
      (void)hipStreamSynchronize(sl->stream);
      (void)hipEventRecord(sl->done, sl->stream);
    }
    inflight.push_back(sl);
    ++st_batches;
    cv_done.notify_one();
  }
  sched_exited = true;
  cv_done.notify_all();
}

void ecl_server_t::release_tree_locked(uint32_t t) {
  TreeEntry& e = trees[t];
  if (!e.removing || e.inflight > 0 || !e.wt) return;
  (void)ecl_wavelet_tree_destroy(e.wt);
  hbm_used -= e.bytes;
  e = TreeEntry{};
}

void ecl_server_t::completer_loop() {
  (void)hipSetDevice(device);
  while (true) {
    Slot* sl = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_done.wait(lk, [&] { return !inflight.empty() || sched_exited; });
      if (inflight.empty()) break;
      sl = inflight.front();
    }
    (void)hipEventSynchronize(sl->done);
    for (const Part& pt : sl->parts)
      parallel_memcpy(pt.r->out + pt.req_off, sl->pin_out + pt.batch_off, pt.count * 8ull);
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    for (const Part& pt : sl->parts) {
      Request* r = pt.r;
      r->done += pt.count;
      if (r->done == r->m && !r->finished) {
        r->finished = true;
        r->t_done = now;
        lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
        st_queries += r->m;
        st_last = now;
        TreeEntry& e = trees[r->tree];
        e.inflight--;
        release_tree_locked(r->tree);
        if (r->detached) requests.erase(r->id);
      }
    }
    inflight.pop_front();
    free_slots.push_back(sl);
    cv_sched.notify_one();
    cv_wait.notify_all();
  }
}

namespace {

ecl_status submit_impl(ecl_server s, uint32_t tree, const ecl_query* q, size_t m, uint64_t* out,
                       bool detached, uint64_t* id_out) {
  return guard([&] {
    require(s && q && out, "NULL argument");
    require(m > 0, "empty request");
    std::lock_guard<std::mutex> lk(s->mu);
    require(!s->stop, "the server is shutting down");
    require(tree < s->trees.size() && s->trees[tree].active && !s->trees[tree].removing,
            "no such tree");
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->tree = tree;
    r->q = q;
    r->m = m;
    r->out = out;
    r->detached = detached;
    r->t_submit = Clock::now();
    if (!s->st_any) {
      s->st_any = true;
      s->st_first = r->t_submit;
    }
    s->trees[tree].inflight++;
    Request* raw = r.get();
    s->requests.emplace(r->id, std::move(r));
    s->queue.push_back(raw);
    s->queued += m;
    if (id_out) *id_out = raw->id;
    s->cv_sched.notify_one();
  });
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  const size_t k = std::min(v.size() - 1, static_cast<size_t>(std::ceil(p * v.size())) - 1);
  std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(k), v.end());
  return v[k];
}

}  // namespace

extern "C" {

ecl_status ecl_server_create(int device, const ecl_server_config* config, ecl_server* out) {
  return guard([&] {
    require(out, "NULL argument");
    set_device(device);
    auto s = std::make_unique<ecl_server_t>();
    s->device = device;
    if (config) s->cfg = *config;
    else s->cfg.coalesce = 1;
    if (s->cfg.max_batch == 0) s->cfg.max_batch = 1ull << 22;
    if (s->cfg.max_inflight == 0) s->cfg.max_inflight = 3;
    if (s->cfg.max_trees == 0) s->cfg.max_trees = 1024;
    if (s->cfg.hbm_budget == 0) s->cfg.hbm_budget = free_device_memory() / 10 * 9;
    require(s->cfg.max_inflight <= 16, "at most 16 batches in flight");
    require(s->cfg.query_lines <= 3, "query_lines must be 0, 1, 2 or 3");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
      s->q_grid = static_cast<uint32_t>(cus) * 8;
    s->trees.resize(s->cfg.max_trees);
    s->d_views = DevBuf(s->cfg.max_trees * sizeof(TreeView));
    ECL_HIP(hipMemset(s->d_views.ptr, 0, s->cfg.max_trees * sizeof(TreeView)));
    if (s->cfg.query_lines >= 3) {
      s->d_params = DevBuf(s->cfg.max_trees * sizeof(TreeQueryParams));
      ECL_HIP(hipMemset(s->d_params.ptr, 0, s->cfg.max_trees * sizeof(TreeQueryParams)));
    }
    ECL_HIP(hipStreamCreateWithFlags(&s->ctl, hipStreamNonBlocking));
    const uint64_t B = s->cfg.max_batch;
    for (uint32_t i = 0; i < s->cfg.max_inflight; ++i) {
      auto sl = std::make_unique<Slot>();
      ECL_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
      ECL_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
      void* p = nullptr;
      ECL_HIP(hipHostMalloc(&p, B * sizeof(ecl_query), hipHostMallocDefault));
      sl->pin_q = static_cast<ecl_query*>(p);
      ECL_HIP(hipHostMalloc(&p, kMaxSegments * 8ull, hipHostMallocDefault));
      sl->pin_seg = static_cast<uint64_t*>(p);
      ECL_HIP(hipHostMalloc(&p, kMaxSegments * 4ull, hipHostMallocDefault));
      sl->pin_tree = static_cast<uint32_t*>(p);
      ECL_HIP(hipHostMalloc(&p, B * 8, hipHostMallocDefault));
      sl->pin_out = static_cast<uint64_t*>(p);
      sl->d_q = DevBuf(B * sizeof(ecl_query));
      sl->d_seg = DevBuf(kMaxSegments * 8ull);
      sl->d_tree = DevBuf(kMaxSegments * 4ull);
      sl->d_out = DevBuf(B * 8);
      s->free_slots.push_back(sl.get());
      s->slots.push_back(std::move(sl));
    }
    ecl_server_t* raw = s.get();
    s->sched = std::thread([raw] { raw->scheduler_loop(); });
    s->completer = std::thread([raw] { raw->completer_loop(); });
    *out = s.release();
  });
}

ecl_status ecl_server_destroy(ecl_server s) {
  return guard([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_sched.notify_all();
    if (s->sched.joinable()) s->sched.join();
    if (s->completer.joinable()) s->completer.join();
    set_device(s->device);
    for (auto& e : s->trees)
      if (e.wt) (void)ecl_wavelet_tree_destroy(e.wt);
    for (auto& sl : s->slots) {
      if (sl->stream) (void)hipStreamDestroy(sl->stream);
      if (sl->done) (void)hipEventDestroy(sl->done);
      if (sl->pin_q) (void)hipHostFree(sl->pin_q);
      if (sl->pin_seg) (void)hipHostFree(sl->pin_seg);
      if (sl->pin_tree) (void)hipHostFree(sl->pin_tree);
      if (sl->pin_out) (void)hipHostFree(sl->pin_out);
    }
    if (s->ctl) (void)hipStreamDestroy(s->ctl);
    delete s;
  });
}

}  // extern "C"

// Bits, directory and metadata of a uint8 tree of n symbols over sigma;
// with lines >= 1 its query lines (128 B per 960 bits) and select hints (64 B
// per 4096 ones or zeros, plus one per Value) of every level; with lines == 2
// also its select lines (per array and Value one 128 B line every r of them,
// r = max(1, floor(800 x density)): at most 4 / 800 lines per bit).
// quad lines and quad select lines (query_lines 3) need a power-of-two
// alphabet with an even number of levels; hex lines a level count that is a
// multiple of 4
static bool quad_layout(uint64_t sigma) {
  const uint32_t L = ecl::detail::ceil_log2(sigma);
  return (sigma & (sigma - 1)) == 0 && L >= 2 && L % 2 == 0;
}

static uint64_t tree_estimate(uint64_t n, uint64_t sigma, uint32_t lines) {
  const uint64_t L = ecl::detail::ceil_log2(sigma);
  uint64_t b = L * ((n + 1023) / 1024 * 128 + (n + 511) / 512 * 2 + (n + 65535) / 65536 * 8 +
                    n / 4096 * 8 + 1024) +
               sigma * 64 + 4096;
  if (lines >= 3 && quad_layout(sigma)) {
    // quad lines, quad select lines (a line every ~85 of each digit at
    // density 1/4: 1.5 bytes per symbol, plus 4 lines per digit), hex lines
    b += L / 2 * ((n / kQuadDigits + 1) * 128 + n / 2 * 3 + 4 * 128 * 2) + 65536;
    if (L % 4 == 0) b += L / 4 * ((n / kHexDigits + 1) * 128) + 65536;
    return b;
  }
  if (lines >= 1)
    b += L * ((n + ecl::detail::kLineBits - 1) / ecl::detail::kLineBits * 128 +
              (n / ecl::detail::kSampleRate + 2) * 64);
  if (lines >= 2) b += L * ((n / 200 + 4) * 128);
  return b;
}

// free device memory as the library sees it (the runtime's free bytes plus
// the library's cached blocks, which its allocations reuse)
static size_t lib_free_memory(int device) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return fr + pooled_bytes(device);
}

// The tree's query lines (and select lines) when the server is configured
// for them; on failure the tree is destroyed and the error returned.
static ecl_status add_query_lines(ecl_server s, ecl_wavelet_tree wt, std::string* msg) {
  if (!s->cfg.query_lines) return ECL_OK;
  ecl_wavelet_tree_info info{};
  ecl_status st = ecl_wavelet_tree_info_get(wt, &info);
  if (st == ECL_OK && s->cfg.query_lines >= 3 && quad_layout(info.alphabet_size)) {
    // the quad layouts (quad select lines build the quad lines first), and
    // the hex lines where the level count is a multiple of 4
    st = ecl_wavelet_tree_build_quad_select_lines(wt);
    if (st == ECL_OK && info.num_levels % 4 == 0) st = ecl_wavelet_tree_build_hex_lines(wt);
  } else if (st == ECL_OK) {
    st = ecl_wavelet_tree_build_query_lines(wt);
    if (st == ECL_OK && s->cfg.query_lines >= 2) st = ecl_wavelet_tree_build_select_lines(wt);
  }
  if (st != ECL_OK) {
    *msg = last_error();
    (void)ecl_wavelet_tree_destroy(wt);
  }
  return st;
}

// Admission: reserves a slot and `estimate` bytes of the budget.
static uint32_t reserve_slot(ecl_server s, uint64_t estimate) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->hbm_used + estimate > s->cfg.hbm_budget)
    throw Error(ECL_ERROR_OUT_OF_MEMORY_TREE, "Not enough memory available for the wavelet tree.");
  uint32_t slot = 0;
  while (slot < s->trees.size() && (s->trees[slot].wt || s->trees[slot].active)) ++slot;
  require(slot < s->trees.size(), "no free tree slot");
  s->trees[slot].active = false;
  s->trees[slot].bytes = estimate;
  s->trees[slot].wt = reinterpret_cast<ecl_wavelet_tree>(1);  // reserved while building
  s->hbm_used += estimate;
  return slot;
}

static void release_slot(ecl_server s, uint32_t slot, uint64_t estimate) {
  std::lock_guard<std::mutex> lk(s->mu);
  s->hbm_used -= estimate;
  s->trees[slot] = TreeEntry{};
}

// Publishes a built tree in its slot: its view goes to the device table and
// the budget is charged what the build actually took.
static void commit_slot(ecl_server s, uint32_t slot, ecl_wavelet_tree wt, uint64_t estimate,
                        size_t fr0, size_t fr1) {
  TreeView v{};
  (void)ecl_wavelet_tree_device_view(wt, &v, sizeof(v));
  ECL_HIP(hipMemcpyAsync(s->d_views.as<TreeView>() + slot, &v, sizeof(v), hipMemcpyHostToDevice,
                         s->ctl));
  TreeQueryParams qp{};
  if (s->cfg.query_lines >= 3) {
    qp = tree_query_params(wt);
    ECL_HIP(hipMemcpyAsync(s->d_params.as<TreeQueryParams>() + slot, &qp, sizeof(qp),
                           hipMemcpyHostToDevice, s->ctl));
  }
  ECL_HIP(hipStreamSynchronize(s->ctl));
  std::lock_guard<std::mutex> lk(s->mu);
  TreeEntry& e = s->trees[slot];
  const uint64_t used = fr0 > fr1 ? fr0 - fr1 : estimate;
  s->hbm_used = s->hbm_used - estimate + used;
  e.wt = wt;
  e.bytes = used;
  e.active = true;
}

extern "C" {

ecl_status ecl_server_add_tree(ecl_server s, const void* text, int text_on_device, size_t n,
                               uint64_t sigma, uint32_t* tree) {
  return guard([&] {
    require(s && text && tree, "NULL argument");
    require(sigma >= 2 && sigma <= 256, "alphabet sizes must be in [2, 256]");
    require(n > 0, "empty text");
    set_device(s->device);
    std::lock_guard<std::mutex> build_lk(s->build_mu);
    const uint64_t estimate = tree_estimate(n, sigma, s->cfg.query_lines);
    const uint32_t slot = reserve_slot(s, estimate);
    std::vector<uint8_t> alpha(sigma);
    for (uint64_t c = 0; c < sigma; ++c) alpha[c] = static_cast<uint8_t>(c);
    ecl_wavelet_tree wt = nullptr;
    const size_t fr0 = lib_free_memory(s->device);
    ecl_status st =
        text_on_device ? ecl_wavelet_tree_create_device(ECL_U8, text, n, alpha.data(), sigma,
                                                        s->device, &wt)
                       : ecl_wavelet_tree_create(ECL_U8, text, n, alpha.data(), sigma, s->device,
                                                 &wt);
    std::string msg = st != ECL_OK ? last_error() : std::string();
    if (st == ECL_OK) st = add_query_lines(s, wt, &msg);
    const size_t fr1 = lib_free_memory(s->device);
    if (st != ECL_OK) {
      release_slot(s, slot, estimate);
      throw Error(st, msg);
    }
    commit_slot(s, slot, wt, estimate, fr0, fr1);
    *tree = slot;
  });
}

ecl_status ecl_server_add_tree_file(ecl_server s, const char* path, uint32_t* tree) {
  return guard([&] {
    require(s && path && tree, "NULL argument");
    int esize = 0;
    uint64_t n = 0, sigma = 0;
    read_tree_file_header(path, &esize, &n, &sigma);
    require(esize == 1 && sigma >= 2 && sigma <= 256,
            "the server holds uint8 trees with alphabets of 2 to 256 symbols");
    set_device(s->device);
    std::lock_guard<std::mutex> build_lk(s->build_mu);
    const uint64_t estimate = tree_estimate(n, sigma, s->cfg.query_lines);
    const uint32_t slot = reserve_slot(s, estimate);
    ecl_wavelet_tree wt = nullptr;
    const size_t fr0 = lib_free_memory(s->device);
    ecl_status st = ecl_wavelet_tree_load(path, s->device, &wt);
    std::string msg = st != ECL_OK ? last_error() : std::string();
    if (st == ECL_OK) {
      ecl_wavelet_tree_info info{};
      (void)ecl_wavelet_tree_info_get(wt, &info);
      if (!info.is_min_alphabet) {
        (void)ecl_wavelet_tree_destroy(wt);
        st = ECL_ERROR_INVALID_ARGUMENT;
        msg = "the server holds trees over the alphabet 0 .. sigma-1";
      }
    }
    if (st == ECL_OK) st = add_query_lines(s, wt, &msg);
    const size_t fr1 = lib_free_memory(s->device);
    if (st != ECL_OK) {
      release_slot(s, slot, estimate);
      throw Error(st, msg);
    }
    commit_slot(s, slot, wt, estimate, fr0, fr1);
    *tree = slot;
  });
}

ecl_status ecl_server_remove_tree(ecl_server s, uint32_t tree) {
  return guard([&] {
    require(s, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    require(tree < s->trees.size() && s->trees[tree].active && !s->trees[tree].removing,
            "no such tree");
    s->trees[tree].removing = true;
    s->release_tree_locked(tree);
  });
}

ecl_status ecl_server_submit(ecl_server s, uint32_t tree, const ecl_query* queries, size_t m,
                             uint64_t* results, uint64_t* request) {
  return submit_impl(s, tree, queries, m, results, false, request);
}

ecl_status ecl_server_wait(ecl_server s, uint64_t request, double* latency_ms) {
  return guard([&] {
    require(s, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(request);
    require(it != s->requests.end() && !it->second->detached, "unknown request");
    Request* r = it->second.get();
    s->cv_wait.wait(lk, [&] { return r->finished; });
    if (latency_ms)
      *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    s->requests.erase(it);
  });
}

ecl_status ecl_server_set_mode(ecl_server s, uint32_t coalesce) {
  return guard([&] {
    require(s, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->cfg.coalesce = coalesce ? 1 : 0;
    s->cv_sched.notify_all();
  });
}

ecl_status ecl_server_stats_get(ecl_server s, ecl_server_stats* out) {
  return guard([&] {
    require(s && out, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    ecl_server_stats r{};
    r.requests = s->lat_ms.size();
    r.queries = s->st_queries;
    r.batches = s->st_batches;
    r.p50_ms = percentile(s->lat_ms, 0.50);
    r.p99_ms = percentile(s->lat_ms, 0.99);
    double sum = 0, mx = 0;
    for (double x : s->lat_ms) { sum += x; mx = std::max(mx, x); }
    r.mean_ms = r.requests ? sum / r.requests : 0.0;
    r.max_ms = mx;
    r.elapsed_s = s->st_any && r.requests
                      ? std::chrono::duration<double>(s->st_last - s->st_first).count()
                      : 0.0;
    r.queries_per_s = r.elapsed_s > 0 ? r.queries / r.elapsed_s : 0.0;
    r.hbm_used = s->hbm_used;
    for (const auto& e : s->trees) r.trees += e.active ? 1 : 0;
    *out = r;
  });
}

ecl_status ecl_server_stats_reset(ecl_server s) {
  return guard([&] {
    require(s, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_ms.clear();
    s->st_queries = 0;
    s->st_batches = 0;
    s->st_any = false;
  });
}

ecl_status ecl_server_load(ecl_server s, const ecl_load_spec* spec, ecl_server_stats* out) {
  return guard([&] {
    require(s && spec && out, "NULL argument");
    require(spec->rate > 0 && spec->seconds > 0, "rate and duration must be positive");
    require(spec->min_size >= 1 && spec->max_size >= spec->min_size, "invalid request sizes");
    set_device(s->device);
    std::vector<uint32_t> live;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (uint32_t t = 0; t < s->trees.size(); ++t)
        if (s->trees[t].active && !s->trees[t].removing) live.push_back(t);
    }
    require(!live.empty(), "no resident trees");
    const uint64_t pool = std::max<uint64_t>(spec->pool ? spec->pool : (1ull << 20), spec->max_size);
    // per-tree query pools drawn on the device, results scratch per tree
    std::vector<std::vector<ecl_query>> qpool(live.size());
    std::vector<std::vector<uint64_t>> rpool(live.size());
    {
      DevBuf d(pool * sizeof(ecl_query));
      for (size_t j = 0; j < live.size(); ++j) {
        ecl_wavelet_tree wt;
        {
          std::lock_guard<std::mutex> lk(s->mu);
          wt = s->trees[live[j]].wt;
        }
        const ecl_status st =
            ecl_wavelet_tree_generate_queries(wt, 3, spec->seed + 7919 * j, pool, d.ptr, nullptr);
        if (st != ECL_OK) throw Error(st, last_error());
        qpool[j].resize(pool);
        rpool[j].resize(pool);
        ECL_HIP(hipDeviceSynchronize());
        d2h(qpool[j].data(), d.ptr, pool);
      }
    }
    ECL_HIP(hipDeviceSynchronize());
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->lat_ms.clear();
      s->st_queries = 0;
      s->st_batches = 0;
      s->st_any = false;
    }
    // open-loop Poisson arrivals (splitmix64-driven, seeded)
    uint64_t state = spec->seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
      state += 0x9E3779B97F4A7C15ull;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    };
    auto unit = [&]() { return (static_cast<double>(next() >> 11) + 0.5) * 0x1.0p-53; };
    const double lmin = std::log(static_cast<double>(spec->min_size));
    const double lmax = std::log(static_cast<double>(spec->max_size) + 1.0);
    const Clock::time_point t0 = Clock::now();
    double t = 0.0;
    uint64_t submitted = 0;
    while (true) {
      t += -std::log(unit()) / spec->rate;
      if (t >= spec->seconds) break;
      const Clock::time_point due =
          t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t));
      std::this_thread::sleep_until(due);
      const size_t j = static_cast<size_t>(next() % live.size());
      uint64_t size = static_cast<uint64_t>(std::exp(lmin + (lmax - lmin) * unit()));
      size = std::min<uint64_t>(std::max<uint64_t>(size, spec->min_size), spec->max_size);
      const uint64_t off = next() % (pool - size + 1);
      const ecl_status st = submit_impl(s, live[j], qpool[j].data() + off, size,
                                        rpool[j].data() + off, true, nullptr);
      if (st != ECL_OK) throw Error(st, last_error());
      ++submitted;
    }
    // wait for the run to drain
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_wait.wait(lk, [&] { return s->lat_ms.size() >= submitted; });
    }
    const ecl_status st = ecl_server_stats_get(s, out);
    if (st != ECL_OK) throw Error(st, last_error());
  });
}

}  // extern "C"
