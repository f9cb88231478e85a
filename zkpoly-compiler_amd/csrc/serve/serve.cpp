// libzkp_serve.so -- continuous-batching server for proof-step requests on one GPU.
//
// Design reference: the reference's multi-request scheduler
// (scheduler/src/scheduler.rs:215-310, scheduler/src/scheduler/core.rs:197-383),
// which admits whole compiled programs by host-memory budget and runs each on
// one card.  Here the unit is one proof step -- an NTT or an MSM request -- and
// admission is by HBM footprint:
//
//  * resident state: twiddle tables for every NTT size the server accepts and
//    one MSM base table (requests of 2^k points use its first 2^k bases);
//  * a dispatcher thread releases requests at their arrival times into a FIFO;
//  * `slots` worker threads, each with its own HIP stream, take requests in
//    arrival order; a worker first reserves the request's HBM footprint from a
//    first-fit pool carved out of one device allocation (waiting while it does
//    not fit), copies the input host->device, runs the kernels through the
//    same C ABI as single calls (ssip_ntt, zkp_msm_device), copies the result
//    back and frees the footprint -- so requests that arrive while others run
//    are admitted as soon as a slot and memory free up, and their kernels
//    overlap on the device;
//  * per request the server records arrival, start (admission) and finish
//    times, from which the caller computes p50 / p99 latency and throughput.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../common/poly_ptr.hpp"

extern "C" {
hipError_t ssip_ntt(PolyPtr x, const unsigned int* twiddle, unsigned int log_len, hipStream_t stream,
                    const unsigned int, const unsigned int);
unsigned long long zkp_msm_workspace_size(unsigned long long len, unsigned int window, unsigned int target_windows);
hipError_t zkp_msm_device(const unsigned int* d_points, const unsigned int* d_scalars, unsigned long long len,
                          unsigned int window, unsigned int target_windows, void* d_workspace,
                          unsigned long long workspace_bytes, hipStream_t stream, unsigned int* h_result, int debug);
}

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, size) with 256-byte granularity.
class Pool {
 public:
  explicit Pool(size_t size) : size_(size) { free_[0] = size; }
  // returns offset or SIZE_MAX
  size_t alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return size_; }

 private:
  size_t size_;
  std::map<size_t, size_t> free_;
};

}  // namespace

extern "C" {

// Server configuration.
typedef struct ZkpServeConfig {
  unsigned int slots;                 // worker threads / streams
  unsigned long long hbm_budget;      // bytes of request memory (pool size)
  unsigned int ntt_min_log, ntt_max_log;
  unsigned int msm_max_log;           // resident base table holds 2^msm_max_log points
  unsigned int msm_window;            // window for every MSM request
  const unsigned int* msm_bases_host; // 2^msm_max_log affine points (host), copied once
  const unsigned int* const* twiddles_host;  // [log - ntt_min_log] -> 2^(log-1) elements (host)
} ZkpServeConfig;

// One request.  kind 0: in-place forward NTT of 2^log_n elements from input
// (host) to output (host, 2^log_n elements); kind 1: MSM of 2^log_n scalars
// (host input) with the resident bases, output = affine point (16 words).
typedef struct ZkpServeRequest {
  unsigned int kind;
  unsigned int log_n;
  double arrival_s;  // offset from the start of the run
  const unsigned int* input;
  unsigned int* output;
  // filled by the server
  double start_s;
  double finish_s;
  int status;  // hipError_t
} ZkpServeRequest;

}  // extern "C"

namespace {

struct Engine {
  ZkpServeConfig cfg{};
  int device = 0;  // the device current at zkp_serve_create; workers run on it
  void* pool_base = nullptr;
  Pool* pool = nullptr;
  std::vector<uint32_t*> d_twiddles;  // per log - ntt_min_log
  uint32_t* d_bases = nullptr;
  std::vector<hipStream_t> streams;

  std::mutex mu;
  std::condition_variable cv_queue, cv_mem;
  std::deque<ZkpServeRequest*> queue;
  bool dispatch_done = false;

  ~Engine() {
    for (auto s : streams) (void)hipStreamDestroy(s);
    for (auto t : d_twiddles)
      if (t) (void)hipFree(t);
    if (d_bases) (void)hipFree(d_bases);
    if (pool_base) (void)hipFree(pool_base);
    delete pool;
  }

  size_t footprint(const ZkpServeRequest& r) const {
    const size_t n = 1ull << r.log_n;
    if (r.kind == 0) return n * 32;
    return n * 32 + zkp_msm_workspace_size(n, cfg.msm_window, (254 + cfg.msm_window - 1) / cfg.msm_window) + 256;
  }

  // blocks until the footprint fits; returns the pool offset
  size_t acquire(size_t bytes) {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      const size_t off = pool->alloc(bytes);
      if (off != SIZE_MAX) return off;
      cv_mem.wait(lk);
    }
  }
  void release(size_t off, size_t bytes) {
    {
      std::lock_guard<std::mutex> lk(mu);
      pool->release(off, bytes);
    }
    cv_mem.notify_all();
  }

  hipError_t process(ZkpServeRequest& r, hipStream_t s) {
    const size_t n = 1ull << r.log_n;
    const size_t bytes = footprint(r);
    const size_t off = acquire(bytes);
    r.start_s = now();
    char* base = reinterpret_cast<char*>(pool_base) + off;
    hipError_t err;
    if (r.kind == 0) {
      uint32_t* x = reinterpret_cast<uint32_t*>(base);
      err = hipMemcpyAsync(x, r.input, n * 32, hipMemcpyHostToDevice, s);
      if (err == hipSuccess) {
        PolyPtr p{x, n, 0, 0, n};
        err = ssip_ntt(p, d_twiddles[r.log_n - cfg.ntt_min_log], r.log_n, s, 8, 8);
      }
      if (err == hipSuccess) err = hipMemcpyAsync(r.output, x, n * 32, hipMemcpyDeviceToHost, s);
    } else {
      uint32_t* sc = reinterpret_cast<uint32_t*>(base);
      char* ws = base + ((n * 32 + 255) & ~(size_t)255);
      const unsigned target = (254 + cfg.msm_window - 1) / cfg.msm_window;
      const size_t ws_bytes = zkp_msm_workspace_size(n, cfg.msm_window, target);
      err = hipMemcpyAsync(sc, r.input, n * 32, hipMemcpyHostToDevice, s);
      if (err == hipSuccess)
        err = zkp_msm_device(d_bases, sc, n, cfg.msm_window, target, ws, ws_bytes, s, r.output, 0);
    }
    // Whatever failed, work already queued on s may still read or write the
    // region: drain the stream before the region goes back to the pool.
    const hipError_t sync = hipStreamSynchronize(s);
    if (err == hipSuccess) err = sync;
    r.finish_s = now();
    release(off, bytes);
    return err;
  }

  Clock::time_point t0;
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

}  // namespace

extern "C" {

// Creates a server: copies the resident twiddles and MSM bases to the device
// and allocates the request pool.  Returns hipErrorInvalidValue for a bad
// configuration.
hipError_t zkp_serve_create(const ZkpServeConfig* cfg, void** out) {
  if (!cfg || !out || cfg->slots == 0 || cfg->ntt_min_log == 0 || cfg->ntt_min_log > cfg->ntt_max_log ||
      cfg->ntt_max_log > 28 || cfg->msm_window < 2 || cfg->msm_window > 24 || !cfg->twiddles_host)
    return hipErrorInvalidValue;
  Engine* e = new Engine();
  e->cfg = *cfg;
  hipError_t err = hipGetDevice(&e->device);
  if (err == hipSuccess) err = hipMalloc(&e->pool_base, cfg->hbm_budget);
  if (err == hipSuccess) e->pool = new Pool(cfg->hbm_budget);
  for (unsigned k = cfg->ntt_min_log; k <= cfg->ntt_max_log && err == hipSuccess; k++) {
    uint32_t* t = nullptr;
    const size_t bytes = (1ull << (k - 1)) * 32;
    err = hipMalloc(&t, bytes);
    e->d_twiddles.push_back(t);
    if (err == hipSuccess) err = hipMemcpy(t, cfg->twiddles_host[k - cfg->ntt_min_log], bytes, hipMemcpyHostToDevice);
  }
  if (err == hipSuccess && cfg->msm_max_log > 0 && cfg->msm_bases_host) {
    const size_t bytes = (1ull << cfg->msm_max_log) * 64;
    err = hipMalloc(&e->d_bases, bytes);
    if (err == hipSuccess) err = hipMemcpy(e->d_bases, cfg->msm_bases_host, bytes, hipMemcpyHostToDevice);
  }
  for (unsigned i = 0; i < cfg->slots && err == hipSuccess; i++) {
    hipStream_t s;
    err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (err == hipSuccess) e->streams.push_back(s);
  }
  if (err != hipSuccess) {
    delete e;
    return err;
  }
  *out = e;
  return hipSuccess;
}

// Largest footprint (bytes) a request of this kind and size reserves.
unsigned long long zkp_serve_footprint(void* engine, unsigned int kind, unsigned int log_n) {
  Engine* e = reinterpret_cast<Engine*>(engine);
  ZkpServeRequest r{};
  r.kind = kind;
  r.log_n = log_n;
  return e ? e->footprint(r) : 0;
}

// Runs a trace: request i is released at arrival_s (seconds after the call
// starts) and served by the first free slot with room in the pool.  Blocks
// until all requests finished; fills start_s / finish_s / status.  Requests
// must be sorted by arrival_s and validated: NTT sizes within the configured
// range, MSM sizes <= msm_max_log, footprints <= hbm_budget.
hipError_t zkp_serve_run(void* engine, ZkpServeRequest* reqs, unsigned int count) {
  Engine* e = reinterpret_cast<Engine*>(engine);
  if (!e || (count && !reqs)) return hipErrorInvalidValue;
  for (unsigned i = 0; i < count; i++) {
    const ZkpServeRequest& r = reqs[i];
    const bool ntt_ok = r.kind == 0 && r.log_n >= e->cfg.ntt_min_log && r.log_n <= e->cfg.ntt_max_log;
    const bool msm_ok = r.kind == 1 && r.log_n <= e->cfg.msm_max_log && e->d_bases;
    if ((!ntt_ok && !msm_ok) || !r.input || !r.output || e->footprint(r) > e->cfg.hbm_budget ||
        (i && reqs[i].arrival_s < reqs[i - 1].arrival_s))
      return hipErrorInvalidValue;
  }
  e->t0 = Clock::now();
  e->dispatch_done = false;
  std::atomic<int> first_error{0};
  std::vector<std::thread> workers;
  for (unsigned w = 0; w < e->cfg.slots; w++) {
    workers.emplace_back([e, w, &first_error] {
      if (hipSetDevice(e->device) != hipSuccess) {
        int expected = 0;
        first_error.compare_exchange_strong(expected, (int)hipErrorInvalidDevice);
      }
      while (true) {
        ZkpServeRequest* r = nullptr;
        {
          std::unique_lock<std::mutex> lk(e->mu);
          e->cv_queue.wait(lk, [e] { return !e->queue.empty() || e->dispatch_done; });
          if (e->queue.empty()) return;
          r = e->queue.front();
          e->queue.pop_front();
        }
        r->status = e->process(*r, e->streams[w]);
        if (r->status != hipSuccess) {
          int expected = 0;
          first_error.compare_exchange_strong(expected, r->status);
        }
      }
    });
  }
  // dispatcher: release each request at its arrival time
  for (unsigned i = 0; i < count; i++) {
    const double wait = reqs[i].arrival_s - e->now();
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    {
      std::lock_guard<std::mutex> lk(e->mu);
      e->queue.push_back(&reqs[i]);
    }
    e->cv_queue.notify_one();
  }
  {
    std::lock_guard<std::mutex> lk(e->mu);
    e->dispatch_done = true;
  }
  e->cv_queue.notify_all();
  for (auto& t : workers) t.join();
  return (hipError_t)first_error.load();
}

void zkp_serve_destroy(void* engine) { delete reinterpret_cast<Engine*>(engine); }

}  // extern "C"
