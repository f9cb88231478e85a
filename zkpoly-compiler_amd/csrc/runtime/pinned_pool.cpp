// libzkp_pinned_pool.so -- host memory pool for the runtime's CPU-side arrays:
// a buddy allocator over large slabs of pinned (page-locked, device-mapped)
// host memory from hipHostMalloc, so that host<->device copies of polynomials,
// MSM points and scalars run asynchronously at the link rate.
//
// Behaviour of the reference's CpuMemoryPool / MemoryPool
// (memory_pool/src/cpu_pool.rs:21-205, memory_pool.rs:153-330,440-536):
//  * every request is rounded up to base_size * 2^k (k <= max_log_factor);
//  * a slab of base_size * 2^max_log_factor bytes is obtained from the system
//    when no free block of the requested class or above exists, and split in
//    halves down to the class;
//  * a freed block merges with its buddy when both are free, repeatedly;
//  * shrink() returns whole free slabs to the system, clear() all of them;
//  * preallocate(n) reserves n slabs up front;
//  * used_bytes() is the memory held from the system (slabs x slab size);
//  * the "mmap" backing (the reference's use_mmap) places slabs in unlinked
//    files under a temporary directory instead of pinned memory.
// Bookkeeping is this library's own: one ordered free set per size class and
// a map from every live allocation to its class; a block's buddy is found by
// flipping the class bit of its offset inside its slab.  All calls take one
// mutex: the runtime calls the pool from several threads.
#include <hip/hip_runtime_api.h>

#include <fcntl.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <unistd.h>

#include <map>
#include <mutex>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

enum Backing : unsigned { kPinned = 0, kMmap = 1 };

struct Slab {
  char* base;
  int fd;  // mmap backing only (-1 otherwise)
};

struct Pool {
  unsigned max_log;
  size_t base_size;
  size_t slab_bytes;
  Backing backing;
  std::string tmp_dir;  // mmap backing
  unsigned file_seq = 0;
  std::mutex mu;
  std::map<char*, Slab> slabs;                  // by base address
  std::vector<std::set<char*>> free_blocks;     // per class k: free block addresses
  std::unordered_map<char*, unsigned> live;     // allocation -> class
};

hipError_t slab_obtain(Pool& p, Slab* out) {
  if (p.backing == kPinned) {
    void* ptr = nullptr;
    const hipError_t e = hipHostMalloc(&ptr, p.slab_bytes, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    if (ptr == nullptr) return hipErrorOutOfMemory;
    *out = Slab{static_cast<char*>(ptr), -1};
    return hipSuccess;
  }
  const std::string path = p.tmp_dir + "/slab_" + std::to_string(p.file_seq++);
  const int fd = open(path.c_str(), O_RDWR | O_CREAT | O_EXCL, 0600);
  if (fd < 0) return hipErrorOutOfMemory;
  unlink(path.c_str());  // the mapping keeps the file alive; nothing is left on disk afterwards
  if (ftruncate(fd, (off_t)p.slab_bytes) != 0) {
    close(fd);
    return hipErrorOutOfMemory;
  }
  void* m = mmap(nullptr, p.slab_bytes, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
  if (m == MAP_FAILED) {
    close(fd);
    return hipErrorOutOfMemory;
  }
  *out = Slab{static_cast<char*>(m), fd};
  return hipSuccess;
}

hipError_t slab_release(Pool& p, const Slab& s) {
  if (p.backing == kPinned) return hipHostFree(s.base);
  munmap(s.base, p.slab_bytes);
  close(s.fd);
  return hipSuccess;
}

// The slab holding address a (a lies in [base, base + slab_bytes)).
const Slab* slab_of(const Pool& p, char* a) {
  auto it = p.slabs.upper_bound(a);
  if (it == p.slabs.begin()) return nullptr;
  --it;
  if (a >= it->first + p.slab_bytes) return nullptr;
  return &it->second;
}

// A free block of class k, splitting a larger one or taking a new slab.
hipError_t take_block(Pool& p, unsigned k, char** out) {
  unsigned j = k;
  while (j <= p.max_log && p.free_blocks[j].empty()) j++;
  if (j > p.max_log) {
    Slab s;
    const hipError_t e = slab_obtain(p, &s);
    if (e != hipSuccess) return e;
    p.slabs.emplace(s.base, s);
    p.free_blocks[p.max_log].insert(s.base);
    j = p.max_log;
  }
  char* b = *p.free_blocks[j].begin();
  p.free_blocks[j].erase(p.free_blocks[j].begin());
  // split down to class k: keep the lower half, free the upper one
  while (j > k) {
    j--;
    p.free_blocks[j].insert(b + (p.base_size << j));
  }
  *out = b;
  return hipSuccess;
}

// Returns block b of class k, merging with free buddies.
void give_block(Pool& p, char* b, unsigned k) {
  const Slab* s = slab_of(p, b);
  char* base = s->base;
  while (k < p.max_log) {
    const size_t off = (size_t)(b - base);
    char* buddy = base + (off ^ (p.base_size << k));
    auto it = p.free_blocks[k].find(buddy);
    if (it == p.free_blocks[k].end()) break;
    p.free_blocks[k].erase(it);
    b = buddy < b ? buddy : b;
    k++;
  }
  p.free_blocks[k].insert(b);
}

unsigned class_of(const Pool& p, size_t bytes) {
  const size_t chunks = (bytes + p.base_size - 1) / p.base_size;
  unsigned k = 0;
  while (((size_t)1 << k) < chunks) k++;
  return k;
}

}  // namespace

extern "C" {

typedef struct zkp_pinned_pool zkp_pinned_pool;

// New pool of slabs of base_size * 2^max_log_factor bytes.  backing 0: pinned
// host memory (hipHostMalloc); 1: file-backed mappings under tmp_dir (NULL:
// $TMPDIR or /tmp), the reference's use_mmap.  Nothing is allocated yet.
hipError_t zkp_pinned_pool_create(unsigned int max_log_factor, unsigned long long base_size, unsigned int backing,
                                  const char* tmp_dir, zkp_pinned_pool** out) {
  if (out == nullptr || base_size == 0 || max_log_factor > 40 || backing > kMmap) return hipErrorInvalidValue;
  if (base_size > (~0ull >> (max_log_factor + 1))) return hipErrorInvalidValue;  // slab size overflows
  auto* p = new Pool;
  p->max_log = max_log_factor;
  p->base_size = (size_t)base_size;
  p->slab_bytes = (size_t)base_size << max_log_factor;
  p->backing = (Backing)backing;
  p->free_blocks.resize(max_log_factor + 1);
  if (p->backing == kMmap) {
    const char* root = tmp_dir ? tmp_dir : getenv("TMPDIR");
    std::string tmpl = std::string(root && *root ? root : "/tmp") + "/zkp_pool_XXXXXX";
    std::vector<char> buf(tmpl.begin(), tmpl.end());
    buf.push_back('\0');
    if (mkdtemp(buf.data()) == nullptr) {
      delete p;
      return hipErrorInvalidValue;
    }
    p->tmp_dir = buf.data();
  }
  *out = reinterpret_cast<zkp_pinned_pool*>(p);
  return hipSuccess;
}

// Block of at least `bytes` (rounded up to base_size * 2^k); bytes == 0 or a
// class above max_log_factor is refused.
hipError_t zkp_pinned_pool_alloc(zkp_pinned_pool* pool, unsigned long long bytes, void** out) {
  if (pool == nullptr || out == nullptr || bytes == 0) return hipErrorInvalidValue;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  if (bytes > p.slab_bytes) return hipErrorInvalidValue;
  const unsigned k = class_of(p, (size_t)bytes);
  std::lock_guard<std::mutex> lock(p.mu);
  char* b = nullptr;
  const hipError_t e = take_block(p, k, &b);
  if (e != hipSuccess) return e;
  p.live.emplace(b, k);
  *out = b;
  return hipSuccess;
}

// Block of exactly base_size * 2^log_factor bytes (MemoryPool::allocate).
hipError_t zkp_pinned_pool_alloc_log(zkp_pinned_pool* pool, unsigned int log_factor, void** out) {
  if (pool == nullptr || out == nullptr) return hipErrorInvalidValue;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  if (log_factor > p.max_log) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(p.mu);
  char* b = nullptr;
  const hipError_t e = take_block(p, log_factor, &b);
  if (e != hipSuccess) return e;
  p.live.emplace(b, log_factor);
  *out = b;
  return hipSuccess;
}

// Returns a block; a pointer the pool did not hand out (or one already
// returned) is refused.
hipError_t zkp_pinned_pool_free(zkp_pinned_pool* pool, void* ptr) {
  if (pool == nullptr || ptr == nullptr) return hipErrorInvalidValue;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  std::lock_guard<std::mutex> lock(p.mu);
  auto it = p.live.find(static_cast<char*>(ptr));
  if (it == p.live.end()) return hipErrorInvalidValue;
  const unsigned k = it->second;
  p.live.erase(it);
  give_block(p, static_cast<char*>(ptr), k);
  return hipSuccess;
}

// Size class of a live block in bytes, or 0 for a pointer the pool does not hold.
unsigned long long zkp_pinned_pool_block_size(zkp_pinned_pool* pool, const void* ptr) {
  if (pool == nullptr || ptr == nullptr) return 0;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  std::lock_guard<std::mutex> lock(p.mu);
  auto it = p.live.find(static_cast<char*>(const_cast<void*>(ptr)));
  return it == p.live.end() ? 0 : (unsigned long long)(p.base_size << it->second);
}

// Reserves `slabs` whole slabs (CpuMemoryPool::preallocate); they stay free.
hipError_t zkp_pinned_pool_preallocate(zkp_pinned_pool* pool, unsigned long long slabs) {
  if (pool == nullptr) return hipErrorInvalidValue;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  std::lock_guard<std::mutex> lock(p.mu);
  for (unsigned long long i = 0; i < slabs; i++) {
    Slab s;
    const hipError_t e = slab_obtain(p, &s);
    if (e != hipSuccess) return e;
    p.slabs.emplace(s.base, s);
    p.free_blocks[p.max_log].insert(s.base);
  }
  return hipSuccess;
}

// Returns every wholly free slab to the system.
hipError_t zkp_pinned_pool_shrink(zkp_pinned_pool* pool) {
  if (pool == nullptr) return hipErrorInvalidValue;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  std::lock_guard<std::mutex> lock(p.mu);
  hipError_t first = hipSuccess;
  auto& top = p.free_blocks[p.max_log];
  for (auto it = top.begin(); it != top.end();) {
    auto s = p.slabs.find(*it);
    const hipError_t e = slab_release(p, s->second);
    if (e != hipSuccess && first == hipSuccess) first = e;
    p.slabs.erase(s);
    it = top.erase(it);
  }
  return first;
}

// Returns every slab to the system; outstanding blocks become invalid.
hipError_t zkp_pinned_pool_clear(zkp_pinned_pool* pool) {
  if (pool == nullptr) return hipErrorInvalidValue;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  std::lock_guard<std::mutex> lock(p.mu);
  hipError_t first = hipSuccess;
  for (auto& kv : p.slabs) {
    const hipError_t e = slab_release(p, kv.second);
    if (e != hipSuccess && first == hipSuccess) first = e;
  }
  p.slabs.clear();
  p.live.clear();
  for (auto& f : p.free_blocks) f.clear();
  return first;
}

// Bytes held from the system (slabs x slab size).
unsigned long long zkp_pinned_pool_used_bytes(zkp_pinned_pool* pool) {
  if (pool == nullptr) return 0;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  std::lock_guard<std::mutex> lock(p.mu);
  return (unsigned long long)p.slabs.size() * p.slab_bytes;
}

// Free blocks of class k (for tests of splitting and merging).
unsigned long long zkp_pinned_pool_free_blocks(zkp_pinned_pool* pool, unsigned int k) {
  if (pool == nullptr) return 0;
  Pool& p = *reinterpret_cast<Pool*>(pool);
  std::lock_guard<std::mutex> lock(p.mu);
  return k <= p.max_log ? (unsigned long long)p.free_blocks[k].size() : 0;
}

// clear(), then the pool itself.
void zkp_pinned_pool_destroy(zkp_pinned_pool* pool) {
  if (pool == nullptr) return;
  (void)zkp_pinned_pool_clear(pool);
  Pool* p = reinterpret_cast<Pool*>(pool);
  if (!p->tmp_dir.empty()) rmdir(p->tmp_dir.c_str());
  delete p;
}

}  // extern "C"
