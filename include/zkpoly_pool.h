/* libzkp_pinned_pool.so -- the runtime's host memory pool: a buddy allocator
 * over slabs of pinned host memory (hipHostMalloc) for the CPU-side arrays
 * the runtime copies to and from the GPU (polynomials, MSM points and
 * scalars), so that those copies run asynchronously at the link rate.
 *
 * Replaces the reference's CpuMemoryPool (memory_pool/src/cpu_pool.rs:21-205)
 * over MemoryPool (memory_pool/src/memory_pool.rs:153-330,440-536), whose
 * slabs come from cudaMallocHost.  Same semantics: requests are rounded up to
 * base_size * 2^k, k <= max_log_factor; slabs of base_size * 2^max_log_factor
 * bytes are taken from the system on demand and split in halves; freed blocks
 * merge with their buddies; shrink returns wholly free slabs, clear all slabs;
 * backing 1 is the reference's use_mmap (file-backed slabs).  Errors are
 * returned (hipErrorInvalidValue for bad arguments, a foreign or doubly freed
 * pointer, or a request above the largest class) where the reference panics.
 * Thread-safe.
 */
#ifndef ZKPOLY_POOL_H
#define ZKPOLY_POOL_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* library: libzkp_pinned_pool.so */

typedef struct zkp_pinned_pool zkp_pinned_pool;

/* CpuMemoryPool::new(max_log_factor, base_size) (cpu_pool.rs:21-27); backing
 * 0 = pinned host memory, 1 = file-backed mappings under tmp_dir (NULL:
 * $TMPDIR or /tmp; CpuMemoryPool::use_mmap, cpu_pool.rs:33-36). */
hipError_t zkp_pinned_pool_create(unsigned int max_log_factor, unsigned long long base_size, unsigned int backing,
                                  const char* tmp_dir, zkp_pinned_pool** out);

/* CpuMemoryPool::allocate (cpu_pool.rs:45-114): a block of at least `bytes`. */
hipError_t zkp_pinned_pool_alloc(zkp_pinned_pool* pool, unsigned long long bytes, void** out);

/* MemoryPool::allocate(log_factor) (memory_pool.rs:440-466). */
hipError_t zkp_pinned_pool_alloc_log(zkp_pinned_pool* pool, unsigned int log_factor, void** out);

/* CpuMemoryPool::free / MemoryPool::deallocate (cpu_pool.rs:164-190,
 * memory_pool.rs:505-536), merging buddies. */
hipError_t zkp_pinned_pool_free(zkp_pinned_pool* pool, void* ptr);

/* Size class in bytes of a live block, 0 if the pool does not hold ptr. */
unsigned long long zkp_pinned_pool_block_size(zkp_pinned_pool* pool, const void* ptr);

/* CpuMemoryPool::preallocate (cpu_pool.rs:116-160): reserves whole slabs. */
hipError_t zkp_pinned_pool_preallocate(zkp_pinned_pool* pool, unsigned long long slabs);

/* CpuMemoryPool::shrink / MemoryPool::shrink (memory_pool.rs:575-690). */
hipError_t zkp_pinned_pool_shrink(zkp_pinned_pool* pool);

/* CpuMemoryPool::clear / MemoryPool::clear (memory_pool.rs:540-572). */
hipError_t zkp_pinned_pool_clear(zkp_pinned_pool* pool);

/* MemoryPool::used_bytes (memory_pool.rs:692-714): slabs held x slab size. */
unsigned long long zkp_pinned_pool_used_bytes(zkp_pinned_pool* pool);

/* Number of free blocks of class k (inspection for tests). */
unsigned long long zkp_pinned_pool_free_blocks(zkp_pinned_pool* pool, unsigned int k);

/* clear(), then releases the pool. */
void zkp_pinned_pool_destroy(zkp_pinned_pool* pool);

#ifdef __cplusplus
}
#endif

#endif /* ZKPOLY_POOL_H */
