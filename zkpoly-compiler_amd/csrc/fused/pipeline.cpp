// libzkp_fused_pipeline.so -- host-chunked streaming of a fused kernel over
// host-resident polynomials.
//
// The native driver behind PipelinedFusedKernel (the reference's closure in
// core/src/fused_kernels.rs:198-516, meta runtime/src/functions.rs:37-50).
// The polynomials live in host memory; the kernel runs on divide_parts chunks
// of chunk_len = len / divide_parts elements each, in caller-provided device
// buffers: two per read-only polynomial (one loading, one computing) and three
// per mutable polynomial (loading, computing, storing back).  Three streams of
// default priority carry the copies in, the kernels and the copies out;
// events order them exactly as the reference does:
//   copy-in of mutable chunk c (skipped for polynomials the kernel only
//   writes) waits for copy-out of the chunk that last used its buffer (c - 3);
//   copy-in of read-only chunk c waits for the kernel that last used its
//   buffer (c - 2); the kernel waits for both copy-ins; the copy-out waits for
//   the kernel.
// Scalars travel once: read-only and mutable scalars in before the first
// chunk, mutable scalars back after the last.  The kernel sees is_first only
// on chunk 0 (coefficient-form rules index the global element 0).  The call is
// synchronous: it returns after the last copy-out, like the reference.
//
// Host polynomials may be rotated (the reference refuses slices, and so does
// this driver); a chunk whose physical range wraps is copied in two pieces.
// For copies that overlap the kernels the host memory must be pinned
// (hipHostMalloc); pageable memory still gives the right result.  When every
// polynomial sits in mapped pinned memory the kernel instead runs once over the
// host arrays directly; otherwise mutable polynomials the kernel only writes,
// when mapped, are written by the kernel directly (see below).
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

extern "C" {
typedef struct PolyPtr {
  uint32_t* ptr;
  size_t len;
  size_t rotate;
  size_t offset;
  size_t whole_len;
} PolyPtr;
typedef struct ConstPolyPtr {
  const uint32_t* ptr;
  size_t len;
  size_t rotate;
  size_t offset;
  size_t whole_len;
} ConstPolyPtr;
typedef hipError_t (*zkp_fused_fn)(const ConstPolyPtr* vars, const PolyPtr* mut_vars, unsigned long long len,
                                   bool is_first, void* local_buffer, hipStream_t stream);
}

namespace {

constexpr size_t ELEM = 32;

struct Err {
  hipError_t first = hipSuccess;
  void operator()(hipError_t e) {
    if (first == hipSuccess && e != hipSuccess) first = e;
  }
};

// Copies logical elements [start, start + count) of a rotated host array of
// `len` elements (logical i at physical (i + len - rotate) % len) to `dev`.
void copy_in(Err& err, const uint32_t* host, size_t len, size_t rotate, size_t start, size_t count, uint32_t* dev,
             hipStream_t s) {
  size_t phys = (start + len - rotate) % len;
  const size_t first = count < len - phys ? count : len - phys;
  err(hipMemcpyAsync(dev, host + phys * 8, first * ELEM, hipMemcpyHostToDevice, s));
  if (first < count) err(hipMemcpyAsync(dev + first * 8, host, (count - first) * ELEM, hipMemcpyHostToDevice, s));
}

void copy_out(Err& err, uint32_t* host, size_t len, size_t rotate, size_t start, size_t count, const uint32_t* dev,
              hipStream_t s) {
  size_t phys = (start + len - rotate) % len;
  const size_t first = count < len - phys ? count : len - phys;
  err(hipMemcpyAsync(host + phys * 8, dev, first * ELEM, hipMemcpyDeviceToHost, s));
  if (first < count) err(hipMemcpyAsync(host, dev + first * 8, (count - first) * ELEM, hipMemcpyDeviceToHost, s));
}

// A host polynomial accepted by the driver: unsliced, `len` elements.
bool plain(const uint32_t* ptr, size_t len, size_t rotate, size_t offset, size_t whole_len, size_t want) {
  return ptr != nullptr && len == want && offset == 0 && (whole_len == 0 || whole_len == len) &&
         (rotate == 0 || rotate < len);
}

// Device address of a host array of `bytes` bytes that kernels can write
// directly (pinned and mapped, as hipHostMalloc memory is), or nullptr: both
// ends of the range must lie in registered host memory at device addresses
// exactly `bytes - 1` apart.
uint32_t* mapped_device_ptr(uint32_t* host, size_t bytes) {
  hipPointerAttribute_t a{}, b{};
  const char* last = reinterpret_cast<const char*>(host) + bytes - 1;
  if (hipPointerGetAttributes(&a, host) != hipSuccess || hipPointerGetAttributes(&b, last) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory: not an error for the caller
    return nullptr;
  }
  if (a.type != hipMemoryTypeHost || b.type != hipMemoryTypeHost || !a.devicePointer || !b.devicePointer)
    return nullptr;
  const char* da = reinterpret_cast<const char*>(a.devicePointer);
  const char* db = reinterpret_cast<const char*>(b.devicePointer);
  return db - da == (ptrdiff_t)(bytes - 1) ? reinterpret_cast<uint32_t*>(a.devicePointer) : nullptr;
}

// Events that may be waited on before their first record are skipped until
// then (a wait on a never-recorded event is a no-op in the reference's CUDA).
struct Ev {
  hipEvent_t e = nullptr;
  bool recorded = false;
};

// The whole length in one launch over mapped host arrays (see the caller).
hipError_t run_direct(zkp_fused_fn kernel, unsigned num_scalars, const uint32_t* const* host_scalars,
                      uint32_t* const* dev_scalars, unsigned num_mut_scalars, uint32_t* const* host_mut_scalars,
                      uint32_t* const* dev_mut_scalars, unsigned num_polys, const ConstPolyPtr* host_polys,
                      const uint32_t* const* in_dev, unsigned num_mut_polys, const PolyPtr* host_mut_polys,
                      uint32_t* const* out_dev, size_t len, void* arg_buffer, void* local_buffer) {
  const unsigned nv = num_scalars + num_polys, nm = num_mut_scalars + num_mut_polys;
  Err err;
  hipStream_t s = nullptr;
  err(hipStreamCreate(&s));
  if (err.first != hipSuccess) return err.first;
  std::vector<ConstPolyPtr> vars(nv);
  std::vector<PolyPtr> muts(nm);
  for (unsigned i = 0; i < num_scalars; i++) vars[i] = ConstPolyPtr{dev_scalars[i], 1, 0, 0, 1};
  for (unsigned i = 0; i < num_polys; i++)
    vars[num_scalars + i] = ConstPolyPtr{in_dev[i], len, host_polys[i].rotate, 0, len};
  for (unsigned i = 0; i < num_mut_scalars; i++) muts[i] = PolyPtr{dev_mut_scalars[i], 1, 0, 0, 1};
  for (unsigned i = 0; i < num_mut_polys; i++)
    muts[num_mut_scalars + i] = PolyPtr{out_dev[i], len, host_mut_polys[i].rotate, 0, len};
  ConstPolyPtr* d_vars = reinterpret_cast<ConstPolyPtr*>(arg_buffer);
  PolyPtr* d_muts = reinterpret_cast<PolyPtr*>(d_vars + nv);
  if (nv) err(hipMemcpyAsync(d_vars, vars.data(), nv * sizeof(ConstPolyPtr), hipMemcpyHostToDevice, s));
  if (nm) err(hipMemcpyAsync(d_muts, muts.data(), nm * sizeof(PolyPtr), hipMemcpyHostToDevice, s));
  for (unsigned i = 0; i < num_mut_scalars; i++)
    err(hipMemcpyAsync(dev_mut_scalars[i], host_mut_scalars[i], ELEM, hipMemcpyHostToDevice, s));
  for (unsigned i = 0; i < num_scalars; i++)
    err(hipMemcpyAsync(dev_scalars[i], host_scalars[i], ELEM, hipMemcpyHostToDevice, s));
  if (err.first == hipSuccess) err(kernel(d_vars, d_muts, len, true, local_buffer, s));
  for (unsigned i = 0; i < num_mut_scalars && err.first == hipSuccess; i++)
    err(hipMemcpyAsync(host_mut_scalars[i], dev_mut_scalars[i], ELEM, hipMemcpyDeviceToHost, s));
  err(hipStreamSynchronize(s));
  err(hipStreamDestroy(s));
  return err.first;
}

}  // namespace

extern "C" {

// Runs `kernel` over host polynomials in divide_parts chunks (see above).
// Device buffers: dev_polys[j * num_polys + i] (j < 2) and
// dev_mut_polys[j * num_mut_polys + i] (j < 3) of chunk_len elements each;
// dev_scalars / dev_mut_scalars of one element each.  mut_poly_read[i] == 0
// marks a mutable polynomial the kernel only writes (every element of every
// chunk), whose chunks are then not copied in; NULL copies every one in, as
// the reference does.  arg_buffer holds
// (2 * (num_scalars + num_polys) + 3 * (num_mut_scalars + num_mut_polys))
// PolyPtr; local_buffer is passed to the kernel.  hipErrorInvalidValue, before
// any work, when a precondition of the reference fails.
hipError_t zkp_fused_pipelined(zkp_fused_fn kernel, unsigned int num_scalars, const uint32_t* const* host_scalars,
                               uint32_t* const* dev_scalars, unsigned int num_mut_scalars,
                               uint32_t* const* host_mut_scalars, uint32_t* const* dev_mut_scalars,
                               unsigned int num_polys, const ConstPolyPtr* host_polys, uint32_t* const* dev_polys,
                               unsigned int num_mut_polys, const PolyPtr* host_mut_polys,
                               uint32_t* const* dev_mut_polys, const unsigned char* mut_poly_read,
                               unsigned long long len, unsigned int divide_parts,
                               void* arg_buffer, size_t arg_buffer_size, void* local_buffer, int device) {
  const unsigned nv = num_scalars + num_polys, nm = num_mut_scalars + num_mut_polys;
  if (kernel == nullptr || arg_buffer == nullptr || divide_parts <= 3 || len == 0 || len % divide_parts != 0)
    return hipErrorInvalidValue;
  if (arg_buffer_size < (size_t)(2 * nv + 3 * nm) * sizeof(PolyPtr)) return hipErrorInvalidValue;
  for (unsigned i = 0; i < num_scalars; i++)
    if (!host_scalars[i] || !dev_scalars[i]) return hipErrorInvalidValue;
  for (unsigned i = 0; i < num_mut_scalars; i++)
    if (!host_mut_scalars[i] || !dev_mut_scalars[i]) return hipErrorInvalidValue;
  for (unsigned i = 0; i < num_polys; i++) {
    const ConstPolyPtr& p = host_polys[i];
    if (!plain(p.ptr, p.len, p.rotate, p.offset, p.whole_len, len)) return hipErrorInvalidValue;
    for (unsigned j = 0; j < 2; j++)
      if (!dev_polys[j * num_polys + i]) return hipErrorInvalidValue;
  }
  for (unsigned i = 0; i < num_mut_polys; i++) {
    const PolyPtr& p = host_mut_polys[i];
    if (!plain(p.ptr, p.len, p.rotate, p.offset, p.whole_len, len)) return hipErrorInvalidValue;
    for (unsigned j = 0; j < 3; j++)
      if (!dev_mut_polys[j * num_mut_polys + i]) return hipErrorInvalidValue;
  }
  const size_t chunk = len / divide_parts;

  Err err;
  err(hipSetDevice(device));
  if (err.first != hipSuccess) return err.first;

  // Every polynomial in mapped pinned memory: the kernel reads and writes the
  // host arrays itself over PCIe in one launch, with no staging and no
  // cross-stream waits.  On MI355X this kept the link rate in every process
  // state measured, while the staged form lost its overlap after some earlier
  // work in the same process (the device chunk buffers then go unused).
  {
    bool direct_all = true;
    std::vector<const uint32_t*> in_dev(num_polys, nullptr);
    std::vector<uint32_t*> out_dev(num_mut_polys, nullptr);
    for (unsigned i = 0; i < num_polys && direct_all; i++)
      direct_all = (in_dev[i] = mapped_device_ptr(const_cast<uint32_t*>(host_polys[i].ptr), len * ELEM)) != nullptr;
    for (unsigned i = 0; i < num_mut_polys && direct_all; i++)
      direct_all = (out_dev[i] = mapped_device_ptr(host_mut_polys[i].ptr, len * ELEM)) != nullptr;
    if (direct_all)
      return run_direct(kernel, num_scalars, host_scalars, dev_scalars, num_mut_scalars, host_mut_scalars,
                        dev_mut_scalars, num_polys, host_polys, in_dev.data(), num_mut_polys, host_mut_polys,
                        out_dev.data(), len, arg_buffer, local_buffer);
  }
  hipStream_t h2d = nullptr, comp = nullptr, d2h = nullptr;
  err(hipStreamCreate(&h2d));
  err(hipStreamCreate(&comp));
  err(hipStreamCreate(&d2h));
  Ev mut_h2d[3], mut_comp[3], mut_d2h[3], h2d_done[2], comp_done[2];
  for (Ev* set : {mut_h2d, mut_comp, mut_d2h})
    for (int j = 0; j < 3; j++) err(hipEventCreateWithFlags(&set[j].e, hipEventDisableTiming));
  for (Ev* set : {h2d_done, comp_done})
    for (int j = 0; j < 2; j++) err(hipEventCreateWithFlags(&set[j].e, hipEventDisableTiming));

  // argument arrays: [scalars..., polys...] per buffer set, as the reference
  // lays them out (fused_kernels.rs:358-408); the host copy stays alive until
  // the final synchronisation
  std::vector<ConstPolyPtr> vars(2 * (size_t)nv);
  for (unsigned j = 0; j < 2; j++) {
    for (unsigned i = 0; i < num_scalars; i++) vars[j * nv + i] = ConstPolyPtr{dev_scalars[i], 1, 0, 0, 1};
    for (unsigned i = 0; i < num_polys; i++)
      vars[j * nv + num_scalars + i] = ConstPolyPtr{dev_polys[j * num_polys + i], chunk, 0, 0, chunk};
  }
  // A mutable polynomial the kernel only writes, held in mapped pinned host
  // memory, is written by the kernel in place over PCIe: no device buffer and
  // no copy back.  (Device-to-host copies on freshly created streams have run
  // at half the link rate here, depending on what the process did before,
  // while kernel writes to mapped memory kept the full rate and overlapped the
  // copies in.)  The mutable argument set of each chunk is therefore its own,
  // copied into the buffer set the chunk uses before its kernel.
  std::vector<uint32_t*> direct(num_mut_polys, nullptr);
  for (unsigned i = 0; i < num_mut_polys; i++)
    if (mut_poly_read && !mut_poly_read[i]) direct[i] = mapped_device_ptr(host_mut_polys[i].ptr, len * ELEM);
  std::vector<PolyPtr> muts((size_t)divide_parts * nm);
  for (unsigned c = 0; c < divide_parts; c++) {
    PolyPtr* m = muts.data() + (size_t)c * nm;
    for (unsigned i = 0; i < num_mut_scalars; i++) m[i] = PolyPtr{dev_mut_scalars[i], 1, 0, 0, 1};
    for (unsigned i = 0; i < num_mut_polys; i++) {
      const PolyPtr& p = host_mut_polys[i];
      const size_t phys = ((size_t)c * chunk + len - p.rotate) % len;  // logical start of the chunk
      m[num_mut_scalars + i] = direct[i] ? PolyPtr{direct[i], chunk, 0, phys, len}
                                         : PolyPtr{dev_mut_polys[(c % 3) * num_mut_polys + i], chunk, 0, 0, chunk};
    }
  }
  ConstPolyPtr* d_vars = reinterpret_cast<ConstPolyPtr*>(arg_buffer);
  PolyPtr* d_muts = reinterpret_cast<PolyPtr*>(d_vars + 2 * nv);
  if (err.first == hipSuccess) {
    if (nv) err(hipMemcpyAsync(d_vars, vars.data(), vars.size() * sizeof(ConstPolyPtr), hipMemcpyHostToDevice, h2d));
    for (unsigned i = 0; i < num_mut_scalars; i++)
      err(hipMemcpyAsync(dev_mut_scalars[i], host_mut_scalars[i], ELEM, hipMemcpyHostToDevice, h2d));
    for (unsigned i = 0; i < num_scalars; i++)
      err(hipMemcpyAsync(dev_scalars[i], host_scalars[i], ELEM, hipMemcpyHostToDevice, h2d));
  }

  auto wait = [&](hipStream_t s, Ev& ev) {
    if (ev.recorded) err(hipStreamWaitEvent(s, ev.e, 0));
  };
  auto record = [&](Ev& ev, hipStream_t s) {
    err(hipEventRecord(ev.e, s));
    ev.recorded = true;
  };

  for (unsigned c = 0, mb = 0, b = 0; c < divide_parts && err.first == hipSuccess; c++) {
    const size_t start = (size_t)c * chunk;
    // mutable chunk in, once its buffer's previous copy-out is done
    wait(h2d, mut_d2h[mb]);
    if (nm)
      err(hipMemcpyAsync(d_muts + (size_t)mb * nm, muts.data() + (size_t)c * nm, nm * sizeof(PolyPtr),
                         hipMemcpyHostToDevice, h2d));
    for (unsigned i = 0; i < num_mut_polys; i++) {
      if (mut_poly_read && !mut_poly_read[i]) continue;  // written in full by the kernel
      const PolyPtr& p = host_mut_polys[i];
      copy_in(err, p.ptr, len, p.rotate, start, chunk, dev_mut_polys[mb * num_mut_polys + i], h2d);
    }
    record(mut_h2d[mb], h2d);
    // read-only chunk in, once the kernel that last read its buffer is done
    wait(h2d, comp_done[b]);
    for (unsigned i = 0; i < num_polys; i++) {
      const ConstPolyPtr& p = host_polys[i];
      copy_in(err, p.ptr, len, p.rotate, start, chunk, dev_polys[b * num_polys + i], h2d);
    }
    record(h2d_done[b], h2d);
    // the kernel on this chunk
    wait(comp, mut_h2d[mb]);
    wait(comp, h2d_done[b]);
    if (err.first == hipSuccess)
      err(kernel(d_vars + (size_t)b * nv, d_muts + (size_t)mb * nm, chunk, c == 0, local_buffer, comp));
    record(comp_done[b], comp);
    record(mut_comp[mb], comp);
    // mutable chunk back
    wait(d2h, mut_comp[mb]);
    for (unsigned i = 0; i < num_mut_polys; i++) {
      if (direct[i]) continue;  // the kernel wrote host memory itself
      const PolyPtr& p = host_mut_polys[i];
      copy_out(err, p.ptr, len, p.rotate, start, chunk, dev_mut_polys[mb * num_mut_polys + i], d2h);
    }
    record(mut_d2h[mb], d2h);
    mb = (mb + 1) % 3;
    b = (b + 1) % 2;
  }
  for (unsigned i = 0; i < num_mut_scalars; i++)
    err(hipMemcpyAsync(host_mut_scalars[i], dev_mut_scalars[i], ELEM, hipMemcpyDeviceToHost, d2h));

  // drain and release, whatever happened above
  for (hipStream_t s : {h2d, comp, d2h})
    if (s) err(hipStreamSynchronize(s));
  for (Ev* set : {mut_h2d, mut_comp, mut_d2h})
    for (int j = 0; j < 3; j++)
      if (set[j].e) err(hipEventDestroy(set[j].e));
  for (Ev* set : {h2d_done, comp_done})
    for (int j = 0; j < 2; j++)
      if (set[j].e) err(hipEventDestroy(set[j].e));
  for (hipStream_t s : {h2d, comp, d2h})
    if (s) err(hipStreamDestroy(s));
  return err.first;
}

}  // extern "C"
