/*
 * gpu_hip.cpp - the GPU_Backend function table (gpu_backend.h:17-71) over
 * HIP for one MI355X (gfx950).  Replaces src/gpu/vulkan/gpu_vulkan.c for
 * ComputeQL's operator path.
 *
 * Design (MI355X-first, not a translation of the Vulkan backend):
 *  - buffers are plain hipMalloc allocations in HBM, rounded up to 256 B so
 *    every kernel may issue 16-byte loads over a column's padded tail;
 *  - a kernel object holds its bound args (the reference's descriptor set
 *    per kernel, gpu_vulkan.h:34-48) but every dispatch snapshots them into
 *    the launch's kernarg block, so rebinding after a dispatch never
 *    retargets it (the hazard prefix_sum2.comp:3-5 works around);
 *  - kernel objects are cached per name and per host thread, so concurrent
 *    operators on different threads never share bound args;
 *  - a batch (gpu_vulkan.c:3062-3116) is a HIP stream taken from a pool;
 *    batches on different threads run concurrently;
 *  - stream slots (gpu_vulkan.h:231-345) are one HIP stream each.
 */
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cqlhip.h"
#include "qk_library.h"

struct GPU_Buffer {
  void* ptr;
  U64 size;       /* bytes asked for */
  U64 alloc_size; /* bytes allocated (padded) */
  U32 flags;
  bool pooled;
  bool host_import;
  bool wrapped; /* pinned host memory seen by the device; not ours to free */
  std::string pool_name;
};

struct GPU_Kernel {
  std::string name;
  const QK_LibraryKernel* lib;
  hipModule_t module;
  hipFunction_t fn;
  QK_Args args;
  U32 lds_bytes;
  U32 static_lds; /* the kernel's own __shared__ bytes, beside lds_bytes */
  U32 missing_mask; /* bindings set to a null buffer (gpu_dispatch.c:260-306) */
  bool generated;
};

struct PendingProfile {
  std::string name;
  hipEvent_t start, stop;
};

/* small transfers of a batch go through pinned host staging so they are
 * asynchronous on the batch stream; readbacks are copied out at batch_end */
struct PendingRead {
  void* user;
  const void* pinned;
  U64 size;
};

/* a validated launch: kernel, args snapshot, grid */
struct PreparedLaunch {
  GPU_Kernel* k;
  QK_Args args;
  U32 blocks;
  U32 block;
  U32 lds;
};

/* one operation of a deferred (graph) batch; writes and reads go through
 * pinned staging so a captured graph's copies have fixed host addresses */
struct BatchOp {
  int kind; /* 0 write, 1 read, 2 kernel */
  void* dev;
  U64 size;
  char* staged;
  void* user;
  PreparedLaunch launch;
};

struct GPU_Batch {
  hipStream_t stream;
  bool failed;
  std::vector<PendingProfile> prof;
  std::vector<char*> blocks; /* pinned staging blocks owned by this batch */
  U64 block_used;
  std::vector<PendingRead> reads;
  /* deferred batches (gpu_hip_batch_begin_graph): operations are recorded
   * and run at batch_end as a cached HIP graph */
  bool deferred = false;
  std::string key;
  std::vector<BatchOp> ops;
  std::vector<char*> big_blocks; /* staging larger than one block */
};

struct GPU_StreamSlot {
  hipStream_t stream;
  hipEvent_t done;
  bool submitted;
  std::vector<PendingProfile> prof;
};

struct GPU_Stream {
  std::vector<GPU_StreamSlot> slots;
  GPU_StreamStats stats;
};

namespace {

struct HipState {
  std::mutex lock;
  bool initialised = false;
  int device = 0;
  hipDeviceProp_t props;
  hipStream_t default_stream = nullptr;
  std::vector<hipStream_t> free_streams;
  GPU_MemoryStats mem = {};
  std::atomic<bool> lost{false};
  std::unordered_map<std::string, GPU_Buffer*> pool;
  /* generated kernels: name -> (module, function), LRU by use tick */
  struct Generated {
    hipModule_t module;
    hipFunction_t fn;
    U64 tick;
  };
  std::unordered_map<std::string, Generated> generated;
  U64 generated_tick = 0;
  U64 generated_compiles = 0;
  U64 generated_evictions = 0;
  U64 kernel_time_us = 0;
  bool profile = false;
  std::map<std::string, std::pair<U64, U64>> profile_table; /* ns, dispatches */
  U64 fault_alloc_at = 0;
  U64 fault_alloc_counter = 0;
  U64 fault_read_at = 0;
  U64 fault_read_counter = 0;
  U64 (*evict_hook)(U64) = nullptr;
};

HipState& st()
{
  static HipState* s = new HipState();
  return *s;
}

thread_local char t_last_error[512];
thread_local char t_profile_tag[64];

void set_error(const char* fmt, const char* a = "", long long b = 0)
{
  snprintf(t_last_error, sizeof(t_last_error), fmt, a, b);
}

bool check(hipError_t e, const char* what)
{
  if (e == hipSuccess) return true;
  snprintf(t_last_error, sizeof(t_last_error), "%s: %s", what, hipGetErrorString(e));
  if (e == hipErrorIllegalAddress || e == hipErrorLaunchFailure || e == hipErrorECCNotCorrectable ||
      e == hipErrorAssert)
    st().lost = true; /* sticky, like gpu_vulkan.c:1523-1538 */
  return false;
}

U64 max_generated()
{
  const char* env = getenv("GPU_HIP_MAX_GENERATED_KERNELS");
  U64 v = env ? strtoull(env, nullptr, 10) : 0;
  return v ? v : 64; /* GPU_VULKAN_MAX_GENERATED_KERNELS default, gpu_vulkan.c:2696-2761 */
}

/* per-thread kernel objects keyed by name */
thread_local std::unordered_map<std::string, GPU_Kernel*>* t_kernels;

std::unordered_map<std::string, GPU_Kernel*>& kernels()
{
  if (!t_kernels) t_kernels = new std::unordered_map<std::string, GPU_Kernel*>();
  return *t_kernels;
}

hipStream_t take_stream()
{
  HipState& s = st();
  std::lock_guard<std::mutex> g(s.lock);
  if (!s.free_streams.empty()) {
    hipStream_t h = s.free_streams.back();
    s.free_streams.pop_back();
    return h;
  }
  hipStream_t h = nullptr;
  if (!check(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreate")) return nullptr;
  return h;
}

void give_stream(hipStream_t h)
{
  if (!h) return;
  HipState& s = st();
  std::lock_guard<std::mutex> g(s.lock);
  s.free_streams.push_back(h);
}

/* synchronous copies and fills run on a pooled non-blocking stream, never
 * on the legacy null stream: batch streams are not ordered against the null
 * stream, and other host threads may be capturing graphs meanwhile */
hipError_t copy_sync(void* dst, const void* src, U64 size, hipMemcpyKind kind)
{
  hipStream_t h = take_stream();
  if (!h) return hipErrorOutOfMemory;
  hipError_t e = hipMemcpyAsync(dst, src, size, kind, h);
  hipError_t e2 = hipStreamSynchronize(h);
  give_stream(h);
  return e != hipSuccess ? e : e2;
}

hipError_t zero_sync(void* p, U64 size)
{
  hipStream_t h = take_stream();
  if (!h) return hipErrorOutOfMemory;
  hipError_t e = hipMemsetAsync(p, 0, size, h);
  hipError_t e2 = hipStreamSynchronize(h);
  give_stream(h);
  return e != hipSuccess ? e : e2;
}

void flush_profile(std::vector<PendingProfile>& prof)
{
  HipState& s = st();
  for (auto& p : prof) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      U64 ns = (U64)((double)ms * 1.0e6);
      std::lock_guard<std::mutex> g(s.lock);
      auto& e = s.profile_table[p.name];
      e.first += ns;
      e.second += 1;
      s.kernel_time_us += ns / 1000;
    }
    hipEventDestroy(p.start);
    hipEventDestroy(p.stop);
  }
  prof.clear();
}

/* validate a dispatch and snapshot its args; *empty when the grid is 0 */
bool prepare(GPU_Kernel* k, U32 global_work_size, U32 local_work_size, PreparedLaunch& out, bool& empty)
{
  empty = false;
  if (!k) {
    set_error("dispatch of a null kernel");
    return false;
  }
  if (k->missing_mask) {
    /* gpu_dispatch.c:260-306/:415-426: a dispatch whose args failed to
     * allocate is skipped and counted */
    set_error("dispatch of %s skipped: an argument buffer is missing", k->name.c_str());
    return false;
  }
  U32 block = k->lib ? k->lib->block : (local_work_size ? local_work_size : 256);
  if (k->lib && local_work_size && local_work_size != block) {
    set_error("%s: workgroup size mismatch", k->name.c_str());
    return false;
  }
  U64 blocks = ((U64)global_work_size + block - 1) / block;
  if (blocks == 0) {
    empty = true;
    return true;
  }
  if (blocks > 0x7fffffffULL) {
    set_error("%s: grid too large", k->name.c_str());
    return false;
  }
  {
    /* a launch whose dynamic and static LDS exceed a workgroup's would
     * abort the queue: refuse it here with the sizes */
    const U64 limit = gpu_hip_max_lds_bytes();
    if ((U64)k->lds_bytes + k->static_lds > limit) {
      snprintf(t_last_error, sizeof(t_last_error), "%s: %u bytes of dynamic LDS and %u static exceed the %llu a workgroup has",
               k->name.c_str(), k->lds_bytes, k->static_lds, (unsigned long long)limit);
      return false;
    }
  }
  out.k = k;
  out.args = k->args; /* snapshot */
  out.args.global_work_size = global_work_size;
  out.args.grid_blocks = (U32)blocks;
  out.blocks = (U32)blocks;
  out.block = block;
  out.lds = k->lds_bytes;
  return true;
}

/* ---- the no-allocation scratch of qe_aggregate_noalloc ------------------
 * A file-scope device array of QE_NOALLOC_SCRATCH_BYTES: each GPU the
 * library runs on holds its own copy until the process ends.  Host code
 * never uses the array's name as an address (that is a host-side stand-in):
 * gpu_hip_init takes each device's address once with hipGetSymbolAddress
 * and keeps it, with a stream of default priority, per device. */
static __device__ __align__(256) unsigned char scratch[QE_NOALLOC_SCRATCH_BYTES];

#define QK_NOALLOC_MAX_DEVICES 64
struct NoallocDevice {
  void* ptr;
  hipStream_t stream;
};
static NoallocDevice g_noalloc[QK_NOALLOC_MAX_DEVICES];

/* once per device, with that device current */
static bool noalloc_setup(int device)
{
  if (device < 0 || device >= QK_NOALLOC_MAX_DEVICES) return false;
  if (g_noalloc[device].ptr) return true;
  void* p = nullptr;
  if (!check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || !p) return false;
  hipStream_t h = nullptr;
  if (!check(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreate")) return false;
  g_noalloc[device].stream = h;
  g_noalloc[device].ptr = p;
  return true;
}

extern "C" U64 gpu_hip_noalloc_scratch(void)
{
  HipState& s = st();
  if (!s.initialised || s.device < 0 || s.device >= QK_NOALLOC_MAX_DEVICES) return 0;
  return (U64)(uintptr_t)g_noalloc[s.device].ptr;
}

/* upload `upload_bytes` to the scratch at upload_offset, launch the library
 * kernels in order with their args (QK_Args by value, `blocks[i]` blocks of
 * the kernel's fixed size) and read `download_bytes` at download_offset
 * back, all on the device's no-allocation stream; returns when done.
 * Nothing is allocated: no buffers, no staging, no events. */
extern "C" B32 gpu_hip_noalloc_run(const void* upload, U64 upload_bytes, U64 upload_offset, const char* const* kernels,
                                   const QK_Args* args, const U32* blocks, U32 steps, void* download,
                                   U64 download_bytes, U64 download_offset)
{
  HipState& s = st();
  if (!s.initialised || s.device < 0 || s.device >= QK_NOALLOC_MAX_DEVICES || !g_noalloc[s.device].ptr) {
    set_error("no-allocation scratch is not set up on this device");
    return 0;
  }
  const NoallocDevice& d = g_noalloc[s.device];
  if (upload_offset + upload_bytes > QE_NOALLOC_SCRATCH_BYTES ||
      download_offset + download_bytes > QE_NOALLOC_SCRATCH_BYTES) {
    set_error("no-allocation scratch: %lld bytes past its end", "",
              (long long)(upload_offset + upload_bytes > download_offset + download_bytes
                              ? upload_offset + upload_bytes : download_offset + download_bytes));
    return 0;
  }
  char* base = (char*)d.ptr;
  bool ok = true;
  if (upload_bytes)
    ok = check(hipMemcpyAsync(base + upload_offset, upload, upload_bytes, hipMemcpyHostToDevice, d.stream),
               "noalloc upload");
  for (U32 i = 0; ok && i < steps; i++) {
    const QK_LibraryKernel* lib = qk_library_find(kernels[i]);
    if (!lib) {
      set_error("no kernel named %s", kernels[i]);
      ok = false;
      break;
    }
    QK_Args a = args[i];
    void* params[1] = {&a};
    ok = check(hipLaunchKernel(lib->host_fn, dim3(blocks[i]), dim3(lib->block), params, 0, d.stream), kernels[i]);
  }
  if (ok && download_bytes)
    ok = check(hipMemcpyAsync(download, base + download_offset, download_bytes, hipMemcpyDeviceToHost, d.stream),
               "noalloc download");
  /* drain the stream even after a failed step, so nothing of this call is
   * still in flight when the next one reuses the scratch */
  const bool synced = check(hipStreamSynchronize(d.stream), "noalloc sync");
  return ok && synced ? 1 : 0;
}

hipError_t enqueue(const PreparedLaunch& p, hipStream_t stream)
{
  QK_Args args = p.args;
  void* params[1] = {&args};
  if (p.k->lib) return hipLaunchKernel(p.k->lib->host_fn, dim3(p.blocks), dim3(p.block), params, p.lds, stream);
  return hipModuleLaunchKernel(p.k->fn, p.blocks, 1, 1, p.block, 1, 1, p.lds, stream, params, nullptr);
}

bool launch(GPU_Kernel* k, U32 global_work_size, U32 local_work_size, hipStream_t stream,
            std::vector<PendingProfile>* prof)
{
  PreparedLaunch p;
  bool empty = false;
  if (!prepare(k, global_work_size, local_work_size, p, empty)) return false;
  if (empty) return true;
  PendingProfile pp;
  bool timed = st().profile && prof;
  if (timed) {
    pp.name = t_profile_tag[0] ? std::string(t_profile_tag) + ":" + k->name : k->name;
    hipEventCreate(&pp.start);
    hipEventCreate(&pp.stop);
    hipEventRecord(pp.start, stream);
  }
  hipError_t e = enqueue(p, stream);
  if (timed) {
    hipEventRecord(pp.stop, stream);
    prof->push_back(pp);
  }
  return check(e, k->name.c_str());
}

GPU_Buffer* alloc_impl(U64 size, U32 flags, void* data)
{
  HipState& s = st();
  if (!s.initialised) gpu_hip_init();
  {
    std::lock_guard<std::mutex> g(s.lock);
    s.mem.alloc_calls++;
    if (s.fault_alloc_at) {
      s.fault_alloc_counter++;
      if (s.fault_alloc_counter == s.fault_alloc_at) {
        s.mem.alloc_failures++;
        s.mem.unrecovered_alloc_failures++;
        s.fault_alloc_at = 0;
        set_error("injected allocation failure");
        return nullptr;
      }
    }
  }
  U64 padded = ((size ? size : 1) + 255) & ~(U64)255;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, padded);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    /* evict pooled buffers and retry once (gpu_vulkan.c:1578) */
    std::vector<GPU_Buffer*> victims;
    {
      std::lock_guard<std::mutex> g(s.lock);
      s.mem.alloc_failures++;
      for (auto& kv : s.pool) victims.push_back(kv.second);
      s.pool.clear();
    }
    for (GPU_Buffer* v : victims) {
      std::lock_guard<std::mutex> g(s.lock);
      s.mem.pool_evictions++;
      s.mem.pool_evicted_bytes += v->alloc_size;
      s.mem.bytes_in_use -= v->alloc_size;
      hipFree(v->ptr);
      delete v;
    }
    e = hipMalloc(&p, padded);
    if (e != hipSuccess) {
      /* then the owner of cached device data (the resident chunk cache)
       * frees what it can, and one more try (g_gpu_evict_hook,
       * query_exec.c:1817-1822) */
      (void)hipGetLastError();
      U64 (*hook)(U64) = s.evict_hook;
      if (hook && hook(padded) > 0) e = hipMalloc(&p, padded);
    }
    std::lock_guard<std::mutex> g(s.lock);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      s.mem.unrecovered_alloc_failures++;
      set_error("hipMalloc of %s bytes failed", "", (long long)padded);
      return nullptr;
    }
    s.mem.alloc_retry_successes++;
  }
  GPU_Buffer* b = new GPU_Buffer();
  b->ptr = p;
  b->size = size;
  b->alloc_size = padded;
  b->flags = flags;
  b->pooled = false;
  b->host_import = false;
  {
    std::lock_guard<std::mutex> g(s.lock);
    s.mem.bytes_in_use += padded;
    if (s.mem.bytes_in_use > s.mem.bytes_peak) s.mem.bytes_peak = s.mem.bytes_in_use;
  }
  if (data && size) {
    if (!check(copy_sync(p, data, size, hipMemcpyHostToDevice), "copy H2D")) {
      gpu_hip_buffer_release(b);
      return nullptr;
    }
  }
  /* new buffers start zeroed (kernels keep their completion tickets in
   * scratch buffers and reset them themselves).  The zeroing completes
   * before the buffer is handed out: the batch streams are not ordered
   * against the stream it runs on, and a kernel's writes into the fresh
   * buffer must not be overwritten by it. */
  hipError_t ze = hipSuccess;
  if (!data || !size) ze = zero_sync(p, padded);
  else if (padded > size) ze = zero_sync((char*)p + size, padded - size);
  if (!check(ze, "alloc zero")) {
    gpu_hip_buffer_release(b);
    return nullptr;
  }
  return b;
}

}  // namespace

extern "C" {

void gpu_hip_init(void)
{
  HipState& s = st();
  std::lock_guard<std::mutex> g(s.lock);
  if (s.initialised) return;
  int count = 0;
  if (!check(hipGetDeviceCount(&count), "hipGetDeviceCount") || count < 1) {
    if (count < 1) set_error("no HIP device");
    return;
  }
  const char* dev = getenv("CQLHIP_DEVICE");
  s.device = dev ? atoi(dev) : 0;
  if (!check(hipSetDevice(s.device), "hipSetDevice")) return;
  if (!check(hipGetDeviceProperties(&s.props, s.device), "hipGetDeviceProperties")) return;
  if (!check(hipStreamCreateWithFlags(&s.default_stream, hipStreamNonBlocking), "hipStreamCreate")) return;
  /* the no-allocation entry point's scratch address on this device; on
   * failure that entry point reports it and every other one works */
  noalloc_setup(s.device);
  s.initialised = true;
}

void gpu_hip_release(void)
{
  HipState& s = st();
  if (!s.initialised) return;
  hipDeviceSynchronize();
  std::lock_guard<std::mutex> g(s.lock);
  for (auto& kv : s.pool) {
    hipFree(kv.second->ptr);
    delete kv.second;
  }
  s.pool.clear();
  for (auto& kv : s.generated) hipModuleUnload(kv.second.module);
  s.generated.clear();
  for (hipStream_t h : s.free_streams) hipStreamDestroy(h);
  s.free_streams.clear();
  /* kernel objects of other threads keep dangling module pointers; the
   * reference re-inits the whole backend on switch (gpu_dispatch.c:130-151) */
  if (t_kernels) {
    for (auto& kv : *t_kernels) delete kv.second;
    t_kernels->clear();
  }
}

void gpu_hip_wait(void)
{
  if (!st().initialised) return;
  check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

U64 gpu_hip_get_executed_kernel_time_microseconds(void) { return st().kernel_time_us; }

U64 gpu_hip_device_total_memory(void)
{
  size_t fr = 0, tot = 0;
  if (!st().initialised) gpu_hip_init();
  hipMemGetInfo(&fr, &tot);
  return tot;
}

U64 gpu_hip_device_free_memory(void)
{
  size_t fr = 0, tot = 0;
  if (!st().initialised) gpu_hip_init();
  hipMemGetInfo(&fr, &tot);
  return fr;
}

/* no per-binding range limit on HIP; report total memory */
U64 gpu_hip_device_max_storage_buffer_range(void) { return gpu_hip_device_total_memory(); }

B32 gpu_hip_device_lost(void) { return st().lost ? 1 : 0; }

void gpu_hip_pool_begin_epoch(void) {}

GPU_MemoryStats gpu_hip_memory_stats(void)
{
  std::lock_guard<std::mutex> g(st().lock);
  return st().mem;
}

GPU_Buffer* gpu_hip_buffer_alloc(U64 size, GPU_BufferFlags flags, void* data)
{
  return alloc_impl(size, (U32)flags, data);
}

GPU_Buffer* gpu_hip_buffer_alloc_pooled(String8 name, U64 size, GPU_BufferFlags flags, void* data)
{
  HipState& s = st();
  std::string key((const char*)name.str, (size_t)name.size);
  GPU_Buffer* hit = nullptr;
  {
    std::lock_guard<std::mutex> g(s.lock);
    auto it = s.pool.find(key);
    if (it != s.pool.end()) {
      if (it->second->alloc_size >= size) {
        hit = it->second;
        s.pool.erase(it);
      }
    }
  }
  if (hit) {
    hit->size = size;
    if (data && size && !check(copy_sync(hit->ptr, data, size, hipMemcpyHostToDevice), "pooled copy")) {
      /* back to the pool, and no buffer for the caller */
      std::lock_guard<std::mutex> g(s.lock);
      s.pool.emplace(key, hit);
      return nullptr;
    }
    hit->pooled = true;
    hit->pool_name = key;
    return hit;
  }
  GPU_Buffer* b = alloc_impl(size, (U32)flags, data);
  if (b) {
    b->pooled = true;
    b->pool_name = key;
  }
  return b;
}

/* Host-pointer import (gpu_vulkan.c:2213-2358).  Registering existing host
 * memory is not available on this platform, so the import is a device copy
 * made once; callers treat it as read-only like the reference does. */
GPU_Buffer* gpu_hip_buffer_import_host_readonly(void* host_ptr, U64 size)
{
  GPU_Buffer* b = alloc_impl(size, GPU_BufferFlag_Read, host_ptr);
  if (b) b->host_import = true;
  return b;
}

GPU_Buffer* gpu_hip_buffer_import_host_readonly_pooled(String8 name, void* host_ptr, U64 size)
{
  return gpu_hip_buffer_alloc_pooled(name, size, GPU_BufferFlag_Read, host_ptr);
}

void gpu_hip_buffer_release(GPU_Buffer* buffer)
{
  if (!buffer) return;
  HipState& s = st();
  if (buffer->pooled) {
    /* back to the pool under its name; a previous holder is freed */
    GPU_Buffer* old = nullptr;
    {
      std::lock_guard<std::mutex> g(s.lock);
      auto it = s.pool.find(buffer->pool_name);
      if (it != s.pool.end()) old = it->second;
      s.pool[buffer->pool_name] = buffer;
    }
    if (old && old != buffer) {
      hipFree(old->ptr);
      std::lock_guard<std::mutex> g(s.lock);
      s.mem.bytes_in_use -= old->alloc_size;
      delete old;
    }
    return;
  }
  if (buffer->wrapped) {
    delete buffer;
    return;
  }
  hipFree(buffer->ptr);
  std::lock_guard<std::mutex> g(s.lock);
  s.mem.bytes_in_use -= buffer->alloc_size;
  delete buffer;
}

/* fault injection (gpu_hip_fault_read_at): true when this device-to-host
 * read is the one set to fail */
static bool fault_read_now()
{
  HipState& s = st();
  std::lock_guard<std::mutex> g(s.lock);
  if (!s.fault_read_at) return false;
  if (++s.fault_read_counter != s.fault_read_at) return false;
  s.fault_read_at = 0;
  return true;
}

B32 gpu_hip_buffer_write_checked(GPU_Buffer* buffer, const void* data, U64 size)
{
  if (!buffer || (!data && size)) {
    set_error("buffer_write: null buffer or data");
    return 0;
  }
  if (size > buffer->alloc_size) {
    set_error("buffer_write: %lld bytes exceed the buffer", "", (long long)size);
    return 0;
  }
  if (!size) return 1;
  /* complete before returning: batch streams are not ordered against the
   * stream this copy runs on */
  return check(copy_sync(buffer->ptr, data, size, hipMemcpyHostToDevice), "buffer_write") ? 1 : 0;
}

B32 gpu_hip_buffer_read_checked(GPU_Buffer* buffer, void* data, U64 size)
{
  if (!buffer || (!data && size)) {
    set_error("buffer_read: null buffer or data");
    return 0;
  }
  if (size > buffer->alloc_size) {
    set_error("buffer_read: %lld bytes exceed the buffer", "", (long long)size);
    return 0;
  }
  if (!size) return 1;
  if (fault_read_now()) {
    set_error("buffer_read: injected read failure");
    return 0;
  }
  return check(copy_sync(data, buffer->ptr, size, hipMemcpyDeviceToHost), "buffer_read") ? 1 : 0;
}

void gpu_hip_buffer_write(GPU_Buffer* buffer, void* data, U64 size)
{
  if (!buffer || !data) return;
  if (size > buffer->alloc_size) size = buffer->alloc_size;
  gpu_hip_buffer_write_checked(buffer, data, size);
}

/* the reference's contract: void, and zeros on failure (gpu_vulkan.c:
 * 3107-3111); the operators use gpu_hip_buffer_read_checked */
void gpu_hip_buffer_read(GPU_Buffer* buffer, void* data, U64 size)
{
  if (!buffer || !data) return;
  if (size > buffer->alloc_size) size = buffer->alloc_size;
  if (!gpu_hip_buffer_read_checked(buffer, data, size)) memset(data, 0, size);
}

GPU_Kernel* gpu_hip_kernel_alloc(String8 name)
{
  HipState& s = st();
  if (!s.initialised) gpu_hip_init();
  std::string key((const char*)name.str, (size_t)name.size);
  auto& ks = kernels();
  auto it = ks.find(key);
  if (it != ks.end()) return it->second;
  const QK_LibraryKernel* lib = qk_library_find(key.c_str());
  if (!lib) {
    /* a generated kernel already loaded under this name */
    std::lock_guard<std::mutex> g(s.lock);
    auto gi = s.generated.find(key);
    if (gi == s.generated.end()) {
      set_error("no kernel named %s", key.c_str());
      return nullptr;
    }
    GPU_Kernel* k = new GPU_Kernel();
    memset(&k->args, 0, sizeof(k->args));
    k->name = key;
    k->lib = nullptr;
    k->module = gi->second.module;
    k->fn = gi->second.fn;
    k->generated = true;
    k->missing_mask = 0;
    k->lds_bytes = 0;
    {
      int v = 0;
      k->static_lds = hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, k->fn) == hipSuccess ? (U32)v : 0u;
    }
    gi->second.tick = ++s.generated_tick;
    ks[key] = k;
    return k;
  }
  GPU_Kernel* k = new GPU_Kernel();
  memset(&k->args, 0, sizeof(k->args));
  k->name = key;
  k->lib = lib;
  k->module = nullptr;
  k->fn = nullptr;
  k->generated = false;
  k->missing_mask = 0;
  k->lds_bytes = 0;
  {
    hipFuncAttributes fa;
    memset(&fa, 0, sizeof(fa));
    k->static_lds = hipFuncGetAttributes(&fa, lib->host_fn) == hipSuccess ? (U32)fa.sharedSizeBytes : 0u;
  }
  ks[key] = k;
  return k;
}

/* A kernel from a code object made at run time by hiprtc (the reference
 * passes SPIR-V from shaderc here; the ABI name is kept).  An empty blob only
 * looks the name up.  Loaded modules live in an LRU of
 * GPU_HIP_MAX_GENERATED_KERNELS (default 64) entries; the kernel function in
 * the code object must be named `qk_main`. */
GPU_Kernel* gpu_hip_kernel_alloc_from_spirv(String8 name, String8 code_object)
{
  HipState& s = st();
  if (!s.initialised) gpu_hip_init();
  std::string key((const char*)name.str, (size_t)name.size);
  {
    std::lock_guard<std::mutex> g(s.lock);
    auto gi = s.generated.find(key);
    if (gi != s.generated.end()) {
      gi->second.tick = ++s.generated_tick;
    }
  }
  auto& ks = kernels();
  auto it = ks.find(key);
  if (it != ks.end() && it->second->generated) {
    std::lock_guard<std::mutex> g(s.lock);
    auto gi = s.generated.find(key);
    if (gi != s.generated.end() && gi->second.module == it->second->module) return it->second;
    /* evicted meanwhile: drop the stale object */
    delete it->second;
    ks.erase(it);
  }
  bool have = false;
  {
    std::lock_guard<std::mutex> g(s.lock);
    have = s.generated.count(key) != 0;
  }
  if (!have) {
    if (code_object.size == 0) return nullptr;
    hipModule_t mod = nullptr;
    if (!check(hipModuleLoadData(&mod, code_object.str), "hipModuleLoadData")) return nullptr;
    hipFunction_t fn = nullptr;
    if (!check(hipModuleGetFunction(&fn, mod, "qk_main"), "hipModuleGetFunction")) {
      hipModuleUnload(mod);
      return nullptr;
    }
    std::lock_guard<std::mutex> g(s.lock);
    if (s.generated.size() >= max_generated()) {
      /* evict the least recently used module; callers on other threads hold
       * kernel objects only for the duration of one operator call */
      auto victim = s.generated.begin();
      for (auto gi = s.generated.begin(); gi != s.generated.end(); ++gi)
        if (gi->second.tick < victim->second.tick) victim = gi;
      hipModuleUnload(victim->second.module);
      auto vk = ks.find(victim->first);
      if (vk != ks.end()) {
        delete vk->second;
        ks.erase(vk);
      }
      s.generated.erase(victim);
      s.generated_evictions++;
    }
    s.generated[key] = HipState::Generated{mod, fn, ++s.generated_tick};
    s.generated_compiles++;
  }
  return gpu_hip_kernel_alloc(name);
}

/* specialization constants have no HIP counterpart: run-time compiled
 * source bakes them in instead */
GPU_Kernel* gpu_hip_kernel_alloc_from_spirv_spec(String8 name, String8 code_object, U32* spec_values, U32 spec_count)
{
  (void)spec_values;
  (void)spec_count;
  return gpu_hip_kernel_alloc_from_spirv(name, code_object);
}

void gpu_hip_generated_kernel_stats(U64* out_compiles, U64* out_evictions)
{
  std::lock_guard<std::mutex> g(st().lock);
  if (out_compiles) *out_compiles = st().generated_compiles;
  if (out_evictions) *out_evictions = st().generated_evictions;
}

/* kernels are cached by name, release is a no-op (gpu_vulkan.c:2771-2774) */
void gpu_hip_kernel_release(GPU_Kernel* kernel) { (void)kernel; }

void gpu_hip_kernel_execute(GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size)
{
  HipState& s = st();
  std::vector<PendingProfile> prof;
  if (launch(kernel, global_work_size, local_work_size, s.default_stream, &prof))
    check(hipStreamSynchronize(s.default_stream), "kernel_execute");
  flush_profile(prof);
}

void gpu_hip_kernel_set_arg_buffer(GPU_Kernel* kernel, U32 index, GPU_Buffer* buffer)
{
  if (!kernel || index >= QK_MAX_BINDINGS) return;
  if (!buffer) {
    kernel->missing_mask |= 1u << index;
    kernel->args.ptr[index] = 0;
    kernel->args.bytes[index] = 0;
    return;
  }
  kernel->missing_mask &= ~(1u << index);
  kernel->args.ptr[index] = (U64)(uintptr_t)buffer->ptr;
  kernel->args.bytes[index] = buffer->size;
}

void gpu_hip_kernel_clear_arg(GPU_Kernel* kernel, U32 index)
{
  if (!kernel || index >= QK_MAX_BINDINGS) return;
  kernel->missing_mask &= ~(1u << index);
  kernel->args.ptr[index] = 0;
  kernel->args.bytes[index] = 0;
}

void gpu_hip_kernel_set_arg_u32(GPU_Kernel* kernel, U32 word, U32 value)
{
  if (!kernel || word >= QK_PUSH_WORDS) return;
  kernel->args.words[word] = value;
}

void gpu_hip_kernel_set_arg_u64(GPU_Kernel* kernel, U32 index, U64 value)
{
  if (!kernel || index >= QK_PUSH_WORDS / 2) return;
  kernel->args.words[2 * index] = (U32)(value & 0xffffffffu);
  kernel->args.words[2 * index + 1] = (U32)(value >> 32);
}

GPU_Batch* gpu_hip_batch_begin(U64 upload_bytes_needed, U64 download_bytes_needed)
{
  (void)upload_bytes_needed;
  (void)download_bytes_needed;
  if (!st().initialised) gpu_hip_init();
  hipStream_t h = take_stream();
  if (!h) return nullptr;
  GPU_Batch* b = new GPU_Batch();
  b->stream = h;
  b->failed = false;
  return b;
}

namespace {
const U64 kStagingBlock = 1 << 20;
const U64 kStagingMax = 256 << 10; /* larger transfers go direct */
std::mutex g_staging_lock;
std::vector<char*> g_staging_free;

char* staging_take(GPU_Batch* b, U64 size)
{
  size = (size + 63) & ~(U64)63;
  if (size > kStagingBlock) {
    /* larger than a staging block: a pinned block of its own */
    char* big = nullptr;
    if (hipHostMalloc((void**)&big, size, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    b->big_blocks.push_back(big);
    return big;
  }
  if (b->blocks.empty() || b->block_used + size > kStagingBlock) {
    char* blk = nullptr;
    {
      std::lock_guard<std::mutex> g(g_staging_lock);
      if (!g_staging_free.empty()) {
        blk = g_staging_free.back();
        g_staging_free.pop_back();
      }
    }
    if (!blk && hipHostMalloc((void**)&blk, kStagingBlock, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    b->blocks.push_back(blk);
    b->block_used = 0;
  }
  char* p = b->blocks.back() + b->block_used;
  b->block_used += size;
  return p;
}

void staging_return_blocks(std::vector<char*>& blocks, std::vector<char*>& big)
{
  {
    std::lock_guard<std::mutex> g(g_staging_lock);
    for (char* blk : blocks) g_staging_free.push_back(blk);
  }
  blocks.clear();
  for (char* p : big) hipHostFree(p);
  big.clear();
}

void staging_return(GPU_Batch* b) { staging_return_blocks(b->blocks, b->big_blocks); }

/* HIP graphs of deferred batches, per host thread (a graph's staging is
 * its own, so concurrent threads never share one), least recently used
 * evicted past kMaxGraphs */
struct GraphEntry {
  std::string key;
  hipGraph_t graph;
  hipGraphExec_t exec;
  std::vector<BatchOp> ops;
  std::vector<char*> blocks, big_blocks;
  U64 tick;
};
const size_t kMaxGraphs = 16;
thread_local std::vector<GraphEntry>* t_graphs;
thread_local U64 t_graph_tick;
std::atomic<U64> g_graph_replays{0}, g_graph_captures{0};

std::vector<GraphEntry>& graphs()
{
  if (!t_graphs) t_graphs = new std::vector<GraphEntry>();
  return *t_graphs;
}

void graph_destroy(GraphEntry& e)
{
  hipGraphExecDestroy(e.exec);
  hipGraphDestroy(e.graph);
  staging_return_blocks(e.blocks, e.big_blocks);
}

/* the same operations up to the bytes written: same buffers, sizes,
 * kernels, args and grids */
bool ops_match(const std::vector<BatchOp>& a, const std::vector<BatchOp>& b)
{
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); i++) {
    const BatchOp &x = a[i], &y = b[i];
    if (x.kind != y.kind) return false;
    if (x.kind != 2) {
      if (x.dev != y.dev || x.size != y.size) return false;
    } else if (x.launch.k != y.launch.k || x.launch.blocks != y.launch.blocks || x.launch.block != y.launch.block ||
               x.launch.lds != y.launch.lds || memcmp(&x.launch.args, &y.launch.args, sizeof(QK_Args)) != 0) {
      return false;
    }
  }
  return true;
}

/* issue the recorded operations on a stream */
bool issue_ops(const std::vector<BatchOp>& ops, hipStream_t stream)
{
  for (const BatchOp& op : ops) {
    hipError_t e;
    if (op.kind == 0) e = hipMemcpyAsync(op.dev, op.staged, op.size, hipMemcpyHostToDevice, stream);
    else if (op.kind == 1) e = hipMemcpyAsync(op.staged, op.dev, op.size, hipMemcpyDeviceToHost, stream);
    else e = enqueue(op.launch, stream);
    if (!check(e, op.kind == 2 ? op.launch.k->name.c_str() : "graph batch copy")) return false;
  }
  return true;
}
}  // namespace

void gpu_hip_batch_buffer_write(GPU_Batch* batch, GPU_Buffer* buffer, void* data, U64 size)
{
  if (!batch || !buffer || !data) return;
  if (size > buffer->alloc_size) size = buffer->alloc_size;
  if (batch->deferred) {
    char* p = staging_take(batch, size);
    if (!p) {
      batch->failed = true;
      return;
    }
    memcpy(p, data, size);
    batch->ops.push_back(BatchOp{0, buffer->ptr, size, p, nullptr, {}});
    return;
  }
  const void* src = data;
  if (size <= kStagingMax) {
    char* p = staging_take(batch, size);
    if (p) {
      memcpy(p, data, size);
      src = p;
    }
  }
  if (!check(hipMemcpyAsync(buffer->ptr, src, size, hipMemcpyHostToDevice, batch->stream), "batch_buffer_write"))
    batch->failed = true;
}

void gpu_hip_batch_buffer_fill(GPU_Batch* batch, GPU_Buffer* buffer, U64 size, U32 value)
{
  if (!batch || !buffer) return;
  if (size > buffer->alloc_size) size = buffer->alloc_size;
  /* a fill kernel rather than hipMemsetAsync, so a captured batch replays
   * correctly */
  GPU_Kernel* k = gpu_hip_kernel_alloc(String8{(U8*)"fill_u32", 8});
  if (!k) {
    batch->failed = true;
    return;
  }
  QK_Args saved = k->args;
  U32 saved_missing = k->missing_mask;
  gpu_hip_kernel_set_arg_buffer(k, 0, buffer);
  gpu_hip_kernel_set_arg_u64(k, 0, size / 4);
  gpu_hip_kernel_set_arg_u64(k, 1, value);
  U64 words = size / 4;
  U64 threads = (words + 3) / 4;
  U64 cap = (U64)st().props.multiProcessorCount * 8 * 256;
  if (threads > cap) threads = cap;
  if (threads == 0) threads = 1;
  if (batch->deferred) {
    PreparedLaunch pl;
    bool empty = false;
    if (!prepare(k, (U32)threads, 256, pl, empty)) batch->failed = true;
    else if (!empty) batch->ops.push_back(BatchOp{2, nullptr, 0, nullptr, nullptr, pl});
  } else if (!launch(k, (U32)threads, 256, batch->stream, &batch->prof)) {
    batch->failed = true;
  }
  k->args = saved;
  k->missing_mask = saved_missing;
}

void gpu_hip_batch_buffer_zero(GPU_Batch* batch, GPU_Buffer* buffer, U64 size)
{
  gpu_hip_batch_buffer_fill(batch, buffer, size, 0);
}

void gpu_hip_batch_kernel_execute(GPU_Batch* batch, GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size)
{
  if (!batch) return;
  if (batch->deferred) {
    PreparedLaunch pl;
    bool empty = false;
    if (!prepare(kernel, global_work_size, local_work_size, pl, empty)) batch->failed = true;
    else if (!empty) batch->ops.push_back(BatchOp{2, nullptr, 0, nullptr, nullptr, pl});
    return;
  }
  if (!launch(kernel, global_work_size, local_work_size, batch->stream, &batch->prof)) batch->failed = true;
}

void gpu_hip_batch_buffer_read(GPU_Batch* batch, GPU_Buffer* buffer, void* out_data, U64 size)
{
  if (!batch) return;
  if (!buffer || !out_data || size > buffer->alloc_size) {
    /* a read that cannot be made fails the batch (its caller must not
     * take the unwritten bytes for a result) */
    set_error("batch_buffer_read: missing buffer or %lld bytes past it", "", (long long)size);
    batch->failed = true;
    return;
  }
  if (fault_read_now()) {
    set_error("batch_buffer_read: injected read failure");
    batch->failed = true;
    return;
  }
  if (batch->deferred) {
    char* p = staging_take(batch, size);
    if (!p) {
      batch->failed = true;
      return;
    }
    batch->ops.push_back(BatchOp{1, buffer->ptr, size, p, out_data, {}});
    return;
  }
  void* dst = out_data;
  if (size <= kStagingMax) {
    char* p = staging_take(batch, size);
    if (p) {
      dst = p;
      batch->reads.push_back(PendingRead{out_data, p, size});
    }
  }
  if (!check(hipMemcpyAsync(dst, buffer->ptr, size, hipMemcpyDeviceToHost, batch->stream), "batch_buffer_read"))
    batch->failed = true;
}

/* a deferred batch: replay the thread's graph for this key when the
 * recorded operations match it (only the bytes written may differ), else
 * capture the operations into a new graph; either way the work runs as one
 * graph launch */
static B32 batch_end_deferred(GPU_Batch* batch)
{
  bool ok = !batch->failed;
  std::vector<GraphEntry>& gs = graphs();
  GraphEntry* hit = nullptr;
  for (auto& e : gs)
    if (e.key == batch->key) hit = &e;
  const std::vector<BatchOp>* src = &batch->ops;
  PendingProfile pp;
  bool timed = st().profile;
  if (timed) {
    pp.name = "graph:" + batch->key.substr(0, batch->key.find(':', 6));
    hipEventCreate(&pp.start);
    hipEventCreate(&pp.stop);
  }
  if (ok && hit && ops_match(hit->ops, batch->ops)) {
    for (size_t i = 0; i < batch->ops.size(); i++)
      if (batch->ops[i].kind == 0) memcpy(hit->ops[i].staged, batch->ops[i].staged, batch->ops[i].size);
    if (timed) hipEventRecord(pp.start, batch->stream);
    ok = check(hipGraphLaunch(hit->exec, batch->stream), "hipGraphLaunch");
    if (timed) hipEventRecord(pp.stop, batch->stream);
    hit->tick = ++t_graph_tick;
    src = &hit->ops;
    g_graph_replays++;
  } else if (ok) {
    if (hit) {
      graph_destroy(*hit);
      gs.erase(gs.begin() + (hit - gs.data()));
      hit = nullptr;
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ok = check(hipStreamBeginCapture(batch->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (ok) {
      bool issued = issue_ops(batch->ops, batch->stream);
      ok = check(hipStreamEndCapture(batch->stream, &graph), "hipStreamEndCapture") && issued;
    }
    if (ok) ok = check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (ok) {
      if (timed) hipEventRecord(pp.start, batch->stream);
      ok = check(hipGraphLaunch(exec, batch->stream), "hipGraphLaunch");
      if (timed) hipEventRecord(pp.stop, batch->stream);
    }
    if (ok) {
      if (gs.size() >= kMaxGraphs) {
        size_t lru = 0;
        for (size_t i = 1; i < gs.size(); i++)
          if (gs[i].tick < gs[lru].tick) lru = i;
        graph_destroy(gs[lru]);
        gs.erase(gs.begin() + lru);
      }
      GraphEntry e;
      e.key = batch->key;
      e.graph = graph;
      e.exec = exec;
      e.ops = batch->ops;
      e.blocks.swap(batch->blocks);
      e.big_blocks.swap(batch->big_blocks);
      e.tick = ++t_graph_tick;
      gs.push_back(e);
      src = &gs.back().ops;
      g_graph_captures++;
    } else {
      if (exec) hipGraphExecDestroy(exec);
      if (graph) hipGraphDestroy(graph);
    }
  }
  ok = check(hipStreamSynchronize(batch->stream), "batch_end") && ok;
  for (size_t i = 0; i < batch->ops.size(); i++) {
    const BatchOp& op = batch->ops[i];
    if (op.kind != 1) continue;
    if (ok) memcpy(op.user, (*src)[i].staged, op.size);
    else memset(op.user, 0, op.size);
  }
  if (timed) {
    if (ok) batch->prof.push_back(pp);
    else {
      hipEventDestroy(pp.start);
      hipEventDestroy(pp.stop);
    }
  }
  staging_return(batch);
  flush_profile(batch->prof);
  give_stream(batch->stream);
  delete batch;
  return ok ? 1 : 0;
}

void* gpu_hip_host_alloc(U64 bytes)
{
  if (!st().initialised) gpu_hip_init();
  void* p = nullptr;
  if (!check(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc")) return nullptr;
  return p;
}

void gpu_hip_host_free(void* p)
{
  if (p) hipHostFree(p);
}

void gpu_hip_set_evict_hook(U64 (*hook)(U64 bytes_wanted)) { st().evict_hook = hook; }

GPU_Buffer* gpu_hip_buffer_view(GPU_Buffer* parent, U64 offset, U64 size)
{
  if (!parent || offset > parent->alloc_size || size > parent->alloc_size - offset) return nullptr;
  GPU_Buffer* b = new GPU_Buffer();
  b->ptr = (char*)parent->ptr + offset;
  b->size = size;
  b->alloc_size = size;
  b->flags = parent->flags;
  b->pooled = false;
  b->host_import = parent->host_import;
  b->wrapped = true;
  return b;
}

GPU_Buffer* gpu_hip_buffer_wrap_host(void* host_ptr, U64 size)
{
  if (!st().initialised) gpu_hip_init();
  if (!host_ptr) return nullptr;
  void* dptr = nullptr;
  if (!check(hipHostGetDevicePointer(&dptr, host_ptr, 0), "hipHostGetDevicePointer") || !dptr) return nullptr;
  GPU_Buffer* b = new GPU_Buffer();
  b->ptr = dptr;
  b->size = size;
  b->alloc_size = size;
  b->flags = GPU_BufferFlag_Read;
  b->pooled = false;
  b->host_import = true;
  b->wrapped = true;
  return b;
}

GPU_Batch* gpu_hip_batch_begin_graph(const char* key)
{
  GPU_Batch* b = gpu_hip_batch_begin(0, 0);
  if (!b) return nullptr;
  b->deferred = true;
  b->key = key ? key : "";
  return b;
}

void gpu_hip_graph_stats(U64* replays, U64* captures)
{
  if (replays) *replays = g_graph_replays.load();
  if (captures) *captures = g_graph_captures.load();
}

B32 gpu_hip_batch_end(GPU_Batch* batch)
{
  if (!batch) return 0;
  if (batch->deferred) return batch_end_deferred(batch);
  bool ok = check(hipStreamSynchronize(batch->stream), "batch_end") && !batch->failed;
  for (auto& r : batch->reads) {
    if (ok) memcpy(r.user, r.pinned, r.size);
    else memset(r.user, 0, r.size); /* readbacks of a failed batch are zero (gpu_vulkan.c:3107-3111) */
  }
  staging_return(batch);
  flush_profile(batch->prof);
  give_stream(batch->stream);
  delete batch;
  return ok ? 1 : 0;
}

GPU_Stream* gpu_hip_stream_alloc(U32 slot_count)
{
  if (!st().initialised) gpu_hip_init();
  if (slot_count == 0) slot_count = 1;
  GPU_Stream* s = new GPU_Stream();
  memset(&s->stats, 0, sizeof(s->stats));
  s->stats.separate_copy_queue = 1;
  s->slots.resize(slot_count);
  for (auto& sl : s->slots) {
    sl.stream = take_stream();
    hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    sl.submitted = false;
  }
  return s;
}

void gpu_hip_stream_release(GPU_Stream* stream)
{
  if (!stream) return;
  for (auto& sl : stream->slots) {
    hipStreamSynchronize(sl.stream);
    flush_profile(sl.prof);
    hipEventDestroy(sl.done);
    give_stream(sl.stream);
  }
  delete stream;
}

B32 gpu_hip_stream_slot_begin(GPU_Stream* stream, U32 slot, U64 upload_bytes, U64 readback_bytes)
{
  (void)upload_bytes;
  (void)readback_bytes;
  if (!stream || slot >= stream->slots.size()) return 0;
  GPU_StreamSlot& sl = stream->slots[slot];
  if (sl.submitted) {
    if (!check(hipEventSynchronize(sl.done), "stream_slot_begin")) return 0;
    sl.submitted = false;
  }
  return 1;
}

B32 gpu_hip_stream_copy(GPU_Stream* stream, U32 slot, GPU_Buffer* destination, void* host_ptr, U64 size)
{
  if (!stream || slot >= stream->slots.size() || !destination || !host_ptr) return 0;
  if (size > destination->alloc_size) size = destination->alloc_size;
  stream->stats.bytes_copied += size;
  return check(hipMemcpyAsync(destination->ptr, host_ptr, size, hipMemcpyHostToDevice, stream->slots[slot].stream),
               "stream_copy")
             ? 1
             : 0;
}

void gpu_hip_stream_fill(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, U64 size, U32 value)
{
  if (!stream || slot >= stream->slots.size()) return;
  GPU_Batch tmp;
  tmp.stream = stream->slots[slot].stream;
  tmp.failed = false;
  gpu_hip_batch_buffer_fill(&tmp, buffer, size, value);
  for (auto& p : tmp.prof) stream->slots[slot].prof.push_back(p);
}

void gpu_hip_stream_zero(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, U64 size)
{
  gpu_hip_stream_fill(stream, slot, buffer, size, 0);
}

void gpu_hip_stream_dispatch(GPU_Stream* stream, U32 slot, GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size)
{
  if (!stream || slot >= stream->slots.size()) return;
  launch(kernel, global_work_size, local_work_size, stream->slots[slot].stream, &stream->slots[slot].prof);
}

void gpu_hip_stream_read(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, void* out_data, U64 size)
{
  if (!stream || slot >= stream->slots.size() || !buffer || !out_data) return;
  if (size > buffer->alloc_size) size = buffer->alloc_size;
  check(hipMemcpyAsync(out_data, buffer->ptr, size, hipMemcpyDeviceToHost, stream->slots[slot].stream), "stream_read");
}

B32 gpu_hip_stream_submit(GPU_Stream* stream, U32 slot)
{
  if (!stream || slot >= stream->slots.size()) return 0;
  GPU_StreamSlot& sl = stream->slots[slot];
  if (!check(hipEventRecord(sl.done, sl.stream), "stream_submit")) return 0;
  sl.submitted = true;
  return 1;
}

B32 gpu_hip_stream_wait(GPU_Stream* stream, U32 slot)
{
  if (!stream || slot >= stream->slots.size()) return 0;
  GPU_StreamSlot& sl = stream->slots[slot];
  bool ok = check(hipStreamSynchronize(sl.stream), "stream_wait");
  sl.submitted = false;
  flush_profile(sl.prof);
  return ok ? 1 : 0;
}

GPU_StreamStats gpu_hip_stream_get_stats(GPU_Stream* stream)
{
  GPU_StreamStats z;
  memset(&z, 0, sizeof(z));
  return stream ? stream->stats : z;
}

void* gpu_hip_scratch_arena(void) { return nullptr; }

String8 gpu_hip_name(void) { return String8{(U8*)"hip", 3}; }

U32 gpu_hip_compute_units(void)
{
  if (!st().initialised) gpu_hip_init();
  return st().initialised ? (U32)st().props.multiProcessorCount : 0;
}

U64 gpu_hip_buffer_size(GPU_Buffer* b) { return b ? b->size : 0; }

void gpu_hip_kernel_set_lds_bytes(GPU_Kernel* kernel, U32 bytes)
{
  if (!kernel) return;
  if (bytes > 65536 && kernel->lib && bytes != kernel->lds_bytes)
    (void)hipFuncSetAttribute(kernel->lib->host_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  kernel->lds_bytes = bytes;
}

U32 gpu_hip_max_lds_bytes(void)
{
  if (!st().initialised) gpu_hip_init();
  if (!st().initialised) return 0;
  size_t optin = st().props.sharedMemPerBlockOptin;
  size_t base = st().props.sharedMemPerBlock;
  return (U32)(optin > base ? optin : base);
}

U32 gpu_hip_kernel_max_lds_bytes(GPU_Kernel* kernel)
{
  const U32 max = gpu_hip_max_lds_bytes();
  if (!kernel) return max;
  return kernel->static_lds < max ? max - kernel->static_lds : 0u;
}

U64 gpu_hip_buffer_device_address(GPU_Buffer* b) { return b ? (U64)(uintptr_t)b->ptr : 0; }

void gpu_hip_profile_enable(B32 enabled) { st().profile = enabled != 0; }

void gpu_hip_profile_tag(const char* tag)
{
  snprintf(t_profile_tag, sizeof(t_profile_tag), "%s", tag ? tag : "");
}

void gpu_hip_profile_reset(void)
{
  std::lock_guard<std::mutex> g(st().lock);
  st().profile_table.clear();
  st().kernel_time_us = 0;
}

U64 gpu_hip_profile_query(const char* name, U64* out_dispatches)
{
  std::lock_guard<std::mutex> g(st().lock);
  U64 ns = 0, n = 0;
  for (auto& kv : st().profile_table) {
    size_t len = name ? strlen(name) : 0;
    bool prefix = len > 0 && name[len - 1] == '*';
    if (!name || !name[0] || kv.first == name || (prefix && kv.first.compare(0, len - 1, name, len - 1) == 0)) {
      ns += kv.second.first;
      n += kv.second.second;
    }
  }
  if (out_dispatches) *out_dispatches = n;
  return ns;
}

void gpu_hip_fault_alloc_at(U64 k)
{
  std::lock_guard<std::mutex> g(st().lock);
  st().fault_alloc_at = k;
  st().fault_alloc_counter = 0;
}

void gpu_hip_fault_read_at(U64 k)
{
  std::lock_guard<std::mutex> g(st().lock);
  st().fault_read_at = k;
  st().fault_read_counter = 0;
}

const char* gpu_hip_last_error(void) { return t_last_error; }

GPU_Backend* gpu_hip_backend(void)
{
  static GPU_Backend b = {
      (U32)sizeof(GPU_Backend),
      gpu_hip_init,
      gpu_hip_release,
      gpu_hip_wait,
      gpu_hip_get_executed_kernel_time_microseconds,
      gpu_hip_device_total_memory,
      gpu_hip_device_free_memory,
      gpu_hip_device_max_storage_buffer_range,
      gpu_hip_device_lost,
      gpu_hip_pool_begin_epoch,
      gpu_hip_memory_stats,
      gpu_hip_buffer_alloc,
      gpu_hip_buffer_alloc_pooled,
      gpu_hip_buffer_import_host_readonly,
      gpu_hip_buffer_import_host_readonly_pooled,
      gpu_hip_buffer_release,
      gpu_hip_buffer_write,
      gpu_hip_buffer_read,
      gpu_hip_kernel_alloc,
      gpu_hip_kernel_alloc_from_spirv,
      gpu_hip_kernel_alloc_from_spirv_spec,
      gpu_hip_generated_kernel_stats,
      gpu_hip_kernel_release,
      gpu_hip_kernel_execute,
      gpu_hip_kernel_set_arg_buffer,
      gpu_hip_kernel_set_arg_u64,
      gpu_hip_batch_begin,
      gpu_hip_batch_buffer_write,
      gpu_hip_batch_buffer_zero,
      gpu_hip_batch_buffer_fill,
      gpu_hip_batch_kernel_execute,
      gpu_hip_batch_buffer_read,
      gpu_hip_batch_end,
      gpu_hip_stream_alloc,
      gpu_hip_stream_release,
      gpu_hip_stream_slot_begin,
      gpu_hip_stream_copy,
      gpu_hip_stream_zero,
      gpu_hip_stream_fill,
      gpu_hip_stream_dispatch,
      gpu_hip_stream_read,
      gpu_hip_stream_submit,
      gpu_hip_stream_wait,
      gpu_hip_stream_get_stats,
      gpu_hip_scratch_arena,
      gpu_hip_name,
  };
  return &b;
}

}  // extern "C"
