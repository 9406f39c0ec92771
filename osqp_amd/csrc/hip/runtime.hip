// runtime.hip — process-wide state of the algebra/hip backend.
//
// OSQP calls osqp_algebra_init_libs() in every osqp_setup and
// osqp_algebra_free_libs() in every osqp_cleanup (src/osqp_api.c:378,1068);
// the CUDA backend frees its global handle on the first cleanup
// (algebra/cuda/algebra_libs.cu:50-56), which breaks other live solvers.
// Here the state is reference-counted.  Device memory comes from a caching
// allocator (size-bucketed free lists) so that per-request setup on a stream
// of small QPs does not pay hipMalloc/hipFree each time.
#include <mutex>
#include <string.h>
#include <unordered_map>
#include <vector>

#include "hip_internal.h"

namespace osqp_hip {

static Runtime g_rt;
static std::mutex g_mu;

Runtime& rt() { return g_rt; }

int check(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return 0;
  if (!g_rt.last_error) {
    fprintf(stderr, "osqp_amd: HIP error %d (%s) in %s at %s:%d\n", (int)e, hipGetErrorString(e), what,
            file, line);
  }
  g_rt.last_error = (int)e;
  return (int)e;
}

// ---- caching allocator ------------------------------------------------------
struct Cache {
  std::unordered_map<size_t, std::vector<void*>> free_lists;  // bucket -> blocks
  std::unordered_map<void*, size_t> sizes;                   // live + cached blocks
  size_t cached_bytes = 0;
};
static Cache g_cache;

static size_t bucket_of(size_t bytes) {
  size_t b = 256;
  while (b < bytes) b <<= 1;
  return b;
}

void* dev_alloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  size_t b = bucket_of(bytes);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.free_lists.find(b);
  if (it != g_cache.free_lists.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    g_cache.cached_bytes -= b;
    return p;
  }
  void* p = nullptr;
  if (hipMalloc(&p, b) != hipSuccess) {
    // release cached blocks and retry once
    HIPCHK(hipStreamSynchronize(g_rt.stream));
    for (auto& kv : g_cache.free_lists)
      for (void* q : kv.second) {
        g_cache.sizes.erase(q);
        (void)hipFree(q);
      }
    g_cache.free_lists.clear();
    g_cache.cached_bytes = 0;
    (void)hipGetLastError();
    if (HIPCHK(hipMalloc(&p, b))) return nullptr;
  }
  g_cache.sizes[p] = b;
  return p;
}

// Blocks go back to the cache; reuse is ordered after earlier work because all
// kernels and copies run on the single backend stream.
void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.sizes.find(p);
  if (it == g_cache.sizes.end()) return;
  g_cache.free_lists[it->second].push_back(p);
  g_cache.cached_bytes += it->second;
}

static void release_cache() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_cache.sizes) (void)hipFree(kv.first);
  g_cache.sizes.clear();
  g_cache.free_lists.clear();
  g_cache.cached_bytes = 0;
}

// src may be host (pageable or pinned) or device memory: unified addressing
// lets hipMemcpyDefault resolve the direction.  A pageable source must stay
// valid only until the call returns, so the copy is completed here.
void upload(void* dst, const void* src, size_t bytes) {
  if (!bytes || !src) return;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, g_rt.stream));
  HIPCHK(hipStreamSynchronize(g_rt.stream));
}

void download(void* dst, const void* src, size_t bytes) {
  if (!bytes || !dst) return;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, g_rt.stream));
  HIPCHK(hipStreamSynchronize(g_rt.stream));
}

}  // namespace osqp_hip

using namespace osqp_hip;

extern "C" {

OSQPInt osqp_amd_ptr_is_device(const void* p) {
  if (!p) return 0;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // plain host memory: not an error for the caller
    return 0;
  }
  return a.type == hipMemoryTypeDevice ? 1 : 0;
}

OSQPInt osqp_amd_copy_to_host(void* dst, const void* src, size_t bytes) {
  if (!bytes) return 0;
  if (!osqp_amd_ptr_is_device(src)) {
    memcpy(dst, src, bytes);
    return 0;
  }
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 1;
}

OSQPInt osqp_algebra_linsys_supported(void) { return OSQP_CAPABILITY_INDIRECT_SOLVER; }

enum osqp_linsys_solver_type osqp_algebra_default_linsys(void) { return OSQP_INDIRECT_SOLVER; }

OSQPInt osqp_algebra_init_libs(OSQPInt device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_rt.refcount > 0) {
    if (device != g_rt.device) return 1;  // one device per process
    g_rt.refcount++;
    return 0;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return 1;
  if (HIPCHK(hipSetDevice(device))) return 1;
  // default priority stream (no priority streams on this platform)
  if (HIPCHK(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking))) return 1;
  // reduction results go from the last block of the reducing kernel straight
  // to host-coherent memory (no device-to-host copy after it)
  if (HIPCHK(hipHostMalloc((void**)&g_rt.h_scalars, 64 * sizeof(double),
                           hipHostMallocMapped | hipHostMallocCoherent)))
    return 1;
  if (HIPCHK(hipHostGetDevicePointer((void**)&g_rt.dh_scalars, g_rt.h_scalars, 0))) return 1;
  memset(g_rt.h_scalars, 0, 64 * sizeof(double));
  g_rt.red_seq = 0.0;
  if (HIPCHK(hipMalloc((void**)&g_rt.d_partials, 4 * kMaxReduceBlocks * sizeof(double)))) return 1;
  if (HIPCHK(hipMalloc((void**)&g_rt.d_ticket, 64 * sizeof(unsigned)))) return 1;
  if (HIPCHK(hipMemset(g_rt.d_ticket, 0, 64 * sizeof(unsigned)))) return 1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    g_rt.num_cu = prop.multiProcessorCount;
    snprintf(g_rt.device_name, sizeof(g_rt.device_name), "%s (%s, %d CUs)", prop.name,
             prop.gcnArchName, prop.multiProcessorCount);
  }
  g_rt.device = device;
  g_rt.refcount = 1;
  g_rt.last_error = 0;
  return 0;
}

void osqp_algebra_free_libs(void) {
  bool last = false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_rt.refcount <= 0) return;
    g_rt.refcount--;
    last = g_rt.refcount == 0;
  }
  if (!last) return;
  HIPCHK(hipStreamSynchronize(g_rt.stream));
  release_cache();
  (void)hipFree(g_rt.d_partials);
  (void)hipFree(g_rt.d_ticket);
  (void)hipHostFree(g_rt.h_scalars);
  (void)hipStreamDestroy(g_rt.stream);
  g_rt.d_partials = nullptr;
  g_rt.d_ticket = nullptr;
  g_rt.h_scalars = nullptr;
  g_rt.dh_scalars = nullptr;
  g_rt.stream = nullptr;
  g_rt.device = -1;
}

OSQPInt osqp_algebra_name(char* name, OSQPInt nameLen) {
  return snprintf(name, (size_t)nameLen, "%s", "HIP gfx950 (MI355X)");
}

OSQPInt osqp_algebra_device_name(char* name, OSQPInt nameLen) {
  return snprintf(name, (size_t)nameLen, "%s", g_rt.device_name[0] ? g_rt.device_name : "none");
}

void osqp_amd_algebra_sync(void) {
  if (g_rt.stream) HIPCHK(hipStreamSynchronize(g_rt.stream));
}

OSQPInt osqp_amd_last_error(void) {
  int e = g_rt.last_error;
  g_rt.last_error = 0;
  return e;
}

}  // extern "C"
