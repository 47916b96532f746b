// HIP backend: device memory, streams, events and kernel dispatch for
// kw_sim.cpp.  The kernels themselves are built from kw_kernels.hip.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <string>

#include "kw_backend.h"
#include "kw_variants.h"
#include "kwhip.h"

namespace kw {

#define KW_DECLARE_XL(SIZE)                                                       \
  KW_DECLARE_VARIANT(2, SIZE)                                                     \
  KW_DECLARE_VARIANT(3, SIZE)                                                     \
  KW_DECLARE_VARIANT(4, SIZE)                                                     \
  KW_DECLARE_VARIANT(5, SIZE)                                                     \
  KW_DECLARE_VARIANT(6, SIZE)                                                     \
  KW_DECLARE_VARIANT(7, SIZE)
#define KW_DECLARE_ALL(SIZE)                                                      \
  KW_DECLARE_VARIANT(0, SIZE)                                                     \
  KW_DECLARE_VARIANT(1, SIZE)                                                     \
  KW_DECLARE_XL(SIZE)
KW_DECLARE_ALL(0)
KW_FIXED_SIZES_ALL(KW_DECLARE_ALL)
KW_FIXED_SIZES_XL(KW_DECLARE_XL)
void launch_k_ew(int, int, size_t, void*, const Batch&, const EwArgs&);

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? -4 : -3;
    throw DeviceError(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}

const char* backend_name() { return "hip-gfx950"; }

void* dev_alloc(size_t bytes) {
  void* p = nullptr;
  check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
  return p;
}
void dev_free(void* p) {
  if (p) (void)hipFree(p);
}
void* dev_host_alloc(size_t bytes) {
  void* p = nullptr;
  check(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
  return p;
}
void dev_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}
void* dev_stream_create() {
  hipStream_t s;
  check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  return (void*)s;
}
void dev_stream_destroy(void* s) {
  if (s) (void)hipStreamDestroy((hipStream_t)s);
}
void dev_h2d(void* dst, const void* src, size_t bytes, void* s) {
  check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s), "h2d");
  check(hipStreamSynchronize((hipStream_t)s), "h2d sync");
}
void dev_d2h(void* dst, const void* src, size_t bytes, void* s) {
  check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s), "d2h");
  check(hipStreamSynchronize((hipStream_t)s), "d2h sync");
}
void dev_sync(void* s) { check(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize"); }
void dev_zero(void* dst, size_t bytes, void* s) {
  check(hipMemsetAsync(dst, 0, bytes, (hipStream_t)s), "hipMemsetAsync");
  check(hipStreamSynchronize((hipStream_t)s), "memset sync");
}

size_t dev_max_lds() {
  static const size_t v = []() {
    int dev = 0, b = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
          "hipDeviceGetAttribute");
    return (size_t)b;
  }();
  return v;
}

// Scratch of kw_run_noalloc: every device buffer of one simulation (fields,
// spectra, operators, staging of the inputs) is carved out of this array, so
// that call makes no device allocation.  In host code `scratch` is only a
// host-side handle of the symbol; its address on a device comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) unsigned char scratch[KW_NOALLOC_MAX_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};
}  // namespace

void* dev_scratch(size_t* bytes) {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) throw DeviceError(-3, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw DeviceError(-3, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[dev] = p;
  }
  if (bytes) *bytes = (size_t)KW_NOALLOC_MAX_BYTES;
  return g_scratch_ptr[dev];
}

void dev_mem_info(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
}

int fixed_variant(int N, PassKind kind) {
  switch (N) {
#define KW_ALL_CASE(SIZE) case SIZE:
    KW_FIXED_SIZES_ALL(KW_ALL_CASE)
    return N;
#undef KW_ALL_CASE
#define KW_XL_CASE(SIZE) case SIZE:
    KW_FIXED_SIZES_XL(KW_XL_CASE)
    return (kind == PK_YFWD || kind == PK_YINV) ? 0 : N;
#undef KW_XL_CASE
    default:
      return 0;
  }
}

void launch_pass(PassKind kind, int fixed, int dim, int nblocks, int threads, size_t lds_bytes,
                 const Batch& B, const PassArgs& A, void* stream) {
  if (nblocks <= 0) return;
#define KW_CASE_KIND(KIND, SIZE) \
  case KIND: launch_v_##KIND##_##SIZE(dim, nblocks, threads, lds_bytes, stream, B, A); break;
#define KW_CASE_ALL(SIZE)         \
  case SIZE:                      \
    switch ((int)kind) {          \
      KW_CASE_KIND(0, SIZE)       \
      KW_CASE_KIND(1, SIZE)       \
      KW_CASE_KIND(2, SIZE)       \
      KW_CASE_KIND(3, SIZE)       \
      KW_CASE_KIND(4, SIZE)       \
      KW_CASE_KIND(5, SIZE)       \
      KW_CASE_KIND(6, SIZE)       \
      KW_CASE_KIND(7, SIZE)       \
    }                             \
    break;
#define KW_CASE_XL(SIZE)                                                          \
  case SIZE:                                                                      \
    switch ((int)kind) {                                                          \
      KW_CASE_KIND(2, SIZE)                                                       \
      KW_CASE_KIND(3, SIZE)                                                       \
      KW_CASE_KIND(4, SIZE)                                                       \
      KW_CASE_KIND(5, SIZE)                                                       \
      KW_CASE_KIND(6, SIZE)                                                       \
      KW_CASE_KIND(7, SIZE)                                                       \
      default: throw DeviceError(-3, "no strided-pass variant for this length"); \
    }                                                                             \
    break;
  switch (fixed) {
    KW_CASE_ALL(0)
    KW_FIXED_SIZES_ALL(KW_CASE_ALL)
    KW_FIXED_SIZES_XL(KW_CASE_XL)
    default: throw DeviceError(-3, "no kernel variant for this length");
  }
#undef KW_CASE_XL
#undef KW_CASE_ALL
#undef KW_CASE_KIND
  check(hipGetLastError(), "kernel launch");
}

void launch_ew(int nblocks, const Batch& B, const EwArgs& A, void* stream) {
  if (nblocks <= 0) return;
  launch_k_ew(nblocks, EW_T, 0, stream, B, A);
  check(hipGetLastError(), "kernel launch");
}

void* dev_event_create() {
  hipEvent_t e;
  check(hipEventCreate(&e), "hipEventCreate");
  return (void*)e;
}
void dev_event_destroy(void* e) {
  if (e) (void)hipEventDestroy((hipEvent_t)e);
}
void dev_event_record(void* e, void* s) {
  check(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord");
}
// thread-local capture: the admission thread may allocate and copy while the
// stepping thread captures
bool dev_capture_begin(void* s) {
  check(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeThreadLocal),
        "hipStreamBeginCapture");
  return true;
}
void* dev_capture_end(void* s) {
  hipGraph_t g = nullptr;
  check(hipStreamEndCapture((hipStream_t)s, &g), "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  check(e, "hipGraphInstantiate");
  return (void*)x;
}
void dev_capture_abort(void* s) {
  hipGraph_t g = nullptr;
  (void)hipStreamEndCapture((hipStream_t)s, &g);
  if (g) (void)hipGraphDestroy(g);
}
void dev_graph_launch(void* x, void* s) {
  check(hipGraphLaunch((hipGraphExec_t)x, (hipStream_t)s), "hipGraphLaunch");
}
void dev_graph_destroy(void* x) {
  if (x) (void)hipGraphExecDestroy((hipGraphExec_t)x);
}

bool dev_event_done(void* e) {
  const hipError_t r = hipEventQuery((hipEvent_t)e);
  if (r == hipErrorNotReady) return false;
  check(r, "hipEventQuery");
  return true;
}
double dev_event_elapsed_ms(void* a, void* b) {
  check(hipEventSynchronize((hipEvent_t)b), "hipEventSynchronize");
  float ms = 0.f;
  check(hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
  return (double)ms;
}

}  // namespace kw
