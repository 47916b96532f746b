// HIP backend: device memory, streams, events and kernel dispatch for
// kw_sim.cpp.  The kernels themselves are built from kw_kernels.hip.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "kw_backend.h"

namespace kw {

void launch_k_yfwd(int, int, size_t, void*, const Batch&, const PassArgs&);
void launch_k_yinv(int, int, size_t, void*, const Batch&, const PassArgs&);
void launch_k_spec(int, int, size_t, void*, const Batch&, const PassArgs&);
void launch_k_x_r2c(int, int, size_t, void*, const Batch&, const PassArgs&);
void launch_k_x_c2r(int, int, size_t, void*, const Batch&, const PassArgs&);
void launch_k_x_grad(int, int, size_t, void*, const Batch&, const PassArgs&);
void launch_k_x_div(int, int, size_t, void*, const Batch&, const PassArgs&);
void launch_k_x_eos(int, int, size_t, void*, const Batch&, const PassArgs&);
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

size_t dev_max_lds() { return 65536; }

void launch_pass(PassKind kind, int nblocks, int threads, size_t lds_bytes, const Batch& B,
                 const PassArgs& A, void* stream) {
  if (nblocks <= 0) return;
  switch (kind) {
    case PK_YFWD: launch_k_yfwd(nblocks, threads, lds_bytes, stream, B, A); break;
    case PK_YINV: launch_k_yinv(nblocks, threads, lds_bytes, stream, B, A); break;
    case PK_SPEC: launch_k_spec(nblocks, threads, lds_bytes, stream, B, A); break;
    case PK_X_R2C: launch_k_x_r2c(nblocks, threads, lds_bytes, stream, B, A); break;
    case PK_X_C2R: launch_k_x_c2r(nblocks, threads, lds_bytes, stream, B, A); break;
    case PK_X_GRAD: launch_k_x_grad(nblocks, threads, lds_bytes, stream, B, A); break;
    case PK_X_DIV: launch_k_x_div(nblocks, threads, lds_bytes, stream, B, A); break;
    case PK_X_EOS: launch_k_x_eos(nblocks, threads, lds_bytes, stream, B, A); break;
  }
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
double dev_event_elapsed_ms(void* a, void* b) {
  check(hipEventSynchronize((hipEvent_t)b), "hipEventSynchronize");
  float ms = 0.f;
  check(hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
  return (double)ms;
}

}  // namespace kw
