// Device backend used by the host orchestration (kw_sim.cpp).  Implemented
// by kw_gpu.hip (HIP on gfx950: the product) and by kw_emu.cpp (the CPU
// lane emulator used to validate kernel indexing without a GPU).
#pragma once

#include <stdexcept>
#include <string>

#include "kw_passes.h"

namespace kw {

struct DeviceError : std::runtime_error {
  int code;
  DeviceError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

const char* backend_name();
void* dev_alloc(size_t bytes);  // throws DeviceError(KW_ENOMEM) on failure
void dev_free(void* p);
// pinned host memory the device reads and writes directly (the engine's
// per-simulation failure flags); freed with dev_host_free
void* dev_host_alloc(size_t bytes);
void dev_host_free(void* p);
void* dev_stream_create();
void dev_stream_destroy(void* s);
void dev_h2d(void* dst, const void* src, size_t bytes, void* stream);
void dev_d2h(void* dst, const void* src, size_t bytes, void* stream);
void dev_sync(void* stream);
void dev_zero(void* dst, size_t bytes, void* stream);  // set-up only (not graph-safe)

// compile-time FFT length of the variant built for pass `kind` on an axis of
// length N, or 0 when only the runtime-length variant exists
int fixed_variant(int N, PassKind kind);

// launches (asynchronous on `stream`)
// `fixed` is the compile-time FFT length of the variant to run (0: the
// runtime-length variant); `dim` selects the 2D/3D body of the x passes
void launch_pass(PassKind kind, int fixed, int dim, int nblocks, int threads, size_t lds_bytes,
                 const Batch& B, const PassArgs& A, void* stream);
void launch_ew(int nblocks, const Batch& B, const EwArgs& A, void* stream);

// event timing (GPU only; the emulator returns wall-clock times)
void* dev_event_create();
void dev_event_destroy(void* e);
void dev_event_record(void* e, void* stream);
double dev_event_elapsed_ms(void* a, void* b);  // synchronises on b
bool dev_event_done(void* e);                   // the work before e has finished

// Graphs of launches: capture what one host thread launches on `stream`
// between begin and end into an executable graph, replay it, destroy it.
// dev_capture_begin returns false when the backend cannot capture (the
// emulator); the launches then have to be issued directly.
bool dev_capture_begin(void* stream);
void* dev_capture_end(void* stream);
void dev_capture_abort(void* stream);  // end a capture and drop what it recorded
void dev_graph_launch(void* exec, void* stream);
void dev_graph_destroy(void* exec);

// largest dynamic LDS per block the kernels may request
size_t dev_max_lds();

// The library's static scratch region on the current device (a file-scope
// device array of *bytes bytes, the same array on every call, held until the
// process ends); its device address is looked up once per device.  Used by
// the entry point that allocates nothing.
void* dev_scratch(size_t* bytes);

// free and total device memory in bytes
void dev_mem_info(size_t* free_bytes, size_t* total_bytes);

}  // namespace kw
