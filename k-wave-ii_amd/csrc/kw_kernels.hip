// gfx950 kernels.  This file is compiled once per (pass kind, fixed length)
// variant (KW_KIND = 0..7 and KW_SIZE = 0 for runtime lengths or one of the
// fixed lengths of kw_variants.h), or with KW_KIND = 8 for the elementwise
// kernel, so that the variants build in parallel; each compile defines its
// kernels and one host launcher.
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>

#include "kw_backend.h"
#include "kw_variants.h"

extern __shared__ kw::cf kw_lds_dyn[];

namespace kw {

struct GpuCtx {
  __device__ __forceinline__ int tid() const { return (int)threadIdx.x; }
  __device__ __forceinline__ int bid() const { return (int)blockIdx.x; }
  __device__ __forceinline__ int nthreads() const { return (int)blockDim.x; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
  // wavefront barrier: LDS writes of the wavefront before it are visible to
  // its LDS reads after it (a wavefront's LDS operations execute in order;
  // the fences keep the compiler from moving accesses across)
  __device__ __forceinline__ void wsync() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __device__ __forceinline__ cf* lds() const { return kw_lds_dyn; }
};

#define KW_CAT2(a, b) a##b
#define KW_CAT(a, b) KW_CAT2(a, b)

// Kernels take the problem descriptors and the block-to-problem table as
// read-only, non-aliased parameters: nothing a kernel stores can change
// them, so the compiler may fetch descriptor fields with scalar loads and
// keep every field base address in scalar registers.
#define KW_KPARAMS \
  const ProbDev* __restrict__ probs, const int32_t* __restrict__ tiles, int32_t nprob
#define KW_KBATCH const Batch B{probs, tiles, nprob}

#if KW_KIND == 8

__global__ __launch_bounds__(256) void k_ew(KW_KPARAMS, EwArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_ew(c, B, A);
}
void launch_k_ew(int nb, int T, size_t lds, void* s, const Batch& B, const EwArgs& A) {
  hipLaunchKernelGGL(k_ew, dim3(nb), dim3(T), lds, (hipStream_t)s, B.probs, B.tile_start,
                     B.nprob, A);
}

#else

// kernel names unique per variant: kv_<kind>_<size>[_d2]
#define kern KW_CAT(KW_CAT(KW_CAT(kv_, KW_KIND), _), KW_SIZE)
#define kern2 KW_CAT(kern, _d2)

#if KW_SIZE == 0
using Sel = PlanRt;
using SelX = PlanRt;
#else
using Sel = PlanCt<KW_SIZE>;
using SelX = PlanCtX<KW_SIZE>;
#endif

// x passes run at most 256 threads per block; strided and last-axis passes
// at most 1024 (32 pencils of 32 threads)
#if KW_KIND == 0
__global__ __launch_bounds__(1024) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_ypass<false, Sel>(c, B, A);
}
#elif KW_KIND == 1
__global__ __launch_bounds__(1024) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_ypass<true, Sel>(c, B, A);
}
#elif KW_KIND == 2
__global__ __launch_bounds__(1024) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_spec<Sel, false>(c, B, A);
}
// the sensor-record modes of the last-axis pass
#define kern_aux KW_CAT(kern, _rec)
__global__ __launch_bounds__(1024) void kern_aux(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_spec<Sel, true>(c, B, A);
}
#define KW_HAS_AUX 1
#elif KW_KIND == 3
__global__ __launch_bounds__(512) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_x_r2c<SelX>(c, B, A);
}
#elif KW_KIND == 4
__global__ __launch_bounds__(512) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_x_c2r<SelX>(c, B, A);
}
#elif KW_KIND == 5
__global__ __launch_bounds__(512) void kern2(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_x_grad<SelX, 2>(c, B, A);
}
__global__ __launch_bounds__(512) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_x_grad<SelX, 3>(c, B, A);
}
#define KW_HAS_DIM2 1
#elif KW_KIND == 6
__global__ __launch_bounds__(512) void kern2(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_x_div<SelX, 2>(c, B, A);
}
__global__ __launch_bounds__(512) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_x_div<SelX, 3>(c, B, A);
}
#define KW_HAS_DIM2 1
#elif KW_KIND == 7
__global__ __launch_bounds__(512) void kern(KW_KPARAMS, PassArgs A) {
  GpuCtx c;
  KW_KBATCH;
  body_x_eos<SelX>(c, B, A);
}
#else
#error "KW_KIND must be 0..8"
#endif

// dynamic LDS above 64 KiB per block (gfx950 has 160 KiB) is requested per
// kernel once, before its first such launch
static void lds_limit(const void* f, std::atomic<int>& done, size_t lds) {
  if (lds <= 65536 || done.load(std::memory_order_acquire)) return;
  const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)dev_max_lds());
  if (e != hipSuccess) throw DeviceError(-3, std::string("hipFuncSetAttribute: ") +
                                                 hipGetErrorString(e));
  done.store(1, std::memory_order_release);
}

void KW_CAT(KW_CAT(KW_CAT(launch_v_, KW_KIND), _), KW_SIZE)(int dim, int nb, int T, size_t lds,
                                                            void* s, const Batch& B,
                                                            const PassArgs& A) {
  static std::atomic<int> done_main{0};
#ifdef KW_HAS_DIM2
  static std::atomic<int> done_d2{0};
  lds_limit((const void*)kern2, done_d2, lds);
#endif
#ifdef KW_HAS_AUX
  static std::atomic<int> done_aux{0};
  lds_limit((const void*)kern_aux, done_aux, lds);
#endif
  lds_limit((const void*)kern, done_main, lds);
#ifdef KW_HAS_DIM2
  if (dim == 2) {
    hipLaunchKernelGGL(kern2, dim3(nb), dim3(T), lds, (hipStream_t)s, B.probs, B.tile_start,
                       B.nprob, A);
    return;
  }
#endif
  (void)dim;
#ifdef KW_HAS_AUX
  if (spec_is_aux(A)) {
    hipLaunchKernelGGL(kern_aux, dim3(nb), dim3(T), lds, (hipStream_t)s, B.probs, B.tile_start,
                       B.nprob, A);
    return;
  }
#endif
  hipLaunchKernelGGL(kern, dim3(nb), dim3(T), lds, (hipStream_t)s, B.probs, B.tile_start,
                     B.nprob, A);
}

#endif

}  // namespace kw
