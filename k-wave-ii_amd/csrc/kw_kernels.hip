// gfx950 kernels.  This file is compiled once per kernel (KW_KERNEL = 0..8)
// so that the kernels build in parallel; each compile defines one kernel and
// its host launcher.
#include <hip/hip_runtime.h>

#include "kw_backend.h"

extern __shared__ kw::cf kw_lds_dyn[];

namespace kw {

struct GpuCtx {
  __device__ __forceinline__ int tid() const { return (int)threadIdx.x; }
  __device__ __forceinline__ int bid() const { return (int)blockIdx.x; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
  __device__ __forceinline__ cf* lds() const { return kw_lds_dyn; }
};

#define KW_DEFINE(ID, NAME, BOUNDS, ARGT, CALL)                                        \
  __global__ __launch_bounds__(BOUNDS) void NAME(Batch B, ARGT A) {                    \
    GpuCtx c;                                                                          \
    CALL;                                                                              \
  }                                                                                    \
  void launch_##NAME(int nb, int T, size_t lds, void* s, const Batch& B, const ARGT& A) { \
    hipLaunchKernelGGL(NAME, dim3(nb), dim3(T), lds, (hipStream_t)s, B, A);              \
  }

// x passes run at most 256 threads per block; strided and last-axis passes
// at most 1024
#if KW_KERNEL == 0
KW_DEFINE(0, k_yfwd, 1024, PassArgs, body_ypass<false>(c, B, A))
#elif KW_KERNEL == 1
KW_DEFINE(1, k_yinv, 1024, PassArgs, body_ypass<true>(c, B, A))
#elif KW_KERNEL == 2
KW_DEFINE(2, k_spec, 1024, PassArgs, body_spec(c, B, A))
#elif KW_KERNEL == 3
KW_DEFINE(3, k_x_r2c, 256, PassArgs, body_x_r2c(c, B, A))
#elif KW_KERNEL == 4
KW_DEFINE(4, k_x_c2r, 256, PassArgs, body_x_c2r(c, B, A))
#elif KW_KERNEL == 5
KW_DEFINE(5, k_x_grad, 256, PassArgs, body_x_grad(c, B, A))
#elif KW_KERNEL == 6
KW_DEFINE(6, k_x_div, 256, PassArgs, body_x_div(c, B, A))
#elif KW_KERNEL == 7
KW_DEFINE(7, k_x_eos, 256, PassArgs, body_x_eos(c, B, A))
#elif KW_KERNEL == 8
KW_DEFINE(8, k_ew, 256, EwArgs, body_ew(c, B, A))
#else
#error "KW_KERNEL must be 0..8"
#endif

}  // namespace kw
