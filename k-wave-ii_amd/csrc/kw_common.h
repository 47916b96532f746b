// Shared definitions for the kwhip kernels.
//
// Every kernel body in this library is a template over an execution context
// `Ctx` that provides the thread id, block id, a block barrier and the block's
// LDS.  Built with hipcc the context maps onto threadIdx/blockIdx/
// __syncthreads/dynamic LDS and the bodies run as gfx950 kernels.  Built with
// g++ and KW_EMU defined, the same bodies run on the CPU under a lane-per-
// thread emulator (kw_emu.cpp) in which every global-memory access is checked
// against the live allocations, so indexing can be validated without a GPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__) && !defined(KW_EMU)
#include <hip/hip_runtime.h>
#define KW_GPU 1
#define KW_BODY __device__ __forceinline__
#define KW_HOSTDEV __host__ __device__ __forceinline__
#else
#include <cmath>
#define KW_BODY inline
#define KW_HOSTDEV inline
#endif

namespace kw {

// complex float in HBM and LDS: (re, im), 8-byte aligned so that loads and
// stores are single 64-bit accesses
struct alignas(8) cf {
  float x, y;
};

KW_HOSTDEV cf cmake(float a, float b) { return cf{a, b}; }
KW_HOSTDEV cf cadd(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
KW_HOSTDEV cf csub(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
KW_HOSTDEV cf cmul(cf a, cf b) { return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
KW_HOSTDEV cf cscale(cf a, float s) { return cf{a.x * s, a.y * s}; }
KW_HOSTDEV cf cconj(cf a) { return cf{a.x, -a.y}; }
// multiply by -i and by +i
KW_HOSTDEV cf cmul_mi(cf a) { return cf{a.y, -a.x}; }
KW_HOSTDEV cf cmul_pi(cf a) { return cf{-a.y, a.x}; }

#ifdef KW_EMU
// Bounds check of one global access against the emulator's allocation
// table; an out-of-bounds access is recorded (the launch then fails) and
// redirected to a scratch word so the emulated kernel can finish.
const void* emu_resolve(const void* p, size_t bytes, const char* what);
template <class T>
inline T* emu_ptr(T* p, const char* what) {
  return (T*)emu_resolve((const void*)p, sizeof(T), what);
}
#define KW_LD(ptr) (*::kw::emu_ptr((ptr), "load"))
#define KW_ST(ptr, val) (*::kw::emu_ptr((ptr), "store") = (val))
#else
#define KW_LD(ptr) (*(ptr))
#define KW_ST(ptr, val) (*(ptr) = (val))
#endif

// Elements held per thread per FFT stage (upper bound).  A pencil of length
// N is processed by G = ceil(N / KW_E) threads.
#ifndef KW_E
#define KW_E 8
#endif

#define KW_MAX_STAGES 12
#define KW_MAX_RADIX 8

}  // namespace kw
