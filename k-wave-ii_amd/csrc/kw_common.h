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
#elif defined(KW_GPU)
// HBM accesses go through the global address space explicitly: the
// pointers come from descriptors in memory, so without the cast the
// compiler emits generic (flat) accesses, which wait on both the vector and
// the LDS counters and serialise loads
#define KW_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T kw_ld(const T* p) {
  return *((const KW_GLOBAL T*)p);
}
template <class T, class V>
__device__ __forceinline__ void kw_st(T* p, V v) {
  *((KW_GLOBAL T*)p) = (T)v;
}
// complex values move as a native 2-float vector: copying the struct through
// an address-space pointer would bind its copy constructor to a generic
// reference and bring the flat access back
typedef float kw_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cf kw_ld(const cf* p) {
  const kw_v2f v = *((const KW_GLOBAL kw_v2f*)p);
  return cf{v.x, v.y};
}
__device__ __forceinline__ void kw_st(cf* p, cf v) {
  const kw_v2f w = {v.x, v.y};
  *((KW_GLOBAL kw_v2f*)p) = w;
}
#define KW_LD(ptr) (::kw::kw_ld(ptr))
#define KW_ST(ptr, val) (::kw::kw_st((ptr), (val)))
#define KW_LDS_PTR __attribute__((address_space(3)))
__device__ __forceinline__ cf lds_ld(const cf* p) {
  const kw_v2f v = *((const KW_LDS_PTR kw_v2f*)p);
  return cf{v.x, v.y};
}
__device__ __forceinline__ void lds_st(cf* p, cf v) {
  const kw_v2f w = {v.x, v.y};
  *((KW_LDS_PTR kw_v2f*)p) = w;
}
#else
#define KW_LD(ptr) (*(ptr))
#define KW_ST(ptr, val) (*(ptr) = (val))
#endif

// Loads and stores at a 32-bit element index from a base pointer: the byte
// offset is formed in 32 bits so that the compiler can address with a
// uniform 64-bit base plus a 32-bit per-lane offset (one VGPR per address
// instead of a 64-bit pointer).  Every array is below 4 GiB (checked on the
// host).
template <class T>
KW_HOSTDEV const T* at32(const T* base, uint32_t i) {
  return (const T*)((const char*)base + (size_t)(uint32_t)(i * (uint32_t)sizeof(T)));
}
template <class T>
KW_HOSTDEV T* at32(T* base, uint32_t i) {
  return (T*)((char*)base + (size_t)(uint32_t)(i * (uint32_t)sizeof(T)));
}
#define KW_LD32(base, i) KW_LD(::kw::at32((base), (uint32_t)(i)))
#define KW_ST32(base, i, v) KW_ST(::kw::at32((base), (uint32_t)(i)), (v))

// Scheduling fence between epilogue phases: each phase batches its own
// loads, but loads are not hoisted across phases (which would hold every
// phase's operands in registers at once).
#ifdef KW_GPU
#define KW_PHASE_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define KW_PHASE_FENCE() ((void)0)
#endif

#ifndef KW_GPU
inline cf lds_ld(const cf* p) { return *p; }
inline void lds_st(cf* p, cf v) { *p = v; }
#endif

// Elements held per thread per FFT stage (upper bound).  A pencil of length
// N is processed by G = ceil(N / KW_E) threads.
#ifndef KW_E
#define KW_E 8
#endif

#define KW_MAX_STAGES 12
#define KW_MAX_RADIX 8

}  // namespace kw
