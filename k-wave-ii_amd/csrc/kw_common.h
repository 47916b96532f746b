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

// Loads and stores at a 32-bit element index from a base pointer.  Every
// array they address is below 4 GiB (checked on the host), so the byte
// offset fits 32 bits.
template <class T>
KW_HOSTDEV const T* at32(const T* base, uint32_t i) {
  return (const T*)((const char*)base + (size_t)(uint32_t)(i * (uint32_t)sizeof(T)));
}
template <class T>
KW_HOSTDEV T* at32(T* base, uint32_t i) {
  return (T*)((char*)base + (size_t)(uint32_t)(i * (uint32_t)sizeof(T)));
}
#ifdef KW_GPU
// On the GPU they are raw buffer instructions: the base (uniform in every
// kernel here) becomes a buffer resource in scalar registers and each lane
// supplies one 32-bit byte offset.  With plain global instructions the
// compiler hoists zero-extended 64-bit offsets out of loops and spends two
// VGPRs per address.  The resource spans the whole 32-bit offset range
// (gfx9 resource word 3: 0x00020000, untyped 32-bit data).
typedef __amdgpu_buffer_rsrc_t kw_rsrc;
__device__ __forceinline__ kw_rsrc kw_res(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)0xffffffff,
                                           0x00020000);
}
template <class T>
__device__ __forceinline__ T kw_ld32(const T* base, uint32_t i) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte elements");
  const int off = (int)(i * (uint32_t)sizeof(T));
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(kw_res(base), off, 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(kw_res(base), off, 0, 0));
  }
}
template <class T, class V>
__device__ __forceinline__ void kw_st32(T* base, uint32_t i, V v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte elements");
  const int off = (int)(i * (uint32_t)sizeof(T));
  const T w = (T)v;
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, w), kw_res(base), off, 0,
                                          0);
  } else {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, w), kw_res(base), off, 0, 0);
  }
}
#define KW_LD32(base, i) (::kw::kw_ld32((base), (uint32_t)(i)))
#define KW_ST32(base, i, v) (::kw::kw_st32((base), (uint32_t)(i), (v)))
#else
#define KW_LD32(base, i) KW_LD(::kw::at32((base), (uint32_t)(i)))
#define KW_ST32(base, i, v) KW_ST(::kw::at32((base), (uint32_t)(i)), (v))
#endif

// The value of v, opaque to the optimiser.  The FFT engines derive every
// LDS, twiddle and element index from the thread's index in its pencil; fed
// a fresh copy of it on entry, each transform computes these indices where
// it uses them, instead of sharing them with earlier transforms of the
// kernel or hoisting them out of the loops over field components, which
// keeps dozens of them live in registers across the whole kernel and cuts
// the waves per SIMD.
#ifdef KW_GPU
__device__ __forceinline__ int kw_fresh(int v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}
#else
inline int kw_fresh(int v) { return v; }
#endif

// Scheduling fence between epilogue phases: each phase batches its own
// loads, but loads are not hoisted across phases (which would hold every
// phase's operands in registers at once).
#ifdef KW_GPU
#define KW_PHASE_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define KW_PHASE_FENCE() ((void)0)
#endif

#ifdef KW_EMU
// every LDS access is checked by the emulator for races between lanes: a
// read of an element written by another lane, or a write of a different
// value over another lane's write, with no barrier of the required scope in
// between
void emu_lds_access(const cf* p, bool write, const cf* value);
inline cf lds_ld(const cf* p) {
  emu_lds_access(p, false, nullptr);
  return *p;
}
inline void lds_st(cf* p, cf v) {
  emu_lds_access(p, true, &v);
  *p = v;
}
#elif !defined(KW_GPU)
inline cf lds_ld(const cf* p) { return *p; }
inline void lds_st(cf* p, cf v) { *p = v; }
#endif

// Barrier of a pencil's threads.  When every pencil of a block lies inside
// one 64-lane wavefront (the G threads of pencil p are tid p*G .. p*G+G-1
// and G divides 64), LDS exchanges within a pencil only need the wavefront
// to be ordered, not the block: `wave` selects a wavefront barrier instead
// of a block barrier.  `wave` must be uniform across the block.
template <class Ctx>
struct PencilCtx {
  Ctx c;  // contexts are small and copied by value
  bool wave;
  KW_BODY int tid() const { return c.tid(); }
  KW_BODY int bid() const { return c.bid(); }
  KW_BODY int nthreads() const { return c.nthreads(); }
  KW_BODY cf* lds() const { return c.lds(); }
  KW_BODY void sync() const {
    if (wave)
      c.wsync();
    else
      c.sync();
  }
};

KW_HOSTDEV bool pencils_in_wave(int G) { return G > 0 && G <= 64 && 64 % G == 0; }

// Elements held per thread per FFT stage (upper bound).  A pencil of length
// N is processed by G = ceil(N / KW_E) threads.
#ifndef KW_E
#define KW_E 8
#endif

#define KW_MAX_STAGES 12
#define KW_MAX_RADIX 8

}  // namespace kw
