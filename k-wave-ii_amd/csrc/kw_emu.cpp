// CPU emulator backend: runs the kernel bodies of kw_passes.h with one
// cooperative fiber per lane (ucontext), one block at a time per worker
// thread, so barriers behave as on the GPU.  Every global load and store is
// checked against the table of live allocations; LDS is filled with NaN at
// the start of every block so reads of unwritten LDS show up in results; a
// block whose lanes pass different sequences of barriers is reported.
//
// Lanes are grouped into 64-lane wavefronts as on gfx950.  Block barriers
// (sync) and wavefront barriers (wsync) both yield to the scheduler, which
// resumes the lanes round-robin; with every lane passing the same sequence
// of barriers (checked) each lane's n-th segment runs after all lanes'
// (n-1)-th segments, which is what both barrier kinds guarantee.  What the
// round-robin order does not show, a missing barrier scope, is checked on
// every LDS read: an element written by a lane of another wavefront must
// have a block barrier between the write and the read, an element written
// by another lane of the same wavefront at least a wavefront barrier.
//
// This backend exists to validate indexing and parity on a machine without
// a GPU.  It is built into libkwhip_emu.so, which only the tests load.
#ifndef KW_EMU
#define KW_EMU 1
#endif

#include <ucontext.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <vector>

#include "kw_backend.h"
#include "kw_variants.h"
#include "kwhip.h"

namespace kw {

namespace {
// the allocation table is read by every emulated access and written by
// allocations, which may come from another host thread (the engine admits
// simulations while it steps others)
std::shared_mutex g_mu;
std::map<uintptr_t, size_t> g_allocs;
std::atomic<int> g_fault{0};
char g_fault_msg[512];
alignas(64) char g_scratch[256];
}  // namespace

const void* emu_resolve(const void* p, size_t bytes, const char* what) {
  const uintptr_t a = (uintptr_t)p;
  bool ok;
  {
    std::shared_lock<std::shared_mutex> lk(g_mu);
    auto it = g_allocs.upper_bound(a);
    ok = it != g_allocs.begin();
    if (ok) {
      --it;
      ok = a >= it->first && a + bytes <= it->first + it->second;
    }
  }
  if (!ok) {
    if (g_fault.exchange(1) == 0)
      snprintf(g_fault_msg, sizeof g_fault_msg, "emulator: out-of-bounds %s of %zu bytes at %p",
               what, bytes, p);
    std::memset(g_scratch, 0, sizeof g_scratch);
    return g_scratch;
  }
  return p;
}

static void emu_check(const void* p, size_t bytes, const char* what) {
  (void)emu_resolve(p, bytes, what);
}

const char* backend_name() { return "emulator"; }

void* dev_alloc(size_t bytes) {
  const size_t n = bytes ? bytes : 1;
  void* p = std::aligned_alloc(256, (n + 255) / 256 * 256);
  if (!p) throw DeviceError(-4, "emulator: allocation failed");
  std::unique_lock<std::shared_mutex> lk(g_mu);
  g_allocs[(uintptr_t)p] = n;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  {
    std::unique_lock<std::shared_mutex> lk(g_mu);
    g_allocs.erase((uintptr_t)p);
  }
  std::free(p);
}

// the emulator's stand-in for the device scratch array: one host block of
// the same size, made on first use, registered as one allocation and kept
// until the process ends
void* dev_scratch(size_t* bytes) {
  static void* block = nullptr;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (!block) block = dev_alloc((size_t)KW_NOALLOC_MAX_BYTES);
  if (bytes) *bytes = (size_t)KW_NOALLOC_MAX_BYTES;
  return block;
}

void* dev_host_alloc(size_t bytes) { return dev_alloc(bytes); }
void dev_host_free(void* p) { dev_free(p); }
void* dev_stream_create() { return nullptr; }
void dev_stream_destroy(void*) {}
void dev_h2d(void* dst, const void* src, size_t bytes, void*) {
  emu_check(dst, bytes, "h2d");
  std::memcpy(dst, src, bytes);
}
void dev_d2h(void* dst, const void* src, size_t bytes, void*) {
  emu_check(src, bytes, "d2h");
  std::memcpy(dst, src, bytes);
}
void dev_sync(void*) {}
void dev_zero(void* dst, size_t bytes, void*) {
  emu_check(dst, bytes, "zero");
  std::memset(dst, 0, bytes);
}
size_t dev_max_lds() { return 163840; }  // gfx950: 160 KiB per block

// the emulator's "device" is host memory; report a fixed 16 GiB
void dev_mem_info(size_t* free_bytes, size_t* total_bytes) {
  size_t used = 0;
  {
    std::shared_lock<std::shared_mutex> lk(g_mu);
    for (auto& a : g_allocs) used += a.second;
  }
  const size_t total = (size_t)16 << 30;
  if (free_bytes) *free_bytes = used < total ? total - used : 0;
  if (total_bytes) *total_bytes = total;
}

namespace {

constexpr size_t kStack = 96 * 1024;

struct Runner;
thread_local Runner* tl_runner = nullptr;

struct EmuCtx {
  Runner* R;
  int t, b, nt;
  cf* L;
  int tid() const { return t; }
  int bid() const { return b; }
  int nthreads() const { return nt; }
  cf* lds() const { return L; }
  void sync() const;
  void wsync() const;
};

struct Lane {
  ucontext_t ctx;
  std::vector<char> stack;
  bool done;
  long syncs;   // block barriers passed
  long wsyncs;  // wavefront barriers passed
  uint64_t seq; // hash of the sequence of barrier kinds
};

struct Runner {
  ucontext_t sched;
  std::vector<Lane> lanes;
  std::vector<cf> lds;
  // last writer of each LDS element: lane, its block-barrier count and its
  // total barrier count at the write
  std::vector<int32_t> lw_lane;
  std::vector<long> lw_sync, lw_tick;
  int cur = 0;
  int bid = 0;
  int nthr = 0;
  const std::function<void(EmuCtx&)>* body = nullptr;

  void yield(int kind) {
    Lane& l = lanes[cur];
    if (kind == 0)
      ++l.syncs;
    else
      ++l.wsyncs;
    l.seq = l.seq * 1000003u + (uint64_t)(kind + 1);
    swapcontext(&l.ctx, &sched);
  }

  static void trampoline() {
    Runner* R = tl_runner;
    const int i = R->cur;
    EmuCtx c{R, i, R->bid, R->nthr, R->lds.data()};
    (*R->body)(c);
    R->lanes[i].done = true;
  }

  void run_block(int b, int T) {
    bid = b;
    nthr = T;
    if ((int)lanes.size() < T) lanes.resize(T);
    const float qnan = std::nanf("");
    for (auto& v : lds) v = cf{qnan, qnan};
    lw_lane.assign(lds.size(), -1);
    lw_sync.assign(lds.size(), 0);
    lw_tick.assign(lds.size(), 0);
    for (int i = 0; i < T; ++i) {
      Lane& l = lanes[i];
      if (l.stack.empty()) l.stack.resize(kStack);
      l.done = false;
      l.syncs = 0;
      l.wsyncs = 0;
      l.seq = 0;
      getcontext(&l.ctx);
      l.ctx.uc_stack.ss_sp = l.stack.data();
      l.ctx.uc_stack.ss_size = l.stack.size();
      l.ctx.uc_link = &sched;
      makecontext(&l.ctx, (void (*)())trampoline, 0);
    }
    for (;;) {
      int alive = 0;
      for (int i = 0; i < T; ++i) {
        if (lanes[i].done) continue;
        cur = i;
        swapcontext(&sched, &lanes[i].ctx);
        if (!lanes[i].done) ++alive;
      }
      if (alive == 0) break;
    }
    for (int i = 1; i < T; ++i) {
      if (lanes[i].syncs != lanes[0].syncs || lanes[i].wsyncs != lanes[0].wsyncs ||
          lanes[i].seq != lanes[0].seq) {
        if (g_fault.exchange(1) == 0)
          snprintf(g_fault_msg, sizeof g_fault_msg,
                   "emulator: block %d lane %d passed %ld block and %ld wavefront barriers, "
                   "lane 0 passed %ld and %ld (or in another order)",
                   b, i, lanes[i].syncs, lanes[i].wsyncs, lanes[0].syncs, lanes[0].wsyncs);
        break;
      }
    }
  }
};

void EmuCtx::sync() const { R->yield(0); }
void EmuCtx::wsync() const { R->yield(1); }

void lds_fault(const char* msg, int b, int lane, long el, int other) {
  if (g_fault.exchange(1) == 0)
    snprintf(g_fault_msg, sizeof g_fault_msg, "emulator: block %d lane %d LDS element %ld: %s %d",
             b, lane, el, msg, other);
}

void emu_launch(int nblocks, int T, size_t lds_bytes,
                const std::function<void(EmuCtx&)>& body) {
  if (nblocks <= 0) return;
  if (T <= 0 || T > 1024) throw DeviceError(-3, "emulator: bad block size");
  int nw = (int)std::thread::hardware_concurrency();
  if (nw < 1) nw = 1;
  if (nw > 8) nw = 8;
  if (nw > nblocks) nw = nblocks;
  std::atomic<int> next{0};
  auto work = [&]() {
    Runner R;
    R.body = &body;
    R.lds.resize(lds_bytes / sizeof(cf) + 1);
    tl_runner = &R;
    for (;;) {
      const int b = next.fetch_add(1);
      if (b >= nblocks) break;
      R.run_block(b, T);
    }
    tl_runner = nullptr;
  };
  std::vector<std::thread> th;
  for (int w = 1; w < nw; ++w) th.emplace_back(work);
  work();
  for (auto& x : th) x.join();
  if (g_fault.load()) {
    g_fault = 0;
    throw DeviceError(-3, g_fault_msg);
  }
}

}  // namespace

void emu_lds_access(const cf* p, bool write, const cf* value) {
  Runner* R = tl_runner;
  if (!R) return;
  const long i = (long)(p - R->lds.data());
  const int me = R->cur;
  if (i < 0 || i >= (long)R->lds.size()) {
    lds_fault("out of bounds of the block's LDS of size", R->bid, me, i, (int)R->lds.size());
    return;
  }
  const Lane& l = R->lanes[me];
  const int w = R->lw_lane[i];
  const bool ordered = w < 0 || w == me ||
                       ((w / 64 != me / 64) ? l.syncs > R->lw_sync[i]
                                            : l.syncs + l.wsyncs > R->lw_tick[i]);
  if (write) {
    // two lanes writing different values with nothing ordering them
    if (!ordered && std::memcmp(value, p, sizeof(cf)) != 0)
      lds_fault("write of a different value without a barrier after the write of lane",
                R->bid, me, i, w);
    R->lw_lane[i] = me;
    R->lw_sync[i] = l.syncs;
    R->lw_tick[i] = l.syncs + l.wsyncs;
    return;
  }
  if (!ordered) lds_fault("read without a barrier after the write of lane", R->bid, me, i, w);
}

template <class Sel, class SelX>
static void emu_pass(PassKind kind, int dim, int nblocks, int threads, size_t lds_bytes,
                     const Batch& B, const PassArgs& A) {
  std::function<void(EmuCtx&)> f;
  switch (kind) {
    case PK_YFWD: f = [&](EmuCtx& c) { body_ypass<false, Sel>(c, B, A); }; break;
    case PK_YINV: f = [&](EmuCtx& c) { body_ypass<true, Sel>(c, B, A); }; break;
    case PK_SPEC:
      if (spec_is_aux(A))
        f = [&](EmuCtx& c) { body_spec<Sel, true>(c, B, A); };
      else
        f = [&](EmuCtx& c) { body_spec<Sel, false>(c, B, A); };
      break;
    case PK_X_R2C: f = [&](EmuCtx& c) { body_x_r2c<SelX>(c, B, A); }; break;
    case PK_X_C2R: f = [&](EmuCtx& c) { body_x_c2r<SelX>(c, B, A); }; break;
    case PK_X_GRAD:
      if (dim == 2)
        f = [&](EmuCtx& c) { body_x_grad<SelX, 2>(c, B, A); };
      else
        f = [&](EmuCtx& c) { body_x_grad<SelX, 3>(c, B, A); };
      break;
    case PK_X_DIV:
      if (dim == 2)
        f = [&](EmuCtx& c) { body_x_div<SelX, 2>(c, B, A); };
      else
        f = [&](EmuCtx& c) { body_x_div<SelX, 3>(c, B, A); };
      break;
    case PK_X_EOS: f = [&](EmuCtx& c) { body_x_eos<SelX>(c, B, A); }; break;
  }
  emu_launch(nblocks, threads, lds_bytes, f);
}

void launch_pass(PassKind kind, int fixed, int dim, int nblocks, int threads, size_t lds_bytes,
                 const Batch& B, const PassArgs& A, void*) {
  switch (fixed) {
    case 0: emu_pass<PlanRt, PlanRt>(kind, dim, nblocks, threads, lds_bytes, B, A); break;
#define KW_EMU_SIZE(SIZE) \
    case SIZE: emu_pass<PlanCt<SIZE>, PlanCtX<SIZE>>(kind, dim, nblocks, threads, lds_bytes, B, A); break;
    KW_EMU_FIXED_SIZES(KW_EMU_SIZE)
#undef KW_EMU_SIZE
    default: throw DeviceError(-3, "no kernel variant for this length");
  }
}

int fixed_variant(int N, PassKind) {
  switch (N) {
#define KW_EMU_CASE(SIZE) case SIZE:
    KW_EMU_FIXED_SIZES(KW_EMU_CASE)
    return N;
#undef KW_EMU_CASE
    default:
      return 0;
  }
}

void launch_ew(int nblocks, const Batch& B, const EwArgs& A, void*) {
  std::function<void(EmuCtx&)> f = [&](EmuCtx& c) { body_ew(c, B, A); };
  emu_launch(nblocks, EW_T, 0, f);
}

struct EmuEvent {
  std::chrono::steady_clock::time_point t;
};
void* dev_event_create() { return new EmuEvent(); }
void dev_event_destroy(void* e) { delete (EmuEvent*)e; }
void dev_event_record(void* e, void*) { ((EmuEvent*)e)->t = std::chrono::steady_clock::now(); }
bool dev_event_done(void*) { return true; }  // emulated launches are synchronous
bool dev_capture_begin(void*) { return false; }
void* dev_capture_end(void*) { throw DeviceError(-3, "emulator: no graph capture"); }
void dev_capture_abort(void*) {}
void dev_graph_launch(void*, void*) { throw DeviceError(-3, "emulator: no graph capture"); }
void dev_graph_destroy(void*) {}
double dev_event_elapsed_ms(void* a, void* b) {
  return std::chrono::duration<double, std::milli>(((EmuEvent*)b)->t - ((EmuEvent*)a)->t)
      .count();
}

}  // namespace kw
