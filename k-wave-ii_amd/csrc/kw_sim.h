// Internal host-side interface of kwhip: the simulation object (device
// allocations, operators, launch geometry, per-launch byte accounting) and
// the step functions, shared by the single-simulation C ABI (kw_sim.cpp) and
// the engine that steps many simulations together (kw_engine.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "kw_backend.h"
#include "kw_variants.h"
#include "kwhip.h"

namespace kw {
[[noreturn]] inline void fail(int code, const std::string& m) { throw DeviceError(code, m); }

constexpr double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// operators (double precision)
// ---------------------------------------------------------------------------

// radices for a length: 8s first, then 4 or 2, then 3s and 5s
inline bool factorise(int N, std::vector<int>& r) {
  r.clear();
  int n = N;
  while (n % 8 == 0) {
    r.push_back(8);
    n /= 8;
  }
  if (n % 4 == 0) {
    r.push_back(4);
    n /= 4;
  }
  if (n % 2 == 0) {
    r.push_back(2);
    n /= 2;
  }
  while (n % 3 == 0) {
    r.push_back(3);
    n /= 3;
  }
  while (n % 5 == 0) {
    r.push_back(5);
    n /= 5;
  }
  // radix 7 in the runtime engine only (pencils of such lengths run with
  // G = ceil(N/7) threads: one radix-7 butterfly each)
  while (n % 7 == 0) {
    r.push_back(7);
    n /= 7;
  }
  // N = 1 (the y axis of a 1-D problem): no stages, the transform is the
  // identity
  return n == 1 && (N == 1 || !r.empty()) && (int)r.size() <= KW_MAX_STAGES;
}

// length of the chirp-z convolution for a transform of length N that does
// not factor into 2, 3, 5 and 7: the smallest 2^a 3^b 5^c >= 2N - 1
inline int chirp_length(int N) {
  for (int L = std::max(1, 2 * N - 1);; ++L) {
    int m = L;
    for (int p : {2, 3, 5})
      while (m % p == 0) m /= p;
    if (m == 1) return L;
  }
}

// length the FFT stages of an axis of length N run at
inline int stage_length(int N) {
  std::vector<int> r;
  return factorise(N, r) ? N : chirp_length(N);
}

// wavenumber of ifftshifted index i (kWaveGrid.m:779-797 then ifftshift):
// (2 pi/d) * (n/N) with n = i for i < ceil(N/2), else i - N
inline double kvec(int i, int N, double d) {
  const int n = (i < (N + 1) / 2) ? i : i - N;
  return (2.0 * kPi / d) * ((double)n / (double)N);
}

// getPML.m:52-86
inline std::vector<double> pml_profile(int N, double d, double dt, double c, int size,
                                       double alpha, bool staggered) {
  std::vector<double> v(N, 1.0);
  for (int x = 1; x <= size && x <= N; ++x) {
    const double xs = staggered ? x + 0.5 : (double)x;
    const double left = alpha * (c / d) * std::pow((xs - size - 1) / (0.0 - size), 4);
    const double right = alpha * (c / d) * std::pow(xs / size, 4);
    v[x - 1] = std::exp(-left * dt / 2.0);
    v[N - size + x - 1] = std::exp(-right * dt / 2.0);
  }
  return v;
}

inline double sinc(double x) { return x == 0.0 ? 1.0 : std::sin(x) / x; }

// f(i) for i < n on up to 16 host threads; `work` is the cost of one item in
// element operations (small loops stay on the calling thread)
template <class F>
void parallel_for(int64_t n, F f, int64_t work = 1) {
  int nt = (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, 16));
  if (n * work < 100000) nt = 1;
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int w = 0; w < nt; ++w) {
    const int64_t a = w * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=]() {
      for (int64_t i = a; i < b; ++i) f(i);
    });
  }
  for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// simulation object
// ---------------------------------------------------------------------------

struct Geometry {
  int dim, Nx, Ny, Nz, Wm, nyq;
  int64_t ncell, nrows, nspec_main, nspec;
};

inline Geometry make_geometry(int dim, int64_t Nx, int64_t Ny, int64_t Nz) {
  Geometry g;
  g.dim = dim;
  g.Nx = (int)Nx;
  g.Ny = (int)Ny;
  g.Nz = dim == 3 ? (int)Nz : 1;
  g.nyq = (g.Nx % 2 == 0) ? 1 : 0;
  g.Wm = g.nyq ? g.Nx / 2 : (g.Nx + 1) / 2;
  g.ncell = (int64_t)g.Nx * g.Ny * g.Nz;
  g.nrows = (int64_t)g.Ny * g.Nz;
  g.nspec_main = (int64_t)g.Wm * g.Ny * g.Nz;
  g.nspec = g.nspec_main + (g.nyq ? g.nrows : 0);
  return g;
}

struct LaunchGeom {
  int nb = 0, T = 0;
  size_t lds = 0;
};

// LDS of a launch of pass k, whose geometry reserves `full` bytes.  The
// strided and last-axis passes start their LDS with the N twiddles of the
// transform, then the pencils.  The last-axis pass holds two pencil buffers
// per pencil (la, lb); with a
// fixed-length plan only the gradient uses lb (its two inverse transforms
// both read la), so the other modes launch with la alone and fit twice as
// many blocks per CU (the velocity split keeps its sum in lb too).
inline size_t launch_lds(PassKind k, int fixed, const PassArgs& a, size_t full) {
  // `full` holds the N twiddles and the two pencil buffers
  const size_t tw = sizeof(cf) * (size_t)fixed;
  return (k == PK_SPEC && fixed && a.mode != SM_GRAD && a.mode != SM_SPLIT)
             ? (full - tw) / 2 + tw
             : full;
}

// strided y passes over several buffers give each buffer blocks of its own
// (KW_YSPLIT=0: one block loops over the buffers)
inline bool ysplit_on() {
  static const bool on = [] {
    const char* e = std::getenv("KW_YSPLIT");
    return !(e && std::strcmp(e, "0") == 0);
  }();
  return on;
}

inline int max_pencils_env() {
  const char* e = std::getenv("KW_MAX_PENCILS");
  if (!e) return 1 << 30;
  const int v = std::atoi(e);
  return v > 0 ? v : 1 << 30;
}

// pencils per block of the strided y pass and of the last-axis pass
// (each pencil row of a block is P consecutive complex values in HBM)
inline int pencils_env(const char* name, int dflt) {
  const char* e = std::getenv(name);
  if (!e) return dflt;
  const int v = std::atoi(e);
  return v > 0 ? v : dflt;
}

// Pencils per block of the three pass families (x passes, strided y
// passes, last-axis passes); 0 takes the built-in rule (256 threads per x
// block, 16 pencils per y and last-axis block).  The tile changes which
// block handles a pencil, never the arithmetic on it, so results do not
// depend on it.
struct Tiles {
  int x = 0, y = 0, l = 0;
};

// Bump allocator over a device region the caller owns (the static scratch
// array of the entry point that allocates nothing): 256-byte aligned
// buffers, released all at once when the region is reused.
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  size_t used = 0;
};

struct Sim {
  Geometry G{};
  ProbDev P{};
  ProbDev* dP = nullptr;
  int32_t* d_tile0 = nullptr;
  Tiles tiles;  // pencils per block chosen for this grid (plan_launches)
  void* stream = nullptr;
  bool own_stream = true;
  Arena* arena = nullptr;  // set: buffers come from the arena, not dev_alloc
  std::vector<void*> allocs;
  int64_t bytes = 0;
  int64_t steps_done = 0;
  int64_t Nt = 0, Nt_eff = 0, n_rec = 0, nsens = 0, record_start = 1;
  bool has_p0 = false, time_rev = false, full_mask = false, directivity = false;
  uint32_t record = 0;
  bool p_grid = false, u_grid[3] = {false, false, false};
  float* skappa = nullptr;
  float* p0 = nullptr;
  std::vector<float*> dir_tab;  // k-space filter per unique directivity angle
  LaunchGeom lx, ly, ll, lcell, lsens;
  // what the dt-dependent operators are built from (a new time step
  // rebuilds κ and the PML profiles: kw_sim_change_dt)
  struct OpParams {
    double d[3] = {0, 0, 0};
    double c_ref = 0;
    int pml_size[3] = {0, 0, 0};
    double pml_alpha[3] = {0, 0, 0};
    bool use_sg = true, use_k = true;
  } op;
  // a time-step change armed for the next step: the tables of the
  // transition step's velocity half (mean step) and those of the new step
  struct DtChange {
    bool armed = false;
    float dt_mean = 0.f, dt_new = 0.f;
    float* kappa_mix = nullptr;  // two-step κ, velocity half
    float* ksplit = nullptr;     // split correction (nullptr: none)
    float* pml_sg_mean[3] = {nullptr, nullptr, nullptr};
    float* kappa_new = nullptr;
    float* pml_new[3] = {nullptr, nullptr, nullptr};
    float* pml_sg_new[3] = {nullptr, nullptr, nullptr};
  } dtc;
  float* scratch_grid = nullptr;  // one real grid (the split correction)
  // axisymmetric problems: the signed inverse radius along the expanded y
  // axis and κ / Ny for the u_r / r term; as_ws: the 'WSWS-FFT' layout
  // (2 No - 2 points, every copy symmetric) instead of 'WSWA-FFT' (4 No)
  bool axisym = false;
  bool as_ws = false;
  // 'WSWS-FFT': the reference's last radial column of u_y and the mirror
  // value the transforms read ([2 Nx]), its coefficient per x, and the
  // staggered radial PML at that column and the one before
  float* as_col = nullptr;
  float* as_k = nullptr;
  double as_a = 1.0, as_ap = 1.0;
  float* as_ir = nullptr;
  float* as_kappa = nullptr;
  // per-kernel event timing
  bool timing = false;
  std::vector<std::pair<void*, void*>> ev_used;
  std::vector<int> ev_kind;
  std::vector<double> ev_bytes;
  std::vector<void*> ev_free;

  ~Sim() {
    for (auto& e : ev_used) {
      dev_event_destroy(e.first);
      dev_event_destroy(e.second);
    }
    for (void* e : ev_free) dev_event_destroy(e);
    for (void* p : allocs) dev_free(p);
    if (own_stream) dev_stream_destroy(stream);
  }

  template <class T>
  T* alloc(size_t n) {
    const size_t nb = n * sizeof(T);
    void* p;
    if (arena) {
      const size_t need = ((nb ? nb : 1) + 255) / 256 * 256;
      if (need > arena->size - arena->used)
        fail(KW_ENOMEM, "the problem needs more device memory than the " +
                            std::to_string(arena->size >> 20) +
                            " MiB scratch of the no-allocation entry point (at least " +
                            std::to_string((arena->used + need + (1 << 20) - 1) >> 20) +
                            " MiB so far)");
      p = arena->base + arena->used;
      arena->used += need;
    } else {
      p = dev_alloc(nb);
      allocs.push_back(p);
    }
    bytes += (int64_t)nb;
    return (T*)p;
  }
  // free a buffer of alloc() before the simulation ends (no-op in an arena)
  void release(const void* q) {
    if (!q || arena) return;
    for (size_t i = 0; i < allocs.size(); ++i)
      if (allocs[i] == q) {
        dev_free(allocs[i]);
        allocs.erase(allocs.begin() + (long)i);
        return;
      }
  }
  template <class T>
  T* alloc_zero(size_t n) {
    T* p = alloc<T>(n);
    dev_zero(p, n * sizeof(T), stream);
    return p;
  }
  template <class T>
  T* upload(const std::vector<T>& v) {
    T* d = alloc<T>(v.size());
    dev_h2d(d, v.data(), v.size() * sizeof(T), stream);
    return d;
  }

  Batch batch() const { return Batch{dP, d_tile0, 1}; }

  void* take_event() {
    if (!ev_free.empty()) {
      void* e = ev_free.back();
      ev_free.pop_back();
      return e;
    }
    return dev_event_create();
  }

  template <class L>
  void timed(int kind, double nbytes, L launch) {
    if (!timing) {
      launch();
      return;
    }
    void* a = take_event();
    void* b = take_event();
    dev_event_record(a, stream);
    launch();
    dev_event_record(b, stream);
    ev_used.push_back({a, b});
    ev_kind.push_back(kind);
    ev_bytes.push_back(nbytes);
  }

  // compile-time FFT length of the kernel variant for a pass (0: runtime)
  int fixed_len(PassKind k) const {
    if (std::getenv("KW_NO_FIXED")) return 0;
    const int N = (k == PK_YFWD || k == PK_YINV) ? G.Ny
                  : (k == PK_SPEC)               ? (G.dim == 3 ? G.Nz : G.Ny)
                                                 : G.Nx;
    return fixed_variant(N, k);
  }

  void pass(PassKind k, const PassArgs& a0) {
    const LaunchGeom* g = &lx;
    if (k == PK_YFWD || k == PK_YINV) g = &ly;
    if (k == PK_SPEC) g = &ll;
    const Batch B = batch();
    const int fx = fixed_len(k);
    PassArgs a = a0;
    int nb = g->nb;
    if ((k == PK_YFWD || k == PK_YINV) && a.nbuf > 1 && ysplit_on()) {
      a.ysplit = nb;
      nb *= a.nbuf;
    }
    timed((int)k, timing ? pass_bytes(k, a) : 0.0,
          [&]() { launch_pass(k, fx, G.dim, nb, g->T, launch_lds(k, fx, a, g->lds), B, a, stream); });
  }

  void ew(const EwArgs& a, int nblocks) {
    const Batch B = batch();
    timed(8, timing ? ew_bytes(a) : 0.0, [&]() { launch_ew(nblocks, B, a, stream); });
  }

  // -- compulsory bytes of one launch: every array it reads or writes, once
  double fb(const Field& f) const { return f.grid() ? 4.0 * G.ncell : 0.0; }
  double src_bytes(const TimeSource& s, int t) const {
    if (s.len < t) return 0.0;
    if (s.mode == SRC_ADDITIVE) return 4.0 * G.ncell;  // K_M or the filtered grid
    return 4.0 * G.ncell;                             // dense point map
  }
  double sensor_bytes(int t) const {
    if (time_rev || t < record_start) return 0.0;
    const double nc = (double)G.ncell;
    const bool first = t == record_start;
    const double stat = first ? 8.0 : 16.0;
    double b = 0.0;
    if (full_mask) {
      if (record & KW_REC_P) b += 4.0 * nc;
      if (record & KW_REC_P_MAX) b += stat * nc;
      if (record & KW_REC_P_MIN) b += stat * nc;
      if (record & KW_REC_P_RMS) b += stat * nc;
    }
    if (record & KW_REC_P_MAX_ALL) b += stat * nc;
    if (record & KW_REC_P_MIN_ALL) b += stat * nc;
    return b;
  }
  double pass_bytes(PassKind k, const PassArgs& a) const {
    const int t = (int)(steps_done + 1);
    const double nc = (double)G.ncell, sp = 8.0 * (double)G.nspec, tab = 4.0 * (double)G.nspec;
    const int dim = G.dim;
    switch (k) {
      case PK_YFWD:
      case PK_YINV:
        return a.nbuf * 2.0 * sp;
      case PK_SPEC:
        if (a.mode == SM_GRAD) return sp + dim * sp + tab;
        if (a.mode == SM_DIV) return 2.0 * dim * sp + tab;
        if (a.mode == SM_ABS) return 4.0 * sp + 2.0 * tab;
        if (a.mode == SM_SHIFT) return 2.0 * sp;
        if (a.mode == SM_SPLIT) return 2.0 * dim * sp + dim * tab;
        return 2.0 * sp + tab;
      case PK_X_R2C:
        return a.nbuf * (4.0 * nc + sp);
      case PK_X_C2R:
        return a.nbuf * (sp + 4.0 * nc) * ((a.flags & XC_ACCUM) ? 1.5 : 1.0);
      case PK_X_GRAD: {
        double b = 0.0;
        const bool init = a.flags & XG_INIT;
        for (int d = 0; d < dim; ++d) {
          b += sp + fb(P.rinv[d]) + 4.0 * nc;
          if (!init) b += 4.0 * nc + sp + src_bytes(P.srcu[d], t);
        }
        return b;
      }
      case PK_X_DIV: {
        double b = dim * (sp + 8.0 * nc) + fb(P.rho0) + dim * src_bytes(P.srcp, t);
        if (P.eos == EOS_ABSORBING) return b + 2.0 * sp + 4.0 * nc;
        b += fb(P.c0) + sp + sensor_bytes(t);
        if (P.eos == EOS_STOKES) b += fb(P.tau);
        if (P.nonlinear) b += fb(P.bona);
        if (a.flags & X_STORE_P) b += 4.0 * nc;
        if (time_rev) b += 4.0 * nc;
        return b;
      }
      case PK_X_EOS: {
        double b = 2.0 * sp + 4.0 * nc + fb(P.c0) + fb(P.tau) + fb(P.eta) + sp + sensor_bytes(t);
        if (P.nonlinear) b += fb(P.bona) + fb(P.rho0);
        if (a.flags & X_STORE_P) b += 4.0 * nc;
        if (time_rev) b += 4.0 * nc;
        return b;
      }
    }
    return 0.0;
  }
  double ew_bytes(const EwArgs& a) const {
    const double nc = (double)G.ncell;
    switch (a.kind) {
      case EW_INIT_P0: return (2.0 + G.dim) * 4.0 * nc;
      case EW_TR_PROLOGUE: return (2.0 + G.dim) * 4.0 * nc;
      case EW_SRC_GRID: return 8.0 * nc;
      case EW_RECORD_POINTS: return 24.0 * (double)nsens;
      case EW_RECORD_GRID: return 24.0 * nc;
      default: return 0.0;
    }
  }

  // -- tiling -------------------------------------------------------------
  void plan_launches() {
    const size_t maxl = dev_max_lds();
    const int capP = max_pencils_env();
    // threads per pencil of the runtime engine, from the length its stages
    // run at: ceil(L/7) for lengths with a factor 7 (one radix-7 butterfly
    // per thread; every other stage then has at most its usual count)
    auto Gof = [](int N) {
      const int L = stage_length(N);
      return L % 7 == 0 ? (L + 6) / 7 : (L + KW_E - 1) / KW_E;
    };
    // x pass: <= 256 threads, LDS Nx twiddles + P * kNPX padded pencils; the
    // fixed-length x kernels run x_threads(Nx) threads per pencil
    {
      const int Gx = fixed_len(PK_X_GRAD) ? x_threads(G.Nx) : Gof(G.Nx);
      const int Lx = stage_length(G.Nx);
      const size_t pencil = sizeof(cf) * (size_t)kNPX * x_lds_stride(Lx);
      const size_t twb = sizeof(cf) * (size_t)Lx;
      int P = std::max(1, std::min(pencils_env("KW_PENCILS_X", tiles.x ? tiles.x : 256 / Gx),
                                   512 / Gx));
      P = std::min<int64_t>(P, (int64_t)((maxl - twb) / pencil));
      P = std::max(1, std::min(P, capP));
      this->P.Gx = Gx;
      this->P.Px = P;
      const int64_t npairs = (G.nrows + 1) / 2;
      lx.nb = (int)((npairs + P - 1) / P);
      lx.T = P * Gx;
      lx.lds = pencil * (size_t)P + twb;
      if (Gx > 512 || lx.lds > maxl) fail(KW_EUNSUP, "grid size Nx too large for the x pass");
    }
    // strided y pass (3D): <= 1024 threads, LDS Ny twiddles + P*(Ny+1)
    if (G.dim == 3) {
      const int Gy = Gof(G.Ny);
      const int Ly = stage_length(G.Ny);
      int P = std::max(1, std::min(pencils_env("KW_PENCILS_Y", tiles.y ? tiles.y : 16), 1024 / Gy));
      P = std::min<int64_t>(P, (int64_t)((maxl - sizeof(cf) * Ly) / (sizeof(cf) * (Ly + 1))));
      P = std::max(1, std::min(P, capP));
      this->P.Gy = Gy;
      this->P.Py = P;
      const int64_t nH = (int64_t)G.Wm * G.Nz;
      ly.nb = (int)((nH + P - 1) / P + (G.nyq ? (G.Nz + P - 1) / P : 0));
      ly.T = P * Gy;
      ly.lds = sizeof(cf) * ((size_t)P * (Ly + 1) + Ly);
      if (Gy > 1024 || ly.lds > maxl) fail(KW_EUNSUP, "grid size Ny too large for the y pass");
    }
    // last-axis pass: <= 1024 threads, LDS N twiddles + 2*P*N
    {
      const int Nl = G.dim == 3 ? G.Nz : G.Ny;
      const int Gl = Gof(Nl);
      const int Ll = stage_length(Nl);
      int P = std::max(1, std::min(pencils_env("KW_PENCILS_L", tiles.l ? tiles.l : 16), 1024 / Gl));
      P = std::min<int64_t>(P, (int64_t)((maxl - sizeof(cf) * Ll) / (2 * sizeof(cf) * Ll)));
      P = std::max(1, std::min(P, capP));
      this->P.Gl = Gl;
      this->P.Pl = P;
      const int64_t nH = G.dim == 3 ? (int64_t)G.Wm * G.Ny : (int64_t)G.Wm;
      const int64_t nQ = G.dim == 3 ? G.Ny : 1;
      ll.nb = (int)((nH + P - 1) / P + (G.nyq ? (nQ + P - 1) / P : 0));
      ll.T = P * Gl;
      ll.lds = sizeof(cf) * (2 * (size_t)P * Ll + Ll);
      if (Gl > 1024 || ll.lds > maxl) fail(KW_EUNSUP, "grid size too large for the last-axis pass");
    }
    lcell.nb = (int)((G.ncell + EW_T - 1) / EW_T);
    lsens.nb = (int)((std::max<int64_t>(nsens, 1) + EW_T - 1) / EW_T);
  }

  // FFT plan of an axis of length N: stages of the radices of N, or for N
  // with a prime factor above 7 Bluestein's chirp-z transform (kw_fft.h),
  // whose chirp and filter transform are computed here in double precision
  // and rounded once
  FftPlan make_plan(int N) {
    FftPlan pl{};
    std::vector<int> r;
    const bool direct = factorise(N, r);
    const int L = direct ? N : chirp_length(N);
    if (!direct && !factorise(L, r)) fail(KW_EUNSUP, "no FFT plan for length " + std::to_string(N));
    pl.N = N;
    pl.L = L;
    pl.nst = (int)r.size();
    for (size_t i = 0; i < r.size(); ++i) pl.radix[i] = r[i];
    std::vector<cf> tw(L);
    for (int j = 0; j < L; ++j) {
      const double a = 2.0 * kPi * (double)j / (double)L;
      tw[j] = cf{(float)std::cos(a), (float)-std::sin(a)};
    }
    pl.tw = upload(tw);
    if (!direct) {
      // w_n = exp(-i pi n^2 / N), with n^2 reduced modulo 2N
      std::vector<std::complex<double>> w(N);
      for (int n = 0; n < N; ++n) {
        const int64_t q = ((int64_t)n * n) % (2 * (int64_t)N);
        const double a = kPi * (double)q / (double)N;
        w[n] = std::complex<double>(std::cos(a), -std::sin(a));
      }
      // filter b_m = conj(w_|m|) for |m| < N, wrapped to length L (symmetric),
      // and its transform B_k = b_0 + 2 sum_m b_m cos(2 pi m k / L)
      std::vector<double> cosl(L);
      for (int j = 0; j < L; ++j) cosl[j] = std::cos(2.0 * kPi * (double)j / (double)L);
      std::vector<std::complex<double>> B(L);
      parallel_for(L, [&](int64_t k) {
        std::complex<double> acc = std::conj(w[0]);
        for (int m = 1; m < N; ++m)
          acc += 2.0 * cosl[(int64_t)m * k % L] * std::conj(w[m]);
        B[k] = acc;
      }, N);
      std::vector<cf> ch(N), bs(L);
      for (int n = 0; n < N; ++n) ch[n] = cf{(float)w[n].real(), (float)w[n].imag()};
      for (int k = 0; k < L; ++k) bs[k] = cf{(float)B[k].real(), (float)B[k].imag()};
      pl.chirp = upload(ch);
      pl.bspec = upload(bs);
    }
    return pl;
  }

  // Table of f(kx, ky, kz) (ifftshifted wavenumbers) over the half
  // spectrum, stored in the order the last-axis pass reads it: element n of
  // pencil p of tile t at (t * N + n) * Pl + p, the tiles of the main block
  // first, then those of the Nyquist plane, padding pencils zero.  A block
  // then reads whole contiguous lines of a table; in the natural layout its
  // Pl floats per element would be half of a 128-byte line whose other half
  // another block (on another L2) fetches again.
  template <class F>
  float* spectral_table(F f) {
    const Geometry g = G;
    const int Pn = P.Pl;
    const int N = g.dim == 3 ? g.Nz : g.Ny;
    const int64_t nH = g.dim == 3 ? (int64_t)g.Wm * g.Ny : (int64_t)g.Wm;
    const int64_t nQ = g.nyq ? (g.dim == 3 ? g.Ny : 1) : 0;
    const int64_t tH = (nH + Pn - 1) / Pn;
    const int64_t nt = ll.nb;
    std::vector<float> t((size_t)(nt * N * Pn), 0.f);
    auto pencil = [&](int64_t e) {
      const int64_t tile = e / Pn;
      const int p = (int)(e % Pn);
      int ix, iy;
      if (tile < tH) {
        const int64_t q = tile * Pn + p;
        if (q >= nH) return;
        ix = (int)(q % g.Wm);
        iy = (int)(q / g.Wm);
      } else {
        const int64_t q = (tile - tH) * Pn + p;
        if (q >= nQ) return;
        ix = g.Nx / 2;
        iy = (int)q;
      }
      for (int n = 0; n < N; ++n) {
        const double v = g.dim == 3 ? f(ix, iy, n) : f(ix, n, 0);
        t[(size_t)((tile * N + n) * Pn + p)] = (float)v;
      }
    };
    parallel_for(nt * Pn, pencil, N);
    return upload(t);
  }
};

// `arena` set: every device buffer comes from it and the simulation runs on
// the device's null stream (no stream is created)
std::unique_ptr<Sim> create_sim(const kw_problem* pr, void* stream = nullptr,
                                Arena* arena = nullptr);
PassArgs bufs(int nb, int b0 = 0, int b1 = 1, int b2 = 2);
void src_grid(Sim& S, int which, int t);
void record_velocity(Sim& S);
void record_directivity(Sim& S);
void step_once(Sim& S);
void step_n(Sim& S, int64_t nsteps);
void extend(Sim& S, int64_t Nt_total);
void change_dt(Sim& S, double dt_prev, double dt_new);
void init_velocity(Sim& S, const double* const u[3]);
// pencils per block for a problem's grid: the autotuner's measurement,
// cached per grid shape and equation of state (see kw_tiles in kwhip.h)
Tiles choose_tiles(const kw_problem& pr, void* stream, bool may_tune);
// while one lives, choose_tiles on this thread only reads the cache (the
// engine's admission threads: a measurement there would hold up admission)
struct TuneCacheOnly {
  bool prev;
  TuneCacheOnly();
  ~TuneCacheOnly();
};
int64_t recorded(const Sim& S);
void read_outputs(Sim& S, kw_outputs* o);

}  // namespace kw

// C ABI error handling shared by the wrappers
inline void kw_set_err(char* err, size_t errlen, const char* m) {
  if (err && errlen) std::snprintf(err, errlen, "%s", m);
}

template <class F>
int kw_guarded(char* err, size_t errlen, F f) {
  try {
    f();
    kw_set_err(err, errlen, "");
    return KW_OK;
  } catch (const kw::DeviceError& e) {
    kw_set_err(err, errlen, e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    kw_set_err(err, errlen, "host allocation failed");
    return KW_ENOMEM;
  } catch (const std::exception& e) {
    kw_set_err(err, errlen, e.what());
    return KW_EDEVICE;
  }
}
