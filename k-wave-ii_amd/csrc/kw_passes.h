// Kernel bodies of the step loop.  One step of kspaceFirstOrder{2D,3D}
// (reference +kwave/+legacy/kspaceFirstOrder3D.m:769-1208 and
// kspaceFirstOrder2D.m:712-1044) runs as the pass sequence
//
//   gradient   : [fft_x(p) from the previous step]  Y+  Z(grad)  Y-  X(grad)
//   divergence : Y+  Z(div)  Y-  X(div)
//   absorbing  : Y+  Z(abs)  Y-  X(eos)
//
// (in 2D the last-axis pass is along y and there are no Y+/Y- passes).
// The last-axis pass does forward FFT, spectral multiply and inverse FFT
// without leaving LDS; the x passes fuse the C2R and R2C transforms with
// every elementwise update of the reference (u, rho, sources, equation of
// state, sensors, time-reversal enforcement).
//
// Bodies are templated on a plan selector per axis (PlanRt for any length,
// PlanCt<N> for a length fixed at compile time) and the x passes on the
// number of dimensions, which is the number of field components they
// transform together.
#pragma once

#include "kw_prob.h"

namespace kw {

// ---------------------------------------------------------------------------
// half-spectrum access for the x passes ("two-for-one": two real rows are
// transformed as one complex pencil z = a + i b)
// ---------------------------------------------------------------------------

// element n (0 <= n < Nx) of the Hermitian extension of row `row`, with the
// imaginary parts of the DC and Nyquist bins dropped, which is what
// real(ifft(.)) does with them; branch-free so the loads batch
KW_BODY cf half_get(const ProbDev& P, const cf* S, uint32_t hrow, uint32_t qrow, int n) {
  // hrow = row * Wm (main block), qrow = offset of the row in the Nyquist plane
  const int N = P.Nx;
  const int h = N >> 1;
  const bool mirror = n > h;
  const bool isq = P.nyq && n == h;
  const uint32_t k = (uint32_t)(mirror ? N - n : n);
  const uint32_t idx = isq ? qrow : hrow + k;
  cf v = KW_LD32(S, idx);
  v.y = (n == 0 || isq) ? 0.f : (mirror ? -v.y : v.y);
  return v;
}

KW_BODY void half_put(const ProbDev& P, cf* S, uint32_t hrow, uint32_t qrow, int k, cf v) {
  const uint32_t idx = k < P.Wm ? hrow + (uint32_t)k : qrow;
  KW_ST32(S, idx, v);
}

// pencils per row pair held in LDS by the x passes: the field components
// are transformed one after another through the same pencil
constexpr int kNPX = 1;

KW_HOSTDEV int x_lds_stride(int Nx) { return Nx + (Nx >> 3) + 1; }

struct XTile {
  int g, G;
  bool active, has1;
  bool wave;  // the G threads of every pencil lie in one wavefront
  int64_t r0, r1;
  // 32-bit offsets of the two rows: first cell (row * Nx), half-spectrum row
  // (row * Wm), Nyquist entry (nspec_main + row); y and z of each row
  uint32_t c0, c1, h0, h1, q0, q1, y0, y1, z0, z1;
  LdsPencils lds;
  TwL tw;  // the x twiddles, staged at the start of the block's LDS
};

template <class XS, class Ctx>
KW_BODY XTile x_tile(Ctx& c, const ProbDev& P, int tile) {
  XTile T;
  constexpr int Gc = XS::fixed ? XS::G : 0;
  T.G = XS::fixed ? Gc : P.Gx;
  T.wave = XS::fixed ? pencils_in_wave(Gc) : pencils_in_wave(P.Gx);
  const int tid = c.tid();
  int p = tid / T.G;
  T.g = tid % T.G;
  const int64_t npairs = (P.nrows + 1) / 2;
  int64_t pair = (int64_t)tile * P.Px + p;
  T.active = p < P.Px && pair < npairs;
  // threads without a row pair read row pair 0 (their loads stay in bounds
  // and unconditional) and never store; they keep their own LDS pencil
  // (a launch shared by several problems sizes its LDS for T / G pencils)
  if (!T.active) pair = 0;
  T.r0 = 2 * pair;
  T.has1 = 2 * pair + 1 < P.nrows;
  T.r1 = T.has1 ? 2 * pair + 1 : T.r0;
  const uint32_t r0 = (uint32_t)T.r0, r1 = (uint32_t)T.r1;
  T.c0 = r0 * (uint32_t)P.Nx;
  T.c1 = r1 * (uint32_t)P.Nx;
  T.h0 = r0 * (uint32_t)P.Wm;
  T.h1 = r1 * (uint32_t)P.Wm;
  T.q0 = (uint32_t)P.nspec_main + r0;
  T.q1 = (uint32_t)P.nspec_main + r1;
  T.y0 = r0 % (uint32_t)P.Ny;
  T.z0 = r0 / (uint32_t)P.Ny;
  T.y1 = r1 % (uint32_t)P.Ny;
  T.z1 = r1 / (uint32_t)P.Ny;
  // LDS: the L twiddles of the x transform (staged by the whole block), then
  // the pencils (L elements each, padded)
  T.tw = stage_twiddles(c, P.fx, c.lds());
  const int stride = x_lds_stride(P.fx.L);
  T.lds = LdsPencils{c.lds() + P.fx.L + (int64_t)p * kNPX * stride, stride, 1, true};
  return T;
}

// LDS pencil j holds Z = fft(a + i b) in natural order; split into the half
// spectra of a (row r0) and b (row r1) and store them to S[j]
template <int NP, class Ctx>
KW_BODY void x_split_store(Ctx& c, const ProbDev& P, const XTile& T, cf* const* S) {
  const int N = P.Nx;
  const int K = N / 2 + 1;
  if (T.active) {
    for (int k = T.g; k < K; k += T.G) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const cf zk = T.lds.get(j, k);
        const cf zm = T.lds.get(j, k == 0 ? 0 : N - k);
        const cf A = cf{0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)};
        const cf B = cf{0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x)};
        // r1 == r0 for a single-row pair: the r0 store comes last and wins
        half_put(P, S[j], T.h1, T.q1, k, B);
        half_put(P, S[j], T.h0, T.q0, k, A);
      }
    }
  }
  c.sync();
}

// forward transform of NP complex pencils already in LDS, then split/store
template <class XS, int NP, class Ctx>
KW_BODY void x_fwd_from_lds(Ctx& c, const ProbDev& P, const XTile& T, cf* const* S) {
  auto src = [&](int j, int n) -> cf { return T.lds.get(j, n); };
  auto dst = [&](int j, int n, cf v, int) { T.lds.set(j, n, v); };
  XS::template fft_tw<NP, false>(c, P.fx, T.tw, T.g, T.G, T.active, src, dst, T.lds);
  x_split_store<NP>(c, P, T, S);
}

// inverse two-for-one load functor: Z_j[n] = A_j[n] + i B_j[n]; row r1 is
// r0 when the pair has a single row (the B part is then zeroed by select)
template <int NP>
struct HalfPairLoad {
  const ProbDev* P;
  const cf* S[NP];
  uint32_t h0, h1, q0, q1;
  bool has1;
  KW_BODY cf operator()(int j, int n) const {
    const cf a = half_get(*P, S[j], h0, q0, n);
    cf b = half_get(*P, S[j], h1, q1, n);
    b.x = has1 ? b.x : 0.f;
    b.y = has1 ? b.y : 0.f;
    return cf{a.x - b.y, a.y + b.x};
  }
};

// 1-D PML factors of axis d for the two rows at x = n: x varies along the
// row, y and z are constant along it
KW_BODY float pml_at(const float* v, const XTile& T, int d, int row, int n) {
  if (d == 0) return KW_LD32(v, (uint32_t)n);
  const uint32_t i = d == 1 ? (row == 0 ? T.y0 : T.y1) : (row == 0 ? T.z0 : T.z1);
  return KW_LD32(v, i);
}

// Time-source update of one cell, for a source form fixed at compile time:
// 0 additive single series (s(t) K_M), 1 additive several series (filtered
// grid), 2 dirichlet (map), 3 additive-no-correction (map).  `s0` is the
// signal value s(t) of form 0, read once per launch by src_value0.
template <int FORM>
KW_BODY float apply_src(const TimeSource& s, uint32_t cell, float v, int t, float s0) {
  const int t0 = t - 1;
  if constexpr (FORM == 0) {
    return v + s0 * KW_LD32(s.km, cell);
  } else if constexpr (FORM == 1) {
    return v + KW_LD32(s.grid, cell);
  } else {
    const int32_t i = KW_LD32(s.map, cell);
    const float sv = s.value(i < 0 ? 0 : i, t0);
    if constexpr (FORM == 2)
      return i >= 0 ? sv : v;
    else
      return i >= 0 ? v + sv : v;
  }
}

template <int FORM>
KW_BODY float src_value0(const TimeSource& s, int t) {
  if constexpr (FORM == 0)
    return s.value(0, t - 1);
  else
    return 0.f;
}

KW_BODY int src_form(const TimeSource& s) {
  if (s.mode == SRC_ADDITIVE) return s.km ? 0 : 1;
  return s.mode == SRC_DIRICHLET ? 2 : 3;
}

// run f(IntC<form>) for the source's form, or f(IntC<-1>) when it is off
template <class F>
KW_BODY void with_src(const TimeSource& s, bool on, F&& f) {
  if (!on) {
    f(IntC<-1>{});
    return;
  }
  switch (src_form(s)) {
    case 0: f(IntC<0>{}); break;
    case 1: f(IntC<1>{}); break;
    case 2: f(IntC<2>{}); break;
    default: f(IntC<3>{}); break;
  }
}

// Time source applied to this thread's values of the two rows (a: row r0,
// b: row r1) of an x pass, in a phase of its own after the update, so that
// the source's loads (up to three dependent loads per cell) are not batched
// together with the update's; nothing when the source is off at step t.
template <class XS>
KW_BODY void x_source(const ProbDev& P, const XTile& T, const TimeSource& src, int t, float* av,
                      float* bv) {
  with_src(src, src.len >= t, [&](auto FORM) {
    constexpr int F = decltype(FORM)::value;
    if constexpr (F >= 0) {
      KW_PHASE_FENCE();
      const float s0 = src_value0<F>(src, t);
      auto f = [&](int slot, int n) {
        av[slot] = apply_src<F>(src, T.c0 + n, av[slot], t, s0);
        bv[slot] = apply_src<F>(src, T.c1 + n, bv[slot], t, s0);
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, f);
    }
  });
}

template <class Ctx>
KW_BODY const ProbDev& batch_prob(Ctx& c, const Batch& B, int& tile) {
  const int pi = find_prob<Ctx>(B, c.bid());
  tile = c.bid() - KW_LD(B.tile_start + pi);
  return B.probs[pi];
}

// ---------------------------------------------------------------------------
// strided pass along y (3D only), forward or inverse, in place
// ---------------------------------------------------------------------------
// the first pass of a step advances the problem's step number (no block of
// that pass reads it; every later kernel of the step does)
KW_BODY void advance_step(const ProbDev& P, const PassArgs& A, int tile, int tid) {
  if (A.advance && tile == 0 && tid == 0) KW_ST(P.t_index, KW_LD(P.t_index) + 1);
}

template <bool INV, class YS, class Ctx>
KW_BODY void body_ypass(Ctx& c, const Batch& B, const PassArgs& A) {
  // the launch's tiles, once per buffer when the buffers are split over
  // blocks (ysplit), else every block loops over the buffers
  int vb = c.bid(), b0 = 0, b1 = A.nbuf;
  if (A.ysplit > 0) {
    b0 = vb / A.ysplit;
    b1 = b0 + 1;
    vb -= b0 * A.ysplit;
  }
  const int pi = find_prob<Ctx>(B, vb);
  const int tile = vb - KW_LD(B.tile_start + pi);
  const ProbDev& P = B.probs[pi];
  if (b0 == 0) advance_step(P, A, tile, c.tid());
  const int N = P.Ny, G = P.Gy, Pn = P.Py;
  // LDS: the L twiddles, then the pencils (L elements each: the length the
  // stages run at)
  const TwL tw = stage_twiddles(c, P.fy, c.lds());
  const int L = P.fy.L;
  cf* const lp = c.lds() + L;
  const int64_t nH = (int64_t)P.Wm * P.Nz;
  const int64_t tH = (nH + Pn - 1) / Pn;
  const int tid = c.tid();
  int p, g;
  uint32_t base, stride;
  bool active;
  LdsPencils lds;
  if (tile < tH) {
    p = tid % Pn;
    g = tid / Pn;
    int64_t q = (int64_t)tile * Pn + p;
    active = q < nH && g < G;
    if (g >= G) g = G - 1;  // a block larger than this problem's tiling
    if (!active) q = 0;     // read a valid pencil, never store
    base = (uint32_t)((q / P.Wm) * ((int64_t)P.Wm * P.Ny) + (q % P.Wm));
    stride = (uint32_t)P.Wm;
    lds = LdsPencils{lp + p, 0, Pn, false};
  } else {
    p = tid / G;
    g = tid % G;
    int64_t q = (tile - tH) * Pn + p;
    active = p < Pn && q < P.Nz;
    if (!active) {
      p = 0;
      q = 0;
    }
    base = (uint32_t)(P.nspec_main + q * P.Ny);
    stride = 1;
    lds = LdsPencils{lp + (int64_t)p * (L + 1), 0, 1, false};
  }
  for (int b = b0; b < b1; ++b) {
    cf* S = P.S[A.buf[b]];
    auto ld = [&](int, int n) -> cf { return KW_LD32(S, base + (uint32_t)n * stride); };
    auto st = [&](int, int n, cf v, int) { KW_ST32(S, base + (uint32_t)n * stride, v); };
    YS::template fft_tw<1, INV>(c, P.fy, tw, g, G, active, ld, st, lds);
  }
}

// ---------------------------------------------------------------------------
// last-axis pass (z in 3D, y in 2D): forward FFT, spectral operator, inverse
// FFT, all in LDS.  The forward transform ends in pencil la; the inverse
// transforms read la.  Operators (kspaceFirstOrder3D.m:629-661, 792-803,
// 907-909, 977-978; createAbsorptionVariables.m:29-84).
// ---------------------------------------------------------------------------
// Sensor-record modes of the last-axis pass (see body_spec).
template <class LS, class Ctx, class Addr, class TAddr, class ToA>
KW_BODY void spec_aux(Ctx& c, const ProbDev& P, const PassArgs& A, const FftPlan& pl,
                      const TwL& tw, int N, int G,
                      int last, int g, bool active, int tile, int64_t tH, int kx, int ky,
                      const LdsPencils& la, const LdsPencils& lb, const LdsPencils& lw, Addr& addr,
                      TAddr& taddr, ToA& to_a) {
  if (A.mode == SM_SHIFT) {
    // S <- ushift_d .* S for the velocity on the non-staggered grid
    // (kspaceFirstOrder_extractSensorData.m:47-56): a per-pencil constant for
    // x and y (3-D), varying along the pencil for the last axis; with a
    // table, table .* ushift_d .* S (the axisymmetric u_r / r term,
    // kspaceFirstOrderAS.m:954-969: kappa and the radial shift in one pass),
    // added to S[buf[1]] when nbuf is 2
    const int d = A.flags;
    cf* S = P.S[A.buf[0]];
    cf* D = A.nbuf == 2 ? P.S[A.buf[1]] : S;
    const cf* sh = P.ushift[d];
    const float* tab = A.table;
    auto ld = [&](int, int n) -> cf { return KW_LD32(S, addr(n)); };
    LS::template fft_tw<1, false>(c, pl, tw, g, G, active, ld, to_a, la);
    const bool along = d == last;
    const cf sc = (along || !active) ? cf{0.f, 0.f} : KW_LD(sh + (d == 0 ? kx : ky));
    auto src = [&](int, int n) -> cf {
      const cf v = cmul(along ? KW_LD32(sh, (uint32_t)n) : sc, la.get(0, n));
      return tab ? cscale(v, KW_LD32(tab, taddr(n))) : v;
    };
    auto st = [&](int, int n, cf v, int) {
      KW_ST32(D, addr(n), D == S ? v : cadd(KW_LD32(D, addr(n)), v));
    };
    LS::template fft_tw<1, true>(c, pl, tw, g, G, active, src, st, lw);
  } else if (A.mode == SM_SPLIT) {
    // compressional part of the velocity (kspaceFirstOrder_extractSensorData.m:
    // 124-292): D = sum_b khat_b shift_b U_b is accumulated in lb (each thread
    // the elements it produces, the same for every component), then
    // component a is ifft(khat_a D).  real(ifftn(.)) keeps nothing of the
    // bins where k_a is the Nyquist wavenumber (the product there is
    // anti-Hermitian), so khat_a is taken as 0 there and every transform
    // stays Hermitian.
    const int nc = P.dim;
    const bool nyq_n = (N & 1) == 0;
    for (int b = 0; b < nc; ++b) {
      const cf* Sb = P.S[A.buf[b]];
      auto ld = [&](int, int n) -> cf { return KW_LD32(Sb, addr(n)); };
      const cf* sh = P.sshift[b];
      const bool along = b == last;
      const cf sc = (along || !active) ? cf{0.f, 0.f} : KW_LD(sh + (b == 0 ? kx : ky));
      const float* kh = P.khat[b];
      auto acc = [&](int, int n, cf v, int) {
        const cf s = along ? KW_LD32(sh, (uint32_t)n) : sc;
        const cf t = cscale(cmul(s, v), KW_LD32(kh, taddr(n)));
        lb.set(0, n, b == 0 ? t : cadd(lb.get(0, n), t));
      };
      LS::template fft_tw<1, false>(c, pl, tw, g, G, active, ld, acc, la);
    }
    for (int a = 0; a < nc; ++a) {
      cf* Sa = P.S[A.buf[a]];
      const float* kh = P.khat[a];
      // is this pencil at the Nyquist wavenumber of axis a (for a != last)
      const bool pen_nyq = a == 0 ? (P.nyq && kx == P.Nx / 2 && tile >= tH)
                                  : (a != last && (P.Ny & 1) == 0 && ky == P.Ny / 2);
      auto src = [&](int, int n) -> cf {
        const bool zero = a == last ? (nyq_n && n == N / 2) : pen_nyq;
        const float k = zero ? 0.f : KW_LD32(kh, taddr(n));
        return cscale(lb.get(0, n), k);
      };
      auto st = [&](int, int n, cf v, int) { KW_ST32(Sa, addr(n), v); };
      LS::template fft_tw<1, true>(c, pl, tw, g, G, active, src, st, la);
    }
  }
}

KW_HOSTDEV bool spec_is_aux(const PassArgs& A) { return A.mode == SM_SHIFT || A.mode == SM_SPLIT; }

// AUX selects the sensor-record modes (SM_SHIFT, SM_SPLIT), built as a
// kernel of their own so that the time step's kernel carries none of their
// code.
template <class LS, bool AUX, class Ctx>
KW_BODY void body_spec(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  advance_step(P, A, tile, c.tid());
  const bool d3 = P.dim == 3;
  const FftPlan& pl = d3 ? P.fz : P.fy;
  const int N = pl.N, G = P.Gl, Pn = P.Pl;
  const int last = d3 ? 2 : 1;
  const int64_t nH = d3 ? (int64_t)P.Wm * P.Ny : (int64_t)P.Wm;
  const int64_t tH = (nH + Pn - 1) / Pn;
  const int tid = c.tid();
  const int p = tid % Pn;
  const bool gin = tid / Pn < G;           // false: block larger than this problem's tiling
  const int g = gin ? tid / Pn : G - 1;
  int64_t q;
  uint32_t base, stride;
  int kx, ky;  // ky = -1: ky is the pencil coordinate (2D)
  bool active;
  if (tile < tH) {
    q = (int64_t)tile * Pn + p;
    active = q < nH && gin;
    base = (uint32_t)q;
    stride = (uint32_t)(d3 ? (int64_t)P.Wm * P.Ny : (int64_t)P.Wm);
    kx = (int)(q % P.Wm);
    ky = d3 ? (int)(q / P.Wm) : -1;
  } else {
    q = (tile - tH) * Pn + p;
    active = q < (d3 ? P.Ny : 1) && gin;
    base = (uint32_t)(P.nspec_main + q);
    stride = (uint32_t)(d3 ? P.Ny : 1);
    kx = P.Nx / 2;
    ky = d3 ? (int)q : -1;
  }
  if (!active) {
    // read pencil 0 of the main block, never store
    q = 0;
    base = 0;
    stride = (uint32_t)(d3 ? (int64_t)P.Wm * P.Ny : (int64_t)P.Wm);
    kx = 0;
    ky = d3 ? 0 : -1;
  }
  // LDS: the L twiddles, then pencil buffers la and lb (L elements each)
  const TwL tw = stage_twiddles(c, pl, c.lds());
  cf* const lp = c.lds() + pl.L;
  LdsPencils la{lp + p, 0, Pn, false};
  LdsPencils lb{lp + (int64_t)Pn * pl.L + p, 0, Pn, false};
  // workspace of the inverse transforms that read la: la itself for a
  // fixed-length plan (every thread of a block then has a pencil or a
  // pencil of its own), lb for runtime plans, whose launches may be shared
  // by problems of different tilings (a thread repeating another's work
  // would read la while its twin overwrites it)
  const LdsPencils& lw = LS::fixed ? la : lb;
  auto addr = [&](int n) -> uint32_t { return base + (uint32_t)n * stride; };
  // operator tables are stored in the order of this pass: element n of
  // pencil p of tile `tile` at (tile N + n) Pn + p (padded to whole tiles,
  // so threads without a pencil read in bounds too)
  const uint32_t trow = (uint32_t)tile * (uint32_t)N;
  auto taddr = [&](int n) -> uint32_t {
    return (trow + (uint32_t)n) * (uint32_t)Pn + (uint32_t)p;
  };
  auto to_a = [&](int, int n, cf v, int) { la.set(0, n, v); };
  // kappa is applied as the forward transform ends, so the gradient's two
  // inverse transforms share one read of it (the kernel is at the 128-VGPR
  // limit of four waves per SIMD, so kappa is not kept in registers)
  auto to_ak = [&](int, int n, cf v, int) {
    la.set(0, n, cscale(v, KW_LD32(P.kappa, taddr(n))));
  };

  if constexpr (AUX) {
    spec_aux<LS>(c, P, A, pl, tw, N, G, last, g, active, tile, tH, kx, ky, la, lb, lw, addr, taddr,
                 to_a);
  } else if (A.mode == SM_GRAD) {
    cf* S0 = P.S[0];
    auto ld = [&](int, int n) -> cf { return KW_LD32(S0, addr(n)); };
    LS::template fft_tw<1, false>(c, pl, tw, g, G, active, ld, to_ak, la);
    // C = ifft(kappa .* S); x (and in 3D y) derivatives are per-pencil
    // constants times C
    const cf dx = active ? KW_LD(P.dpos[0] + kx) : cf{0.f, 0.f};
    const cf dy = (active && d3) ? KW_LD(P.dpos[1] + ky) : cf{0.f, 0.f};
    auto src1 = [&](int, int n) -> cf { return la.get(0, n); };
    auto st1 = [&](int, int n, cf v, int) {
      KW_ST32(P.S[0], addr(n), cmul(dx, v));
      if (d3) KW_ST32(P.S[1], addr(n), cmul(dy, v));
    };
    LS::template fft_tw<1, true>(c, pl, tw, g, G, active, src1, st1, lb);
    // derivative along the pencil axis
    const cf* dl = P.dpos[last];
    auto src2 = [&](int, int n) -> cf { return cmul(KW_LD32(dl, (uint32_t)n), la.get(0, n)); };
    cf* So = P.S[last];
    auto st2 = [&](int, int n, cf v, int) { KW_ST32(So, addr(n), v); };
    LS::template fft_tw<1, true>(c, pl, tw, g, G, active, src2, st2, lb);
  } else if (A.mode == SM_DIV) {
    for (int d = 0; d < P.dim; ++d) {
      cf* S = P.S[d];
      auto ld = [&](int, int n) -> cf { return KW_LD32(S, addr(n)); };
      LS::template fft_tw<1, false>(c, pl, tw, g, G, active, ld, to_ak, la);
      const cf dc = (d == last || !active) ? cf{0.f, 0.f}
                                           : KW_LD(P.dneg[d] + (d == 0 ? kx : ky));
      const cf* dl = P.dneg[d];
      const bool along = d == last;
      auto src = [&](int, int n) -> cf {
        return cmul(along ? KW_LD32(dl, (uint32_t)n) : dc, la.get(0, n));
      };
      auto st = [&](int, int n, cf v, int) { KW_ST32(S, addr(n), v); };
      LS::template fft_tw<1, true>(c, pl, tw, g, G, active, src, st, lw);
    }
  } else {
    // SM_ABS: S0 <- nabla1 .* S0, S1 <- nabla2 .* S1;  SM_FILTER: one table
    const int nb = A.mode == SM_ABS ? 2 : 1;
    for (int j = 0; j < nb; ++j) {
      cf* S = P.S[A.mode == SM_ABS ? j : A.buf[0]];
      const float* tab = A.mode == SM_ABS ? (j == 0 ? P.nabla1 : P.nabla2) : A.table;
      auto ld = [&](int, int n) -> cf { return KW_LD32(S, addr(n)); };
      LS::template fft_tw<1, false>(c, pl, tw, g, G, active, ld, to_a, la);
      auto src = [&](int, int n) -> cf { return cscale(la.get(0, n), KW_LD32(tab, taddr(n))); };
      auto st = [&](int, int n, cf v, int) { KW_ST32(S, addr(n), v); };
      LS::template fft_tw<1, true>(c, pl, tw, g, G, active, src, st, lw);
    }
  }
}

// ---------------------------------------------------------------------------
// x passes
// ---------------------------------------------------------------------------

// real grid(s) -> half spectrum (PassArgs.real[b] -> S[buf[b]])
template <class XS, class Ctx>
KW_BODY void body_x_r2c(Ctx& cb, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(cb, B, tile);
  const XTile T = x_tile<XS>(cb, P, tile);
  PencilCtx<Ctx> c{cb, T.wave};
  // A.table: a factor per y row (the axisymmetric inverse radius), applied
  // as the rows are read
  const float* rs = A.table;
  const float s0 = rs ? KW_LD(rs + (T.c0 / (uint32_t)P.Nx) % (uint32_t)P.Ny) : 1.f;
  const float s1 = rs ? KW_LD(rs + (T.c1 / (uint32_t)P.Nx) % (uint32_t)P.Ny) : 1.f;
  for (int b = 0; b < A.nbuf; ++b) {
    const float* f = A.real[b] ? A.real[b] : P.p;
    auto src = [&](int, int n) -> cf {
      const float a = KW_LD32(f, T.c0 + n);
      const float bb = KW_LD32(f, T.c1 + n);
      return cf{a * s0, T.has1 ? bb * s1 : 0.f};
    };
    auto dst = [&](int j, int n, cf v, int) { T.lds.set(j, n, v); };
    XS::template fft_tw<1, false>(c, P.fx, T.tw, T.g, T.G, T.active, src, dst, T.lds);
    cf* S[1] = {P.S[A.buf[b]]};
    x_split_store<1>(c, P, T, S);
  }
}

// half spectrum -> real grid (S[buf[b]] -> PassArgs.real[b]); XC_ACCUM adds
template <class XS, class Ctx>
KW_BODY void body_x_c2r(Ctx& cb, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(cb, B, tile);
  const XTile T = x_tile<XS>(cb, P, tile);
  PencilCtx<Ctx> c{cb, T.wave};
  const bool acc = A.flags & XC_ACCUM;
  for (int b = 0; b < A.nbuf; ++b) {
    float* f = A.real[b];
    HalfPairLoad<1> src{&P, {P.S[A.buf[b]]}, T.h0, T.h1, T.q0, T.q1, T.has1};
    auto dst = [&](int, int n, cf v, int) {
      const uint32_t i0 = T.c0 + n, i1 = T.c1 + n;
      const float o0 = acc ? KW_LD32(f, i0) : 0.f;
      const float o1 = acc ? KW_LD32(f, i1) : 0.f;
      KW_ST32(f, i1, o1 + v.y);
      KW_ST32(f, i0, o0 + v.x);
    };
    XS::template fft_tw<1, true>(c, P.fx, T.tw, T.g, T.G, T.active, src, dst, T.lds);
  }
}

// inverse x transform of the half spectrum S into this thread's last-stage
// outputs gv[slot] (real parts: row r0, imaginary parts: row r1)
template <class XS, class PC>
KW_BODY void x_inverse(PC& c, const ProbDev& P, const XTile& T, const cf* S, cf* gv) {
  HalfPairLoad<1> src{&P, {S}, T.h0, T.h1, T.q0, T.q1, T.has1};
  auto dst = [&](int, int, cf v, int slot) { gv[slot] = v; };
  KW_PHASE_FENCE();
  XS::template fft_tw<1, true>(c, P.fx, T.tw, T.g, T.G, T.active, src, dst, T.lds);
  KW_PHASE_FENCE();
}

// this thread's values (a: row r0, b: row r1) to LDS pencil 0, then the
// forward x transform of the pencil into the half spectrum S
template <class XS, class PC>
KW_BODY void x_forward(PC& c, const ProbDev& P, const XTile& T, const float* av, const float* bv,
                       cf* S) {
  KW_PHASE_FENCE();
  if (T.active) {
    auto put = [&](int slot, int n) { T.lds.set(0, n, cf{av[slot], T.has1 ? bv[slot] : 0.f}); };
    XS::last_slots(P.fx, T.g, T.G, T.active, put);
  }
  c.sync();
  KW_PHASE_FENCE();
  cf* Sa[1] = {S};
  x_fwd_from_lds<XS, 1>(c, P, T, Sa);
}

// end of the gradient: u_d = pml_sg_d (pml_sg_d u_d - dt rho0_sg_d^-1 g_d)
// (kspaceFirstOrder3D.m:792-803), velocity sources (:806-886), then
// S[d] = fft_x(u_d) for the divergence.  XG_INIT: p0 initial condition
// u_d = dt rho0_sg_d^-1 g_d / 2 (:1023-1037), no fft.
template <class XS, int DIM, class Ctx>
KW_BODY void body_x_grad(Ctx& cb, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(cb, B, tile);
  const XTile T = x_tile<XS>(cb, P, tile);
  PencilCtx<Ctx> c{cb, T.wave};
  const int t = KW_LD(P.t_index);
  const bool init = A.flags & XG_INIT;
  const float dt = P.dt;
  // one component at a time: inverse transform, update, forward transform
#pragma unroll 1
  for (int d = 0; d < DIM; ++d) {
    cf gv[kMaxSlots];
    x_inverse<XS>(c, P, T, P.S[d], gv);
    float* u = P.u[d];
    const Field ri = P.rinv[d];
    const float* pv = P.pml_sg[d];
    float av[kMaxSlots], bv[kMaxSlots];
    auto run = [&](auto INIT) {
      auto upd = [&](int slot, int n) {
        const uint32_t c0 = T.c0 + n, c1 = T.c1 + n;
        float a, b;
        if constexpr (INIT.value) {
          a = dt * ri.at(c0) * gv[slot].x / 2.f;
          b = dt * ri.at(c1) * gv[slot].y / 2.f;
        } else {
          const float p0 = pml_at(pv, T, d, 0, n), p1 = pml_at(pv, T, d, 1, n);
          a = p0 * (p0 * KW_LD32(u, c0) - dt * ri.at(c0) * gv[slot].x);
          b = p1 * (p1 * KW_LD32(u, c1) - dt * ri.at(c1) * gv[slot].y);
        }
        av[slot] = a;
        bv[slot] = b;
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, upd);
    };
    if (init) {
      run(BoolC<true>{});
    } else {
      run(BoolC<false>{});
      // velocity source (kspaceFirstOrder3D.m:806-886) as a phase of its own
      x_source<XS>(P, T, P.srcu[d], t, av, bv);
    }
    KW_PHASE_FENCE();
    if (T.active) {
      auto put = [&](int slot, int n) {
        KW_ST32(u, T.c1 + n, bv[slot]);
        KW_ST32(u, T.c0 + n, av[slot]);
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, put);
    }
    if (!init) x_forward<XS>(c, P, T, av, bv, P.S[d]);
  }
}

// equation of state for one cell (kspaceFirstOrder3D.m:965-1019) for an
// equation of state and nonlinearity fixed at compile time; Aterm and Bterm
// are the transformed-back absorbing terms, sdu the velocity divergence sum
// (stokes)
template <int EOS, bool NL>
KW_BODY float eos_p(const ProbDev& P, uint32_t cell, float rs, float sdu, float Aterm,
                    float Bterm) {
  const float c0 = P.c0.at(cell);
  float inner;
  if constexpr (EOS == EOS_ABSORBING) {
    inner = (rs + P.tau.at(cell) * Aterm) - P.eta.at(cell) * Bterm;
  } else if constexpr (EOS == EOS_STOKES) {
    inner = rs + P.tau.at(cell) * P.rho0.at(cell) * sdu;
  } else {
    inner = rs;
  }
  if constexpr (NL) inner = inner + P.bona.at(cell) * (rs * rs) / (2.f * P.rho0.at(cell));
  return (c0 * c0) * inner;
}

// run f(IntC<eos>, BoolC<nonlinear>) with both fixed at compile time
template <class F>
KW_BODY void with_eos(const ProbDev& P, F&& f) {
  with_flag(P.nonlinear != 0, [&](auto NL) {
    if (P.eos == EOS_ABSORBING)
      f(IntC<EOS_ABSORBING>{}, NL);
    else if (P.eos == EOS_STOKES)
      f(IntC<EOS_STOKES>{}, NL);
    else
      f(IntC<EOS_LOSSLESS>{}, NL);
  });
}

// Density component set on the sensor cells by time reversal: p / (dim c0^2)
// (kspaceFirstOrder3D.m:772-788, kspaceFirstOrder2D.m:715-729,
// kspaceFirstOrder1D.m:597-611)
KW_BODY float tr_rho(const ProbDev& P, float p, float c2) {
  return P.dim == 3 ? p / (3.f * c2) : P.dim == 2 ? (0.5f * p) / c2 : p / c2;
}

// New p of the cells of this thread's slots (pv[slot] = rows r0, r1): time
// reversal, sensors and the optional store, each as one batched phase over
// the slots; finally the values go to LDS pencil 0 for fft_x(p).
template <class XS, class Ctx>
KW_BODY void x_finish_p(Ctx& c, const ProbDev& P, const XTile& T, cf* pv, int t, bool store) {
  const int N = P.Nx;
  {
    // non-finite pressure (NaN fails every comparison): the first step it
    // appears at is flagged for the engine's per-request failure check
    bool fin = true;
    auto f = [&](int slot, int) {
      fin = fin && (__builtin_fabsf(pv[slot].x) <= 3.402823466e38f) &&
            (__builtin_fabsf(pv[slot].y) <= 3.402823466e38f);
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, f);
    if (!fin && T.active && KW_LD(P.bad) == 0) KW_ST(P.bad, (int32_t)t);
  }
  if (P.time_rev) {
    // cells on the sensor take the boundary data of step t+1 of the reversed
    // series (original column Nt - t) and rho_d = p / (dim c0^2)
    int32_t iv[kMaxSlots][2];
    float rv[kMaxSlots][2];
    auto f = [&](int slot, int n) {
      const uint32_t cells[2] = {T.c0 + n, T.c1 + n};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int32_t i = KW_LD32(P.tr_map, cells[h]);
        const float p =
            KW_LD(P.tr_data + (int64_t)(i < 0 ? 0 : i) + (int64_t)(P.Nt - t - 1) * P.nsens);
        const float c0 = P.c0.at(cells[h]);
        const float c2 = c0 * c0;
        rv[slot][h] = tr_rho(P, p, c2);
        iv[slot][h] = i;
        float& pp = h == 0 ? pv[slot].x : pv[slot].y;
        pp = i >= 0 ? p : pp;
      }
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, f);
    if (T.active) {
      auto put = [&](int slot, int n) {
        for (int h = 1; h >= 0; --h) {
          if (iv[slot][h] < 0) continue;
          const uint32_t cell = (h == 0 ? T.c0 : T.c1) + n;
          for (int d = 0; d < P.dim; ++d) KW_ST32(P.rho[d], cell, rv[slot][h]);
        }
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, put);
    }
  }
  const bool rec = !P.time_rev && t >= P.record_start;
  const int fi = t - P.record_start + 1;
  const uint32_t flags = P.record;
  auto stat = [&](double* a, auto KIND) {
    // KIND 0: max, 1: min, 2: sum of squares
    constexpr int K = decltype(KIND)::value;
    KW_PHASE_FENCE();
    double n0v[kMaxSlots], n1v[kMaxSlots];
    auto f = [&](int slot, int n) {
      const double v0 = (double)pv[slot].x, v1 = (double)pv[slot].y;
      const double o0 = KW_LD32(a, T.c0 + n), o1 = KW_LD32(a, T.c1 + n);
      if constexpr (K == 2) {
        n0v[slot] = (fi == 1 ? 0.0 : o0) + v0 * v0;
        n1v[slot] = (fi == 1 ? 0.0 : o1) + v1 * v1;
      } else if constexpr (K == 0) {
        n0v[slot] = fi == 1 ? v0 : (v0 > o0 ? v0 : o0);
        n1v[slot] = fi == 1 ? v1 : (v1 > o1 ? v1 : o1);
      } else {
        n0v[slot] = fi == 1 ? v0 : (v0 < o0 ? v0 : o0);
        n1v[slot] = fi == 1 ? v1 : (v1 < o1 ? v1 : o1);
      }
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, f);
    if (T.active) {
      auto put = [&](int slot, int n) {
        KW_ST32(a, T.c1 + n, n1v[slot]);
        KW_ST32(a, T.c0 + n, n0v[slot]);
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, put);
    }
  };
  if (rec) {
    if (P.full_mask) {
      // sensor mask == whole grid: point statistics fused here
      if ((flags & REC_P) && T.active) {
        auto f = [&](int slot, int n) {
          const int64_t off = (int64_t)(fi - 1) * P.ncell;
          KW_ST(P.p_ts + off + (int64_t)T.c1 + n, pv[slot].y);
          KW_ST(P.p_ts + off + (int64_t)T.c0 + n, pv[slot].x);
        };
        XS::last_slots(P.fx, T.g, T.G, T.active, f);
      }
      if (flags & REC_P_MAX) stat(P.p_max, IntC<0>{});
      if (flags & REC_P_MIN) stat(P.p_min, IntC<1>{});
      if (flags & REC_P_RMS) stat(P.p_rms2, IntC<2>{});
    }
    if (flags & REC_P_MAX_ALL) stat(P.p_max_all, IntC<0>{});
    if (flags & REC_P_MIN_ALL) stat(P.p_min_all, IntC<1>{});
  }
  if (T.active) {
    auto fin = [&](int slot, int n) {
      if (store) {
        KW_ST32(P.p, T.c1 + n, pv[slot].y);
        KW_ST32(P.p, T.c0 + n, pv[slot].x);
      }
      T.lds.set(0, n, cf{pv[slot].x, T.has1 ? pv[slot].y : 0.f});
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, fin);
  }
  c.sync();
}

// end of the divergence: rho_d update (:912-927), pressure source
// (:930-962); then either the lossless/stokes EOS with sensors and fft_x(p)
// into S[0], or (absorbing) S[0] = fft_x(rho0 * sum du), S[1] = fft_x(sum rho)
// and rsum = sum rho
template <class XS, int DIM, class Ctx>
KW_BODY void body_x_div(Ctx& cb, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(cb, B, tile);
  const XTile T = x_tile<XS>(cb, P, tile);
  PencilCtx<Ctx> c{cb, T.wave};
  const int t = KW_LD(P.t_index);
  const float dt = P.dt;
  const bool psrc = P.srcp.len >= t;
  const bool store_p = (A.flags & X_STORE_P) || t >= P.p_store_from || t == P.p_store_at;

  // sums over the components: velocity divergence (absorbing / stokes EOS)
  // and density; mass conservation coefficient k = rho0 (linear) or
  // 2 sum(rho) + rho0 (nonlinear, from the densities before this update),
  // kept in registers across the components
  float du0[kMaxSlots], du1[kMaxSlots], rs0[kMaxSlots], rs1[kMaxSlots];
  float k0[kMaxSlots], k1[kMaxSlots];
  // a single-series additive pressure source adds s(t) K_M to every density
  // component: K_M is read once here and kept in registers for all of them
  float sk0[kMaxSlots], sk1[kMaxSlots];
  const bool km_once = psrc && src_form(P.srcp) == 0;
  if (km_once) {
    KW_PHASE_FENCE();
    const float s0 = src_value0<0>(P.srcp, t);
    auto f = [&](int slot, int n) {
      sk0[slot] = s0 * KW_LD32(P.srcp.km, T.c0 + n);
      sk1[slot] = s0 * KW_LD32(P.srcp.km, T.c1 + n);
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, f);
    KW_PHASE_FENCE();
  }
  const bool nonlin = P.nonlinear != 0;
  with_flag(nonlin, [&](auto NL) {
    constexpr bool nl = decltype(NL)::value;
    auto f = [&](int slot, int n) {
      const uint32_t c0 = T.c0 + n, c1 = T.c1 + n;
      du0[slot] = du1[slot] = rs0[slot] = rs1[slot] = 0.f;
      k0[slot] = P.rho0.at(c0);
      k1[slot] = P.rho0.at(c1);
      if constexpr (nl) {
        float s0 = KW_LD32(P.rho[0], c0) + KW_LD32(P.rho[1], c0);
        float s1 = KW_LD32(P.rho[0], c1) + KW_LD32(P.rho[1], c1);
        if (DIM == 3) {
          s0 = s0 + KW_LD32(P.rho[DIM - 1], c0);
          s1 = s1 + KW_LD32(P.rho[DIM - 1], c1);
        }
        k0[slot] = 2.f * s0 + k0[slot];
        k1[slot] = 2.f * s1 + k1[slot];
      }
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, f);
#pragma unroll 1
    for (int d = 0; d < DIM; ++d) {
      cf gv[kMaxSlots];
      x_inverse<XS>(c, P, T, P.S[d], gv);
      float* rho = P.rho[d];
      const float* pv = P.pml[d];
      float av[kMaxSlots], bv[kMaxSlots];
      auto f = [&](int slot, int n) {
        const uint32_t c0 = T.c0 + n, c1 = T.c1 + n;
        const float p0 = pml_at(pv, T, d, 0, n), p1 = pml_at(pv, T, d, 1, n);
        const cf g = gv[slot];
        av[slot] = p0 * (p0 * KW_LD32(rho, c0) - dt * k0[slot] * g.x);
        bv[slot] = p1 * (p1 * KW_LD32(rho, c1) - dt * k1[slot] * g.y);
        du0[slot] = du0[slot] + g.x;
        du1[slot] = du1[slot] + g.y;
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, f);
      // pressure source (kspaceFirstOrder3D.m:930-962) as a phase of its own;
      // a 1-D problem has its density in rho_x alone
      // (kspaceFirstOrder1D.m:716-743)
      if (km_once && d < P.dim) {
        auto f = [&](int slot, int) {
          av[slot] = av[slot] + sk0[slot];
          bv[slot] = bv[slot] + sk1[slot];
        };
        XS::last_slots(P.fx, T.g, T.G, T.active, f);
      } else if (psrc && d < P.dim) {
        x_source<XS>(P, T, P.srcp, t, av, bv);
      }
      auto acc = [&](int slot, int) {
        rs0[slot] = rs0[slot] + av[slot];
        rs1[slot] = rs1[slot] + bv[slot];
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, acc);
      KW_PHASE_FENCE();
      if (T.active) {
        auto put = [&](int slot, int n) {
          KW_ST32(rho, T.c1 + n, bv[slot]);
          KW_ST32(rho, T.c0 + n, av[slot]);
        };
        XS::last_slots(P.fx, T.g, T.G, T.active, put);
      }
    }
  });
  if (P.eos == EOS_ABSORBING) {
    float av[kMaxSlots], bv[kMaxSlots];
    // rho0 sum du (k is rho0 in the linear case)
    auto ab = [&](int slot, int n) {
      av[slot] = (nonlin ? P.rho0.at(T.c0 + n) : k0[slot]) * du0[slot];
      bv[slot] = (nonlin ? P.rho0.at(T.c1 + n) : k1[slot]) * du1[slot];
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, ab);
    if (T.active) {
      auto put = [&](int slot, int n) {
        KW_ST32(P.rsum, T.c1 + n, rs1[slot]);
        KW_ST32(P.rsum, T.c0 + n, rs0[slot]);
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, put);
    }
    x_forward<XS>(c, P, T, av, bv, P.S[0]);
    x_forward<XS>(c, P, T, rs0, rs1, P.S[1]);
  } else {
    cf pv[kMaxSlots];
    with_eos(P, [&](auto EOS, auto NL) {
      auto f = [&](int slot, int n) {
        constexpr int E = decltype(EOS)::value;
        constexpr bool L = decltype(NL)::value;
        pv[slot].x = eos_p<E, L>(P, T.c0 + n, rs0[slot], du0[slot], 0.f, 0.f);
        pv[slot].y = eos_p<E, L>(P, T.c1 + n, rs1[slot], du1[slot], 0.f, 0.f);
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, f);
    });
    x_finish_p<XS>(c, P, T, pv, t, store_p);
    cf* S0[1] = {P.S[0]};
    x_fwd_from_lds<XS, 1>(c, P, T, S0);
  }
}

// end of the absorbing EOS: p = c0^2 (rsum + tau A - eta B [+ BonA term])
// with A = ifft(nabla1 fft(rho0 sum du)), B = ifft(nabla2 fft(rsum)); sensors;
// S[0] = fft_x(p)
template <class XS, class Ctx>
KW_BODY void body_x_eos(Ctx& cb, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(cb, B, tile);
  const XTile T = x_tile<XS>(cb, P, tile);
  PencilCtx<Ctx> c{cb, T.wave};
  const int t = KW_LD(P.t_index);
  const bool store_p = (A.flags & X_STORE_P) || t >= P.p_store_from || t == P.p_store_at;
  // rsum + tau A first, then - eta B and the nonlinear term; rsum is kept
  // across the second inverse only for the nonlinear term
  cf pv[kMaxSlots];
  with_flag(P.nonlinear != 0, [&](auto NL) {
    constexpr bool nl = decltype(NL)::value;
    cf part[kMaxSlots], rs[nl ? kMaxSlots : 1];
    {
      cf gv[kMaxSlots];
      x_inverse<XS>(c, P, T, P.S[0], gv);
      auto f = [&](int slot, int n) {
        const uint32_t c0 = T.c0 + n, c1 = T.c1 + n;
        const cf r = cf{KW_LD32(P.rsum, c0), KW_LD32(P.rsum, c1)};
        if constexpr (nl) rs[slot] = r;
        part[slot] = cf{r.x + P.tau.at(c0) * gv[slot].x, r.y + P.tau.at(c1) * gv[slot].y};
      };
      XS::last_slots(P.fx, T.g, T.G, T.active, f);
    }
    cf gv[kMaxSlots];
    x_inverse<XS>(c, P, T, P.S[1], gv);
    auto f = [&](int slot, int n) {
      const uint32_t cells[2] = {T.c0 + n, T.c1 + n};
      const float pa[2] = {part[slot].x, part[slot].y};
      const float ga[2] = {gv[slot].x, gv[slot].y};
      float out[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t cell = cells[h];
        float inner = pa[h] - P.eta.at(cell) * ga[h];
        if constexpr (nl) {
          const float r = h == 0 ? rs[slot].x : rs[slot].y;
          inner = inner + P.bona.at(cell) * (r * r) / (2.f * P.rho0.at(cell));
        }
        const float c0 = P.c0.at(cell);
        out[h] = (c0 * c0) * inner;
      }
      pv[slot] = cf{out[0], out[1]};
    };
    XS::last_slots(P.fx, T.g, T.G, T.active, f);
  });
  x_finish_p<XS>(c, P, T, pv, t, store_p);
  cf* S0[1] = {P.S[0]};
  x_fwd_from_lds<XS, 1>(c, P, T, S0);
}

// ---------------------------------------------------------------------------
// elementwise kernels (block of EW_T threads, one cell per thread)
// ---------------------------------------------------------------------------
constexpr int EW_T = 256;

enum EwKind : int32_t {
  EW_INIT_P0 = 0,    // p = p0, rho_d = p0 / (dim c0^2)
  EW_TR_PROLOGUE,    // time reversal: enforce the first data column
  EW_SRC_GRID,       // per-step scatter grid of a many-series additive source
  EW_RECORD_POINTS,  // sensor statistics at sensor points (general mask)
  EW_RECORD_GRID,    // whole-grid statistics (*_all) not fused elsewhere
  EW_FILL,           // zero a float buffer (graph-safe memset)
  EW_RECORD_UNS,     // non-staggered velocity component `which` at the sensor points
  EW_RECORD_SPLIT,   // compressional / shear parts of component `which`
  EW_RECORD_DIR,     // directional p of the sensor points of angle `which`
  EW_AXIS_MUL,       // ptr = src * prof(axis `which`)
  EW_AXIS_FOLD,      // ptr += alpha * src / prof(axis `which`)
  EW_RECORD_RHO,     // sum_d rho_d at the sensor points
  EW_ADD,            // ptr[i] += src[i] for i < n (floats)
  EW_AS_MIRROR,      // axisymmetric: copies 2-4 of a p-like field from copy 1
  EW_AS_WS_COL,      // 'WSWS-FFT': the last radial column of u_y (which 0, 1, 2)
};

// EwArgs.rec value for "the problem's own statistics" (launches shared by
// several problems)
constexpr uint32_t kRecAuto = 0xffffffffu;

struct EwArgs {
  int32_t kind;
  int32_t which;         // EW_SRC_GRID: 0 = p source, 1..3 = u source d-1
  int64_t n;             // EW_FILL length (in floats)
  float* ptr;            // EW_FILL target
  const float* p0;       // EW_INIT_P0; EW_AXIS_*: the source grid
  uint32_t rec;          // EW_RECORD_*: which statistics to update
  const float* prof;     // EW_AXIS_*: 1-D profile along axis `which`
  float alpha;           // EW_AXIS_FOLD; EW_AS_WS_COL: the PML there
  float beta;            // EW_AS_WS_COL: the PML of the column before
};

// coordinate of cell i along axis a (column-major, x fastest)
KW_BODY int64_t axis_coord(const ProbDev& P, int64_t i, int a) {
  if (a == 0) return i % P.Nx;
  if (a == 1) return (i / P.Nx) % P.Ny;
  return i / ((int64_t)P.Nx * P.Ny);
}

// Field f at sensor point i: the grid point of a binary mask (`cell`), or
// for Cartesian points interpolated linearly sum(f(tri) .* bc, 2) over the
// simplex vertices in order (kspaceFirstOrder_extractSensorData.m:301-457)
KW_BODY float sensor_value(const ProbDev& P, const float* f, int64_t i, int64_t cell) {
  if (P.sens_tri == nullptr) return KW_LD(f + cell);
  const int K = P.dim + 1;
  const int64_t o = i * K;
  float s = KW_LD(f + KW_LD(P.sens_tri + o)) * KW_LD(P.sens_bc + o);
  for (int k = 1; k < K; ++k) s += KW_LD(f + KW_LD(P.sens_tri + o + k)) * KW_LD(P.sens_bc + o + k);
  return s;
}

template <class Ctx>
KW_BODY void body_ew(Ctx& c, const Batch& B, const EwArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const int64_t i = (int64_t)tile * EW_T + c.tid();
  const int t = KW_LD(P.t_index);
  switch (A.kind) {
    case EW_AXIS_MUL: {
      if (i >= P.ncell) return;
      KW_ST(A.ptr + i, KW_LD(A.p0 + i) * KW_LD(A.prof + axis_coord(P, i, A.which)));
      return;
    }
    case EW_AXIS_FOLD: {
      if (i >= P.ncell) return;
      const float f = KW_LD(A.prof + axis_coord(P, i, A.which));
      KW_ST(A.ptr + i, KW_LD(A.ptr + i) + A.alpha * KW_LD(A.p0 + i) / f);
      return;
    }
    case EW_AS_MIRROR: {
      // the expanded radial axis holds four copies of No points: a p-like
      // field is [f, 0, -flip(f(2:end)), -f, 0, flip(f(2:end))]
      // (kspaceFirstOrderAS.m:703-707); nonlinear terms do not keep that
      // symmetry, so copies 2-4 are rewritten from copy 1 after each update
      // (A.ptr, or with which = 3 rho_x, rho_y and p in one launch)
      if (i >= P.ncell) return;
      const int No = P.Ny / 4;
      const int64_t x = i % P.Nx, y = i / P.Nx;
      const int b = (int)(y / No), r = (int)(y % No);
      if (b == 0) return;
      const int nf = A.which == 3 ? 3 : 1;
      for (int q = 0; q < nf; ++q) {
        float* f = A.which == 3 ? (q == 0 ? P.rho[0] : q == 1 ? P.rho[1] : P.p) : A.ptr;
        float v = 0.f;
        if (b == 2) {
          v = -KW_LD(f + x + (int64_t)P.Nx * r);
        } else if (r != 0) {
          const float m = KW_LD(f + x + (int64_t)P.Nx * (No - r));
          v = b == 1 ? -m : m;
        }
        KW_ST(f + i, v);
      }
      return;
    }
    case EW_AS_WS_COL: {
      // 'WSWS-FFT' (kspaceFirstOrderAS.m:966-977) drops the last staggered
      // radial point of u_y from its expanded layout (its place holds
      // -u_y of the point before) but still updates and records it.  Row
      // Ny/2 of the expanded grid holds the mirror value while the step's
      // transforms read it, and the reference's value while it is recorded:
      // which 0 puts the mirror value m back before a step; 2 (after the
      // initial condition) and 1 (after an update) recover the gradient
      // there from the mirror value's own update and advance the
      // reference's value u with it, with the PML alpha (beta for the
      // mirror) and prof = (1 / rho0_sgy there) / (1 / rho0_sgy of the
      // mirror).  ptr = [u (Nx), m (Nx)].
      if (i >= P.Nx) return;
      float* row = P.u[1] + (int64_t)(P.Ny / 2) * P.Nx;
      float* uc = A.ptr;
      float* mc = A.ptr + P.Nx;
      if (A.which == 0) {
        KW_ST(row + i, KW_LD(mc + i));
        return;
      }
      const float mn = KW_LD(row + i);
      const float k = KW_LD(A.prof + i);
      // an additive radial velocity source adds the same filtered value to
      // the mirror point and to the reference's last column
      // (kspaceFirstOrderAS.m:879-905 trims the filtered expanded source
      // to the radial grid, that column included)
      float sv = 0.f;
      const TimeSource& us = P.srcu[1];
      if (A.which == 1 && us.mode == SRC_ADDITIVE && us.len >= KW_LD(P.t_index)) {
        const int64_t c = (int64_t)(P.Ny / 2) * P.Nx + i;
        sv = us.grid ? KW_LD(us.grid + c)
                     : KW_LD(us.km + c) * us.value(0, KW_LD(P.t_index) - 1);
      }
      const float un = A.which == 2 ? k * mn
                                    : A.alpha * (A.alpha * KW_LD(uc + i) -
                                                 k * (A.beta * KW_LD(mc + i) - (mn - sv) / A.beta)) +
                                          sv;
      KW_ST(mc + i, mn);
      KW_ST(uc + i, un);
      KW_ST(row + i, un);
      return;
    }
    case EW_ADD: {
      if (i >= A.n) return;
      KW_ST(A.ptr + i, KW_LD(A.ptr + i) + KW_LD(A.p0 + i));
      return;
    }
    case EW_RECORD_RHO: {
      if (i >= P.nsens || t < P.record_start) return;
      const int fi = t - P.record_start + 1;
      const int64_t cell = P.sens_idx ? (int64_t)KW_LD(P.sens_idx + i) : i;
      float r = sensor_value(P, P.rho[0], i, cell);
      for (int d = 1; d < P.dim; ++d) r += sensor_value(P, P.rho[d], i, cell);
      KW_ST(P.rho_ts + i + (int64_t)(fi - 1) * P.nsens, r);
      return;
    }
    case EW_INIT_P0: {
      if (i >= P.ncell) return;
      const float p0 = KW_LD(A.p0 + i);
      const float c0 = P.c0.at(i);
      const float r = p0 / ((float)P.dim * (c0 * c0));
      KW_ST(P.p + i, p0);
      for (int d = 0; d < P.dim; ++d) KW_ST(P.rho[d] + i, r);
      return;
    }
    case EW_TR_PROLOGUE: {
      // p = 0 with p(mask) = reversed column 1 = original column Nt
      // (kspaceFirstOrder3D.m:772-788 at t_index = 1)
      if (i >= P.ncell) return;
      const int32_t k = KW_LD(P.tr_map + i);
      float p = 0.f;
      if (k >= 0) {
        p = KW_LD(P.tr_data + (int64_t)k + (int64_t)(P.Nt - 1) * P.nsens);
        const float c0 = P.c0.at(i);
        const float c2 = c0 * c0;
        const float r = tr_rho(P, p, c2);
        for (int d = 0; d < P.dim; ++d) KW_ST(P.rho[d] + i, r);
      }
      KW_ST(P.p + i, p);
      return;
    }
    case EW_SRC_GRID: {
      // launched before the step's first pass: t_index still holds the
      // previous step
      if (i >= P.ncell) return;
      const TimeSource& s = A.which == 0 ? P.srcp : P.srcu[A.which - 1];
      const int32_t k = KW_LD(s.map + i);
      const int ts = t + 1;
      const float v = (k >= 0 && s.len >= ts) ? s.value(k, ts - 1) : 0.f;
      KW_ST((float*)s.grid + i, v);
      return;
    }
    case EW_RECORD_POINTS: {
      if (i >= P.nsens || t < P.record_start) return;
      const int fi = t - P.record_start + 1;
      const int64_t cell = P.sens_idx ? (int64_t)KW_LD(P.sens_idx + i) : i;
      const uint32_t rec = A.rec == kRecAuto ? P.rec_points : A.rec;
      const float pv = sensor_value(P, P.p, i, cell);
      const double pd = (double)pv;
      if (rec & REC_P) KW_ST(P.p_ts + i + (int64_t)(fi - 1) * P.nsens, pv);
      if (rec & REC_P_MAX) {
        const double m = fi == 1 ? pd : KW_LD(P.p_max + i);
        KW_ST(P.p_max + i, pd > m ? pd : m);
      }
      if (rec & REC_P_MIN) {
        const double m = fi == 1 ? pd : KW_LD(P.p_min + i);
        KW_ST(P.p_min + i, pd < m ? pd : m);
      }
      if (rec & REC_P_RMS) {
        const double s = fi == 1 ? 0.0 : KW_LD(P.p_rms2 + i);
        KW_ST(P.p_rms2 + i, s + pd * pd);
      }
      for (int d = 0; d < P.dim; ++d) {
        const float uv = sensor_value(P, P.u[d], i, cell);
        const double ud = (double)uv;
        if (rec & REC_U) KW_ST(P.u_ts[d] + i + (int64_t)(fi - 1) * P.nsens, uv);
        if (rec & REC_U_MAX) {
          const double m = fi == 1 ? ud : KW_LD(P.u_max[d] + i);
          KW_ST(P.u_max[d] + i, ud > m ? ud : m);
        }
        if (rec & REC_U_MIN) {
          const double m = fi == 1 ? ud : KW_LD(P.u_min[d] + i);
          KW_ST(P.u_min[d] + i, ud < m ? ud : m);
        }
        if (rec & REC_U_RMS) {
          const double s = fi == 1 ? 0.0 : KW_LD(P.u_rms2[d] + i);
          KW_ST(P.u_rms2[d] + i, s + ud * ud);
        }
      }
      return;
    }
    case EW_RECORD_GRID: {
      if (i >= P.ncell || t < P.record_start) return;
      const int fi = t - P.record_start + 1;
      const uint32_t rec = A.rec == kRecAuto ? P.rec_grid : A.rec;
      if (rec & (REC_P_MAX_ALL | REC_P_MIN_ALL)) {
        const double pd = (double)KW_LD(P.p + i);
        if (rec & REC_P_MAX_ALL) {
          const double m = fi == 1 ? pd : KW_LD(P.p_max_all + i);
          KW_ST(P.p_max_all + i, pd > m ? pd : m);
        }
        if (rec & REC_P_MIN_ALL) {
          const double m = fi == 1 ? pd : KW_LD(P.p_min_all + i);
          KW_ST(P.p_min_all + i, pd < m ? pd : m);
        }
      }
      for (int d = 0; d < P.dim; ++d) {
        if (!(rec & (REC_U_MAX_ALL | REC_U_MIN_ALL))) break;
        const double ud = (double)KW_LD(P.u[d] + i);
        if (rec & REC_U_MAX_ALL) {
          const double m = fi == 1 ? ud : KW_LD(P.u_max_all[d] + i);
          KW_ST(P.u_max_all[d] + i, ud > m ? ud : m);
        }
        if (rec & REC_U_MIN_ALL) {
          const double m = fi == 1 ? ud : KW_LD(P.u_min_all[d] + i);
          KW_ST(P.u_min_all[d] + i, ud < m ? ud : m);
        }
      }
      return;
    }
    case EW_RECORD_UNS: {
      // kspaceFirstOrder_extractSensorData.m:111-122: component `which` of
      // the shifted velocity (in P.uns) at the sensor points
      if (i >= P.nsens || t < P.record_start || P.uns == nullptr) return;
      const int fi = t - P.record_start + 1;
      const int64_t cell = P.sens_idx ? (int64_t)KW_LD(P.sens_idx + i) : i;
      KW_ST(P.uns_ts[A.which] + i + (int64_t)(fi - 1) * P.nsens, sensor_value(P, P.uns, i, cell));
      return;
    }
    case EW_RECORD_SPLIT: {
      // part p of component `which` is in P.usp, the component on the
      // non-staggered grid in P.uns; part s is their difference
      if (i >= P.nsens || t < P.record_start || P.usp == nullptr) return;
      const int fi = t - P.record_start + 1;
      const int64_t cell = P.sens_idx ? (int64_t)KW_LD(P.sens_idx + i) : i;
      const int64_t o = i + (int64_t)(fi - 1) * P.nsens;
      const float up = KW_LD(P.usp + cell);
      const float un = KW_LD(P.uns + cell);
      KW_ST(P.usplit_p[A.which] + o, up);
      KW_ST(P.usplit_s[A.which] + o, un - up);
      if (P.uns_ts[A.which] != nullptr) KW_ST(P.uns_ts[A.which] + o, un);
      return;
    }
    case EW_RECORD_DIR: {
      // kspaceFirstOrder2D.m:945-947: the p time series of the points whose
      // angle is number `which`, from P.dirp (p filtered for that angle)
      if (i >= P.nsens || t < P.record_start || P.dir_id == nullptr) return;
      if (KW_LD(P.dir_id + i) != A.which) return;
      const int fi = t - P.record_start + 1;
      const int64_t cell = P.sens_idx ? (int64_t)KW_LD(P.sens_idx + i) : i;
      KW_ST(P.p_ts + i + (int64_t)(fi - 1) * P.nsens, KW_LD(P.dirp + cell));
      return;
    }
    case EW_FILL: {
      if (i < A.n) KW_ST(A.ptr + i, 0.f);
      return;
    }
  }
}

}  // namespace kw
