// Mixed-radix (2, 3, 4, 5, 8) Stockham FFT over pencils staged in LDS.
//
// A pencil of length N is transformed by G threads.  Stage s with radix R
// and Ns = product of the radices before it performs N/R butterflies; the
// butterfly b reads elements b + r*N/R (r < R), twiddles them by
// w^(k*r) with w = exp(-+2 pi i / (Ns*R)) and k = b mod Ns, applies a radix-R
// DFT and writes element (b - k)*R + k + r*Ns.  After the last stage the
// pencil is in natural order (self-sorting, no bit reversal).  Each stage
// reads all of its inputs, synchronises the block, then writes its outputs,
// so a single LDS buffer per pencil is transformed in place.
//
// A thread works on NP pencils at once (for example the x, y and z
// components of a field): their loads are issued together, the twiddles and
// barriers are shared, so NP independent transforms cost one transform's
// latency.
//
// The first stage takes its inputs from a caller functor `src(j, n)` (pencil
// j, element n) and the last stage hands its outputs to
// `dst(j, n, value, slot)`, which is how loads from HBM (with any fused
// pre-processing) and stores to HBM (with any fused epilogue) are folded into
// the FFT.  `slot` is a compile-time-constant index of the output within the
// calling thread, identical for every transform with the same plan, so
// epilogues can keep per-output values in registers across transforms.
//
// Two engines share the stage code: `fft_rt` takes a runtime plan (any
// length that factors into 2, 3 and 5; one code path per radix) and
// `fft_ct<N>` a length fixed at compile time, for which every index is a
// constant and the stages unroll into straight-line code.
//
// Twiddles come from a per-length table of the N roots of unity
// exp(-2 pi i j / N) computed in double precision on the host and rounded
// once to float.  No normalisation is applied in either direction.
#pragma once

#include "kw_common.h"

namespace kw {

// A transform of length N.  For N = 2^a 3^b 5^c 7^d the stages run at
// length L = N.  Any other N runs as Bluestein's chirp-z transform: with the
// chirp w_n = exp(-i pi n^2 / N), X_k = w_k sum_n (x_n w_n) conj(w_(k-n)),
// a circular convolution of length L >= 2N - 1 (L = 2^a 3^b 5^c) computed
// with two length-L transforms in LDS; `chirp` (N entries) and `bspec`
// (L entries: the length-L transform of the chirp filter conj(w_m),
// m = -(N-1) .. N-1, wrapped) are set, and the inverse uses their conjugates.
struct FftPlan {
  int32_t N;   // length of the transform
  int32_t L;   // length the stages run at (N, or the chirp-z length)
  int32_t nst;
  int32_t radix[KW_MAX_STAGES];
  const cf* tw;     // L twiddles exp(-2 pi i j / L)
  const cf* chirp;  // Bluestein only (else NULL): w_n, n < N
  const cf* bspec;  // Bluestein only: transform of the chirp filter, L entries
};

// factorisation rule shared by the host planner and the compile-time plans:
// radix-8 stages first, then one radix-4 or radix-2 stage, then 3s and 5s
constexpr int ct_nst(int N) {
  int n = N, c = 0;
  while (n % 8 == 0) { n /= 8; ++c; }
  if (n % 4 == 0) { n /= 4; ++c; }
  if (n % 2 == 0) { n /= 2; ++c; }
  while (n % 3 == 0) { n /= 3; ++c; }
  while (n % 5 == 0) { n /= 5; ++c; }
  return n == 1 ? c : -1;
}
constexpr int ct_radix(int N, int s) {
  int n = N, c = 0;
  while (n % 8 == 0) { n /= 8; if (c++ == s) return 8; }
  if (n % 4 == 0) { n /= 4; if (c++ == s) return 4; }
  if (n % 2 == 0) { n /= 2; if (c++ == s) return 2; }
  while (n % 3 == 0) { n /= 3; if (c++ == s) return 3; }
  while (n % 5 == 0) { n /= 5; if (c++ == s) return 5; }
  return 0;
}
// product of the radices before stage s
constexpr int ct_ns(int N, int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= ct_radix(N, i);
  return p;
}

// upper bound of butterflies per thread for the runtime engine
template <int R>
struct MaxB {
  static constexpr int v = (KW_E + R - 1) / R;
};
// lengths with a factor 7 run with G = ceil(N/7) threads (see kw_sim.h), so
// a radix-7 stage has one butterfly per thread
template <>
struct MaxB<7> {
  static constexpr int v = 1;
};
// largest number of last-stage outputs one thread can own with G = ceil(N/KW_E)
constexpr int kMaxSlots = 10;
static_assert(MaxB<2>::v * 2 <= kMaxSlots && MaxB<3>::v * 3 <= kMaxSlots &&
                  MaxB<4>::v * 4 <= kMaxSlots && MaxB<5>::v * 5 <= kMaxSlots &&
                  MaxB<8>::v * 8 <= kMaxSlots,
              "slot bound");

// Twiddle table of a plan as the stages read it: in HBM (each read a vector
// memory load through the same address path as the data streams), or
// copied once per block into LDS (stage_twiddles).
struct TwG {
  const cf* p;
};
struct TwL {
  const cf* p;
};

template <bool INV>
KW_BODY cf twiddle(const TwG& tw, int idx) {
  cf w = KW_LD32(tw.p, (uint32_t)idx);
  if (INV) w.y = -w.y;
  return w;
}
template <bool INV>
KW_BODY cf twiddle(const TwL& tw, int idx) {
  cf w = lds_ld(tw.p + idx);
  if (INV) w.y = -w.y;
  return w;
}

// the L twiddles of a plan copied into the block's LDS at `dst` by all
// threads of the block, followed by a block barrier
template <class Ctx>
KW_BODY TwL stage_twiddles(Ctx& c, const FftPlan& pl, cf* dst) {
  for (int i = c.tid(); i < pl.L; i += c.nthreads()) lds_st(dst + i, KW_LD32(pl.tw, (uint32_t)i));
  c.sync();
  return TwL{dst};
}

// radix-R DFT in place; forward uses exp(-2 pi i rs/R), inverse exp(+...)
template <int R, bool INV>
KW_BODY void dft(cf* v);

template <>
KW_BODY void dft<2, false>(cf* v) {
  cf a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <>
KW_BODY void dft<2, true>(cf* v) {
  dft<2, false>(v);
}

template <bool INV>
KW_BODY void dft3(cf* v) {
  const float s3 = 0.866025403784438646763723170752936183f;  // sin(2 pi/3)
  cf t = cadd(v[1], v[2]);
  cf d = csub(v[1], v[2]);
  cf m = cf{v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y};
  cf sd = cscale(d, s3);
  // forward: X1 = m - i sd, X2 = m + i sd
  cf isd = INV ? cmul_pi(sd) : cmul_mi(sd);
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, isd);
  v[2] = csub(m, isd);
}
template <>
KW_BODY void dft<3, false>(cf* v) { dft3<false>(v); }
template <>
KW_BODY void dft<3, true>(cf* v) { dft3<true>(v); }

template <bool INV>
KW_BODY void dft4(cf* v) {
  cf t0 = cadd(v[0], v[2]);
  cf t1 = csub(v[0], v[2]);
  cf t2 = cadd(v[1], v[3]);
  cf t3 = csub(v[1], v[3]);
  // forward: X1 = t1 - i t3, X3 = t1 + i t3
  cf it3 = INV ? cmul_pi(t3) : cmul_mi(t3);
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, it3);
  v[3] = csub(t1, it3);
}
template <>
KW_BODY void dft<4, false>(cf* v) { dft4<false>(v); }
template <>
KW_BODY void dft<4, true>(cf* v) { dft4<true>(v); }

template <bool INV>
KW_BODY void dft5(cf* v) {
  const float c1 = 0.309016994374947424102293417182819059f;   // cos(2 pi/5)
  const float c2 = -0.809016994374947424102293417182819059f;  // cos(4 pi/5)
  const float s1 = 0.951056516295153572116439333379382143f;   // sin(2 pi/5)
  const float s2 = 0.587785252292473129168705954639072769f;   // sin(4 pi/5)
  cf a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  cf a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  cf x0 = v[0];
  cf m1 = cf{x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
  cf m2 = cf{x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
  cf n1 = cf{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
  cf n2 = cf{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
  // forward: X1 = m1 - i n1, X4 = m1 + i n1, X2 = m2 - i n2, X3 = m2 + i n2
  cf in1 = INV ? cmul_pi(n1) : cmul_mi(n1);
  cf in2 = INV ? cmul_pi(n2) : cmul_mi(n2);
  v[0] = cf{x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
  v[1] = cadd(m1, in1);
  v[4] = csub(m1, in1);
  v[2] = cadd(m2, in2);
  v[3] = csub(m2, in2);
}
template <>
KW_BODY void dft<5, false>(cf* v) { dft5<false>(v); }
template <>
KW_BODY void dft<5, true>(cf* v) { dft5<true>(v); }

// radix 7 from the pairs a_k = v_k + v_(7-k), b_k = v_k - v_(7-k):
// X_j = v_0 + sum_k cos(2 pi jk/7) a_k -+ i sum_k sin(2 pi jk/7) b_k
template <bool INV>
KW_BODY void dft7(cf* v) {
  const float c1 = 0.6234898018587336f, c2 = -0.22252093395631434f, c3 = -0.9009688679024191f;
  const float s1 = 0.7818314824680298f, s2 = 0.9749279121818236f, s3 = 0.43388373911755823f;
  const cf a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
  const cf a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
  const cf a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
  const cf x0 = v[0];
  auto lin = [](cf z, float p, cf a, float q, cf b, float r, cf c) {
    return cf{z.x + p * a.x + q * b.x + r * c.x, z.y + p * a.y + q * b.y + r * c.y};
  };
  const cf zero{0.f, 0.f};
  const cf m1 = lin(x0, c1, a1, c2, a2, c3, a3);
  const cf m2 = lin(x0, c2, a1, c3, a2, c1, a3);
  const cf m3 = lin(x0, c3, a1, c1, a2, c2, a3);
  const cf n1 = lin(zero, s1, b1, s2, b2, s3, b3);
  const cf n2 = lin(zero, s2, b1, -s3, b2, -s1, b3);
  const cf n3 = lin(zero, s3, b1, -s1, b2, s2, b3);
  // forward: X_j = m_j - i n_j, X_(7-j) = m_j + i n_j
  const cf i1 = INV ? cmul_pi(n1) : cmul_mi(n1);
  const cf i2 = INV ? cmul_pi(n2) : cmul_mi(n2);
  const cf i3 = INV ? cmul_pi(n3) : cmul_mi(n3);
  v[0] = cf{x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y};
  v[1] = cadd(m1, i1);
  v[6] = csub(m1, i1);
  v[2] = cadd(m2, i2);
  v[5] = csub(m2, i2);
  v[3] = cadd(m3, i3);
  v[4] = csub(m3, i3);
}
template <>
KW_BODY void dft<7, false>(cf* v) { dft7<false>(v); }
template <>
KW_BODY void dft<7, true>(cf* v) { dft7<true>(v); }

template <bool INV>
KW_BODY void dft8(cf* v) {
  const float h = 0.707106781186547524400844362104849039f;
  cf e[4] = {v[0], v[2], v[4], v[6]};
  cf o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  // o[k] *= w8^k, w8 = exp(-+ 2 pi i/8)
  cf w1 = INV ? cf{h, h} : cf{h, -h};
  cf w3 = INV ? cf{-h, h} : cf{-h, -h};
  o[1] = cmul(o[1], w1);
  o[2] = INV ? cmul_pi(o[2]) : cmul_mi(o[2]);
  o[3] = cmul(o[3], w3);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}
template <>
KW_BODY void dft<8, false>(cf* v) { dft8<false>(v); }
template <>
KW_BODY void dft<8, true>(cf* v) { dft8<true>(v); }

KW_BODY int mod_ns(int b, int Ns) {
  // b mod Ns; Ns is a compile-time constant in the fixed-length engine and
  // a power of two for every stage before the first radix-3/5 stage
  return (Ns & (Ns - 1)) == 0 ? (b & (Ns - 1)) : b % Ns;
}

// One Stockham stage over NP pencils (see the file comment).  MB is the
// number of butterflies per thread (exact for the fixed-length engine, an
// upper bound for the runtime engine).  `src` must be safe to call for
// every thread, including threads with active == false.
template <int R, int NP, bool INV, int MB, class Ctx, class TW, class Src, class Dst>
KW_BODY void stage(Ctx& c, int N, int Ns, int G, int g, bool active, const TW& tw, Src& src,
                   Dst& dst) {
  const int nb = N / R;
  const int twstep = N / (Ns * R);
  cf v[NP][MB][R];
  // loads and arithmetic run for every thread (callers make src safe for
  // threads without a pencil) so that they are not split into conditional
  // regions; only the stores are guarded by `active`
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = g + m * G;
    if (MB * G == nb || b < nb) {
      cf w[R];
      if (Ns > 1) {
        const int k = mod_ns(b, Ns);
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = twiddle<INV>(tw, k * r * twstep);
      }
#pragma unroll
      for (int j = 0; j < NP; ++j) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[j][m][r] = src(j, b + r * nb);
      }
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        if (Ns > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) v[j][m][r] = cmul(v[j][m][r], w[r]);
        }
        dft<R, INV>(v[j][m]);
      }
    }
  }
  c.sync();
#pragma unroll
  for (int j = 0; j < NP; ++j) {
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const int b = g + m * G;
      if (active && (MB * G == nb || b < nb)) {
        const int k = mod_ns(b, Ns);
        const int j0 = (b - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst(j, j0 + r * Ns, v[j][m][r], m * R + r);
      }
    }
  }
  c.sync();
}

// LDS view of NP pencils: element n of pencil j at base[j*sj + n*sn + pad]
// with pad = (n >> 3) when `padded` (breaks the power-of-two strides of the
// Stockham writes across LDS banks)
struct LdsPencils {
  cf* base;
  int sj, sn;
  bool padded;
  KW_BODY int at(int j, int n) const { return j * sj + n * sn + (padded ? (n >> 3) : 0); }
  KW_BODY cf get(int j, int n) const { return lds_ld(base + at(j, n)); }
  KW_BODY void set(int j, int n, cf v) const { lds_st(base + at(j, n), v); }
};

// ---- runtime-length engine --------------------------------------------------
// Any N = 2^a 3^b 5^c 7^d.  The transform is staged entirely in LDS: a strided,
// coalesced pass loads the pencil into LDS through the caller's functor, the
// Stockham stages run from LDS to LDS, and a strided pass hands element
// n = g + i*G to the caller as slot i.  The caller's functors are
// instantiated once per call site instead of once per radix and stage
// position, which would multiply the code size past the instruction cache.

constexpr int kRtSlots = 10;  // ceil(N / G) for G = ceil(N / KW_E)
static_assert(kRtSlots <= kMaxSlots, "slot bound");

// Threads without a pencil (active == false) may share another pencil's LDS
// to keep their loads in bounds, so they never store, here as in the
// fixed-length engine.
template <int NP, bool INV, class Ctx, class TW>
KW_BODY void fft_stages_rt(Ctx& c, const FftPlan& pl, const TW& tw, int g, int G, bool active,
                           const LdsPencils& lds) {
  auto lsrc = [&](int j, int n) -> cf { return lds.get(j, n); };
  auto ldst = [&](int j, int n, cf v, int) { lds.set(j, n, v); };
  int Ns = 1;
  const int L = pl.L;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.radix[s];
    switch (R) {
      case 2: stage<2, NP, INV, MaxB<2>::v>(c, L, Ns, G, g, active, tw, lsrc, ldst); break;
      case 3: stage<3, NP, INV, MaxB<3>::v>(c, L, Ns, G, g, active, tw, lsrc, ldst); break;
      case 4: stage<4, NP, INV, MaxB<4>::v>(c, L, Ns, G, g, active, tw, lsrc, ldst); break;
      case 5: stage<5, NP, INV, MaxB<5>::v>(c, L, Ns, G, g, active, tw, lsrc, ldst); break;
      case 7: stage<7, NP, INV, MaxB<7>::v>(c, L, Ns, G, g, active, tw, lsrc, ldst); break;
      default: stage<8, NP, INV, MaxB<8>::v>(c, L, Ns, G, g, active, tw, lsrc, ldst); break;
    }
    Ns *= R;
  }
}

// Bluestein's transform of length N (see FftPlan) over the length-L stages:
// the chirped input zero-padded to L, a forward transform, the product with
// the filter's transform, an inverse transform, and the chirped outputs
// scaled by 1/L; the inverse transform conjugates the chirp and the filter.
// Each thread reads and writes only its own elements n = g + i G between
// barriers, as the plain runtime engine does.
template <int NP, bool INV, class Ctx, class TW, class Src, class Dst>
KW_BODY void fft_bluestein(Ctx& c, const FftPlan& pl, const TW& tw, int g, int G, bool active,
                           Src& src, Dst& dst, const LdsPencils& lds) {
  const int N = pl.N, L = pl.L;
  auto chirp = [&](int n) -> cf {
    cf w = KW_LD32(pl.chirp, (uint32_t)n);
    if (INV) w.y = -w.y;
    return w;
  };
#pragma unroll
  for (int i = 0; i < kRtSlots; ++i) {
    const int n = g + i * G;
    if (n < L) {
      const cf w = n < N ? chirp(n) : cf{0.f, 0.f};
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const cf v = n < N ? cmul(src(j, n), w) : cf{0.f, 0.f};
        if (active) lds.set(j, n, v);
      }
    }
  }
  c.sync();
  fft_stages_rt<NP, false>(c, pl, tw, g, G, active, lds);
#pragma unroll
  for (int i = 0; i < kRtSlots; ++i) {
    const int n = g + i * G;
    if (active && n < L) {
      cf b = KW_LD32(pl.bspec, (uint32_t)n);
      if (INV) b.y = -b.y;
#pragma unroll
      for (int j = 0; j < NP; ++j) lds.set(j, n, cmul(lds.get(j, n), b));
    }
  }
  c.sync();
  fft_stages_rt<NP, true>(c, pl, tw, g, G, active, lds);
  const float inv_l = 1.0f / (float)L;
#pragma unroll
  for (int i = 0; i < kRtSlots; ++i) {
    const int n = g + i * G;
    if (active && n < N) {
      const cf w = cscale(chirp(n), inv_l);
#pragma unroll
      for (int j = 0; j < NP; ++j) dst(j, n, cmul(lds.get(j, n), w), i);
    }
  }
  c.sync();
}

template <int NP, bool INV, class Ctx, class TW, class Src, class Dst>
KW_BODY void fft_rt(Ctx& c, const FftPlan& pl, const TW& tw, int g, int G, bool active, Src& src,
                    Dst& dst, const LdsPencils& lds) {
  const int N = pl.N;
  g = kw_fresh(g);
  if (pl.chirp) {
    fft_bluestein<NP, INV>(c, pl, tw, g, G, active, src, dst, lds);
    return;
  }
#pragma unroll
  for (int i = 0; i < kRtSlots; ++i) {
    const int n = g + i * G;
    if (n < N) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const cf v = src(j, n);
        if (active) lds.set(j, n, v);
      }
    }
  }
  c.sync();
  fft_stages_rt<NP, INV>(c, pl, tw, g, G, active, lds);
#pragma unroll
  for (int i = 0; i < kRtSlots; ++i) {
    const int n = g + i * G;
    if (active && n < N) {
#pragma unroll
      for (int j = 0; j < NP; ++j) dst(j, n, lds.get(j, n), i);
    }
  }
  c.sync();
}

// the (slot, n) pairs this thread owns after a runtime-length transform
template <class F>
KW_BODY void last_slots_rt(const FftPlan& pl, int g, int G, bool, F& f) {
  g = kw_fresh(g);
#pragma unroll
  for (int i = 0; i < kRtSlots; ++i) {
    const int n = g + i * G;
    if (n < pl.N) f(i, n);
  }
}

// ---- fixed-length engine ------------------------------------------------------

template <int N, int G, int S, int NST, int NP, bool INV>
struct CtStages {
  static constexpr int R = ct_radix(N, S);
  static constexpr int Ns = ct_ns(N, S);
  static constexpr int MB = (N / R + G - 1) / G;
  template <class Ctx, class TW, class Src, class Dst, class LSrc, class LDst>
  KW_BODY static void run(Ctx& c, int g, bool active, const TW& tw, Src& src0, Dst& dstN,
                          LSrc& lsrc, LDst& ldst) {
    if constexpr (S == 0 && NST == 1) {
      stage<R, NP, INV, MB>(c, N, Ns, G, g, active, tw, src0, dstN);
    } else if constexpr (S == 0) {
      stage<R, NP, INV, MB>(c, N, Ns, G, g, active, tw, src0, ldst);
    } else if constexpr (S == NST - 1) {
      stage<R, NP, INV, MB>(c, N, Ns, G, g, active, tw, lsrc, dstN);
    } else {
      stage<R, NP, INV, MB>(c, N, Ns, G, g, active, tw, lsrc, ldst);
    }
    if constexpr (S + 1 < NST)
      CtStages<N, G, S + 1, NST, NP, INV>::run(c, g, active, tw, src0, dstN, lsrc, ldst);
  }
};

template <int N, int G, int NP, bool INV, class Ctx, class TW, class Src, class Dst>
KW_BODY void fft_ct(Ctx& c, const TW& tw, int g, bool active, Src& src0, Dst& dstN,
                    const LdsPencils& lds) {
  static_assert(ct_nst(N) > 0, "length must factor into 2, 3 and 5");
  g = kw_fresh(g);
  auto lsrc = [&](int j, int n) -> cf { return lds.get(j, n); };
  auto ldst = [&](int j, int n, cf v, int) { lds.set(j, n, v); };
  CtStages<N, G, 0, ct_nst(N), NP, INV>::run(c, g, active, tw, src0, dstN, lsrc, ldst);
}

template <int N, int G, class F>
KW_BODY void last_slots_ct(int g, bool, F& f) {
  constexpr int S = ct_nst(N) - 1;
  constexpr int R = ct_radix(N, S);
  constexpr int Ns = N / R;
  constexpr int nb = N / R;
  constexpr int MB = (nb + G - 1) / G;
  g = kw_fresh(g);
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = g + m * G;
    if (MB * G == nb || b < nb) {
      const int k = mod_ns(b, Ns);
      const int j0 = (b - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) f(m * R + r, j0 + r * Ns);
    }
  }
}

// ---- plan selectors -----------------------------------------------------------
// A kernel body templated on a selector runs the runtime engine (PlanRt) or
// the fixed-length engine (PlanCt<N>) for one axis.

struct PlanRt {
  static constexpr int fixed = 0;
  static constexpr int G = 0;  // threads per pencil: chosen at run time
  template <int NP, bool INV, class Ctx, class Src, class Dst>
  KW_BODY static void fft(Ctx& c, const FftPlan& pl, int g, int G, bool active, Src& src,
                          Dst& dst, const LdsPencils& lds) {
    fft_rt<NP, INV>(c, pl, TwG{pl.tw}, g, G, active, src, dst, lds);
  }
  // the same with the twiddles read from `tw` (for example staged in LDS)
  template <int NP, bool INV, class Ctx, class TW, class Src, class Dst>
  KW_BODY static void fft_tw(Ctx& c, const FftPlan& pl, const TW& tw, int g, int G, bool active,
                             Src& src, Dst& dst, const LdsPencils& lds) {
    fft_rt<NP, INV>(c, pl, tw, g, G, active, src, dst, lds);
  }
  template <class F>
  KW_BODY static void last_slots(const FftPlan& pl, int g, int G, bool active, F& f) {
    last_slots_rt(pl, g, G, active, f);
  }
};

// threads per pencil of a fixed-length plan: ceil(N / KW_E)
KW_HOSTDEV constexpr int ct_threads(int N) { return (N + KW_E - 1) / KW_E; }

// threads per pencil of the fixed-length x passes: ceil(N / KW_E) rounded up
// to a power of two when that is at most 64, so that every pencil lies in
// one wavefront and the stages synchronise with wavefront barriers instead
// of block barriers (the extra lanes idle in stages with fewer butterflies
// than lanes)
KW_HOSTDEV constexpr int x_threads(int N) {
  int g = ct_threads(N), p = 1;
  if (g > 64) return g;
  while (p < g) p <<= 1;
  return p;
}

template <int N_, int G_ = ct_threads(N_)>
struct PlanCt {
  static constexpr int fixed = N_;
  // threads per pencil; when N is not a multiple of G some threads get
  // fewer butterflies (the stages guard them)
  static constexpr int G = G_;
  static constexpr int last_R = ct_radix(N_, ct_nst(N_) - 1);
  static_assert(((N_ / last_R + G - 1) / G) * last_R <= kMaxSlots, "slot bound");
  template <int NP, bool INV, class Ctx, class Src, class Dst>
  KW_BODY static void fft(Ctx& c, const FftPlan& pl, int g, int, bool active, Src& src, Dst& dst,
                          const LdsPencils& lds) {
    fft_ct<N_, G, NP, INV>(c, TwG{pl.tw}, g, active, src, dst, lds);
  }
  template <int NP, bool INV, class Ctx, class TW, class Src, class Dst>
  KW_BODY static void fft_tw(Ctx& c, const FftPlan&, const TW& tw, int g, int, bool active,
                             Src& src, Dst& dst, const LdsPencils& lds) {
    fft_ct<N_, G, NP, INV>(c, tw, g, active, src, dst, lds);
  }
  template <class F>
  KW_BODY static void last_slots(const FftPlan&, int g, int, bool active, F& f) {
    last_slots_ct<N_, G>(g, active, f);
  }
};

// the plan of the x passes for a fixed length
template <int N>
using PlanCtX = PlanCt<N, x_threads(N)>;

}  // namespace kw
