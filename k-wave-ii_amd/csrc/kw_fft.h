// Mixed-radix (2, 3, 4, 5, 8) Stockham FFT over pencils staged in LDS.
//
// A pencil of length N is transformed by G threads.  Stage s with radix R
// and Ns = product of the radices before it performs N/R butterflies; the
// butterfly b reads elements b + r*N/R (r < R), twiddles them by
// w^(k*r) with w = exp(-+2 pi i / (Ns*R)) and k = b mod Ns, applies a radix-R
// DFT and writes element (b - k)*R + k + r*Ns.  After the last stage the
// pencil is in natural order (self-sorting, no bit reversal).  Each stage
// reads all of its inputs, synchronises the block, then writes its outputs,
// so a single LDS buffer is transformed in place.
//
// The first stage takes its inputs from a caller functor `src(n)` and the
// last stage hands its outputs to `dst(n, value, slot)`, which is how loads
// from HBM (with any fused pre-processing) and stores to HBM (with any fused
// epilogue) are folded into the FFT.  `slot` is a compile-time-constant index
// of the output within the calling thread, identical for every transform
// with the same plan, so epilogues can keep per-output values in registers
// across several transforms.
//
// Twiddles come from a per-length table of the N roots of unity
// exp(-2 pi i j / N) computed in double precision on the host and rounded
// once to float.  No normalisation is applied in either direction.
#pragma once

#include "kw_common.h"

namespace kw {

struct FftPlan {
  int32_t N;
  int32_t nst;
  int32_t radix[KW_MAX_STAGES];
  const cf* tw;  // N entries
};

template <int R>
struct MaxB {
  static constexpr int v = (KW_E + R - 1) / R;
};
// largest number of last-stage outputs one thread can own (over all radices)
constexpr int kMaxSlots = 10;
static_assert(MaxB<2>::v * 2 <= kMaxSlots && MaxB<3>::v * 3 <= kMaxSlots &&
                  MaxB<4>::v * 4 <= kMaxSlots && MaxB<5>::v * 5 <= kMaxSlots &&
                  MaxB<8>::v * 8 <= kMaxSlots,
              "slot bound");

template <bool INV>
KW_BODY cf twiddle(const cf* tw, int idx) {
  cf w = KW_LD(tw + idx);
  if (INV) w.y = -w.y;
  return w;
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

// One Stockham stage for one pencil (see the file comment).
template <int R, bool INV, class Ctx, class Src, class Dst>
KW_BODY void stage(Ctx& c, int N, int Ns, int g, int G, bool active, const cf* tw, Src& src,
                   Dst& dst) {
  constexpr int MB = MaxB<R>::v;
  const int nb = N / R;
  const int twstep = N / (Ns * R);
  cf v[MB][R];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = g + m * G;
    if (active && b < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[m][r] = src(b + r * nb);
      if (Ns > 1) {
        const int k = b % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) v[m][r] = cmul(v[m][r], twiddle<INV>(tw, k * r * twstep));
      }
      dft<R, INV>(v[m]);
    }
  }
  c.sync();
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = g + m * G;
    if (active && b < nb) {
      const int k = b % Ns;
      const int j0 = (b - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst(j0 + r * Ns, v[m][r], m * R + r);
    }
  }
  c.sync();
}

template <bool INV, class Ctx, class Src, class Dst>
KW_BODY void stage_dispatch(int R, Ctx& c, int N, int Ns, int g, int G, bool active,
                            const cf* tw, Src& src, Dst& dst) {
  switch (R) {
    case 2: stage<2, INV>(c, N, Ns, g, G, active, tw, src, dst); break;
    case 3: stage<3, INV>(c, N, Ns, g, G, active, tw, src, dst); break;
    case 4: stage<4, INV>(c, N, Ns, g, G, active, tw, src, dst); break;
    case 5: stage<5, INV>(c, N, Ns, g, G, active, tw, src, dst); break;
    default: stage<8, INV>(c, N, Ns, g, G, active, tw, src, dst); break;
  }
}

// LDS view of one pencil: element n at base[n * sn]
struct LdsPencil {
  cf* base;
  int sn;
  KW_BODY cf get(int n) const { return base[n * sn]; }
  KW_BODY void set(int n, cf v) const { base[n * sn] = v; }
};

// Transform one pencil (called by all threads of the block; `active` is
// false for threads without a pencil, which still take part in barriers).
template <bool INV, class Ctx, class Src, class Dst>
KW_BODY void fft(Ctx& c, const FftPlan& pl, int g, int G, bool active, Src& src0, Dst& dstN,
                 const LdsPencil& lds) {
  auto lsrc = [&](int n) -> cf { return lds.get(n); };
  auto ldst = [&](int n, cf v, int) { lds.set(n, v); };
  int Ns = 1;
  const int nst = pl.nst;
  for (int s = 0; s < nst; ++s) {
    const int R = pl.radix[s];
    const bool first = (s == 0), last = (s == nst - 1);
    if (first && last)
      stage_dispatch<INV>(R, c, pl.N, Ns, g, G, active, pl.tw, src0, dstN);
    else if (first)
      stage_dispatch<INV>(R, c, pl.N, Ns, g, G, active, pl.tw, src0, ldst);
    else if (last)
      stage_dispatch<INV>(R, c, pl.N, Ns, g, G, active, pl.tw, lsrc, dstN);
    else
      stage_dispatch<INV>(R, c, pl.N, Ns, g, G, active, pl.tw, lsrc, ldst);
    Ns *= R;
  }
}

// Visit the (slot, n) pairs that this thread owns in the last stage of `pl`
// (the same pairs `dst` receives), in slot order.
template <int R, class F>
KW_BODY void last_slots_r(const FftPlan& pl, int g, int G, bool active, F& f) {
  constexpr int MB = MaxB<R>::v;
  const int N = pl.N;
  const int nb = N / R;
  const int Ns = N / R;  // product of the earlier radices
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = g + m * G;
    if (active && b < nb) {
      const int k = b % Ns;
      const int j0 = (b - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) f(m * R + r, j0 + r * Ns);
    }
  }
}

template <class F>
KW_BODY void last_slots(const FftPlan& pl, int g, int G, bool active, F& f) {
  switch (pl.radix[pl.nst - 1]) {
    case 2: last_slots_r<2>(pl, g, G, active, f); break;
    case 3: last_slots_r<3>(pl, g, G, active, f); break;
    case 4: last_slots_r<4>(pl, g, G, active, f); break;
    case 5: last_slots_r<5>(pl, g, G, active, f); break;
    default: last_slots_r<8>(pl, g, G, active, f); break;
  }
}

}  // namespace kw
