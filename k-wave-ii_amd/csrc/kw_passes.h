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
#pragma once

#include "kw_prob.h"

namespace kw {

// ---------------------------------------------------------------------------
// half-spectrum access for the x passes ("two-for-one": two real rows are
// transformed as one complex pencil z = a + i b)
// ---------------------------------------------------------------------------

// element n (0 <= n < Nx) of the Hermitian extension of row `row`, with the
// imaginary parts of the DC and Nyquist bins dropped, which is what
// real(ifft(.)) does with them
KW_BODY cf half_get(const ProbDev& P, const cf* S, int64_t row, int n) {
  const int N = P.Nx;
  cf v;
  if (n < P.Wm) {
    v = KW_LD(S + row * P.Wm + n);
    if (n == 0) v.y = 0.f;
  } else if (P.nyq && n == (N >> 1)) {
    v = KW_LD(S + P.nspec_main + row);
    v.y = 0.f;
  } else {
    v = cconj(KW_LD(S + row * P.Wm + (N - n)));
  }
  return v;
}

KW_BODY void half_put(const ProbDev& P, cf* S, int64_t row, int k, cf v) {
  if (k < P.Wm)
    KW_ST(S + row * P.Wm + k, v);
  else
    KW_ST(S + P.nspec_main + row, v);
}

struct XTile {
  int g, G;
  bool active, has1;
  int64_t r0, r1;
  LdsPencil lds;
};

template <class Ctx>
KW_BODY XTile x_tile(Ctx& c, const ProbDev& P, int tile) {
  XTile T;
  T.G = P.Gx;
  const int tid = c.tid();
  int p = tid / T.G;
  T.g = tid % T.G;
  const int64_t npairs = (P.nrows + 1) / 2;
  const int64_t pair = (int64_t)tile * P.Px + p;
  T.active = p < P.Px && pair < npairs;
  if (p >= P.Px) p = 0;
  T.r0 = 2 * pair;
  T.r1 = 2 * pair + 1;
  T.has1 = T.r1 < P.nrows;
  T.lds = LdsPencil{c.lds() + (int64_t)p * (P.Nx + 1), 1};
  return T;
}

// LDS holds Z = fft(a + i b) in natural order; split into the half spectra
// of a (row r0) and b (row r1) and store them.
template <class Ctx>
KW_BODY void x_split_store(Ctx& c, const ProbDev& P, const XTile& T, cf* S) {
  const int N = P.Nx;
  const int K = N / 2 + 1;
  if (T.active) {
    for (int k = T.g; k < K; k += T.G) {
      const cf zk = T.lds.get(k);
      const cf zm = T.lds.get(k == 0 ? 0 : N - k);
      const cf A = cf{0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)};
      const cf B = cf{0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x)};
      half_put(P, S, T.r0, k, A);
      if (T.has1) half_put(P, S, T.r1, k, B);
    }
  }
  c.sync();
}

// forward transform of the complex pencil already in LDS, then split/store
template <class Ctx>
KW_BODY void x_fwd_from_lds(Ctx& c, const ProbDev& P, const XTile& T, cf* S) {
  auto src = [&](int n) -> cf { return T.lds.get(n); };
  auto dst = [&](int n, cf v, int) { T.lds.set(n, v); };
  fft<false>(c, P.fx, T.g, T.G, T.active, src, dst, T.lds);
  x_split_store(c, P, T, S);
}

// inverse two-for-one load functor: Z[n] = A[n] + i B[n]
struct HalfPairLoad {
  const ProbDev* P;
  const cf* S;
  int64_t r0, r1;
  bool has1;
  KW_BODY cf operator()(int n) const {
    const cf a = half_get(*P, S, r0, n);
    const cf b = has1 ? half_get(*P, S, r1, n) : cf{0.f, 0.f};
    return cf{a.x - b.y, a.y + b.x};
  }
};

KW_BODY float pml_row(const float* v, const ProbDev& P, int d, int64_t row) {
  // value of the 1-D profile of axis d (d = 1: y, d = 2: z) for a row
  const int64_t y = row % P.Ny;
  const int64_t z = row / P.Ny;
  return KW_LD(v + (d == 1 ? y : z));
}

// time-reversal enforcement, sensors and optional store for a new p value
KW_BODY float finish_p(const ProbDev& P, int64_t cell, float p, int t, bool store) {
  if (P.time_rev) {
    const int32_t i = KW_LD(P.tr_map + cell);
    if (i >= 0) {
      // data of step t+1 of the reversed series = original column Nt - t
      p = KW_LD(P.tr_data + (int64_t)i + (int64_t)(P.Nt - t - 1) * P.nsens);
      const float c0 = P.c0.at(cell);
      const float c2 = c0 * c0;
      float r;
      if (P.dim == 3)
        r = p / (3.f * c2);
      else
        r = (0.5f * p) / c2;
      for (int d = 0; d < P.dim; ++d) KW_ST(P.rho[d] + cell, r);
    }
  }
  if (!P.time_rev && t >= P.record_start) {
    const int fi = t - P.record_start + 1;
    const double pd = (double)p;
    const uint32_t rec = P.record;
    if (P.full_mask) {
      // sensor mask == whole grid: point statistics fused here
      if (rec & REC_P) KW_ST(P.p_ts + cell + (int64_t)(fi - 1) * P.ncell, p);
      if (rec & REC_P_MAX) {
        double m = fi == 1 ? pd : KW_LD(P.p_max + cell);
        KW_ST(P.p_max + cell, pd > m ? pd : m);
      }
      if (rec & REC_P_MIN) {
        double m = fi == 1 ? pd : KW_LD(P.p_min + cell);
        KW_ST(P.p_min + cell, pd < m ? pd : m);
      }
      if (rec & REC_P_RMS) {
        double s = fi == 1 ? 0.0 : KW_LD(P.p_rms2 + cell);
        KW_ST(P.p_rms2 + cell, s + pd * pd);
      }
    }
    if (rec & REC_P_MAX_ALL) {
      double m = fi == 1 ? pd : KW_LD(P.p_max_all + cell);
      KW_ST(P.p_max_all + cell, pd > m ? pd : m);
    }
    if (rec & REC_P_MIN_ALL) {
      double m = fi == 1 ? pd : KW_LD(P.p_min_all + cell);
      KW_ST(P.p_min_all + cell, pd < m ? pd : m);
    }
  }
  if (store) KW_ST(P.p + cell, p);
  return p;
}

// value added to / imposed on a field by a time source at one cell
// returns the new value
KW_BODY float apply_src(const TimeSource& s, int64_t cell, float v, int t) {
  const int t0 = t - 1;
  if (s.mode == SRC_ADDITIVE) {
    if (s.km) return v + s.value(0, t0) * KW_LD(s.km + cell);
    return v + KW_LD(s.grid + cell);
  }
  const int32_t i = KW_LD(s.map + cell);
  if (i < 0) return v;
  const float sv = s.value(i, t0);
  return s.mode == SRC_DIRICHLET ? sv : v + sv;
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
template <bool INV, class Ctx>
KW_BODY void body_ypass(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const int N = P.Ny, G = P.Gy, Pn = P.Py;
  const int64_t nH = (int64_t)P.Wm * P.Nz;
  const int64_t tH = (nH + Pn - 1) / Pn;
  const int tid = c.tid();
  int p, g;
  int64_t base, stride;
  bool active;
  LdsPencil lds;
  if (tile < tH) {
    p = tid % Pn;
    g = tid / Pn;
    const int64_t q = (int64_t)tile * Pn + p;
    active = q < nH;
    base = (q / P.Wm) * ((int64_t)P.Wm * P.Ny) + (q % P.Wm);
    stride = P.Wm;
    lds = LdsPencil{c.lds() + p, Pn};
  } else {
    p = tid / G;
    g = tid % G;
    const int64_t q = (tile - tH) * Pn + p;
    active = p < Pn && q < P.Nz;
    if (p >= Pn) p = 0;
    base = P.nspec_main + q * P.Ny;
    stride = 1;
    lds = LdsPencil{c.lds() + (int64_t)p * (N + 1), 1};
  }
  for (int b = 0; b < A.nbuf; ++b) {
    cf* S = P.S[A.buf[b]];
    auto ld = [&](int n) -> cf { return KW_LD(S + base + n * stride); };
    auto st = [&](int n, cf v, int) { KW_ST(S + base + n * stride, v); };
    fft<INV>(c, P.fy, g, G, active, ld, st, lds);
  }
}

// ---------------------------------------------------------------------------
// last-axis pass (z in 3D, y in 2D): forward FFT, spectral operator, inverse
// FFT, all in LDS.  Operators (kspaceFirstOrder3D.m:629-661, 792-803,
// 907-909, 977-978; createAbsorptionVariables.m:29-84).
// ---------------------------------------------------------------------------
template <class Ctx>
KW_BODY void body_spec(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const bool d3 = P.dim == 3;
  const FftPlan& pl = d3 ? P.fz : P.fy;
  const int N = pl.N, G = P.Gl, Pn = P.Pl;
  const int last = d3 ? 2 : 1;
  const int64_t nH = d3 ? (int64_t)P.Wm * P.Ny : (int64_t)P.Wm;
  const int64_t tH = (nH + Pn - 1) / Pn;
  const int tid = c.tid();
  const int p = tid % Pn;
  const int g = tid / Pn;
  int64_t q, base, stride;
  int kx, ky;  // ky = -1: ky is the pencil coordinate (2D)
  bool active;
  if (tile < tH) {
    q = (int64_t)tile * Pn + p;
    active = q < nH;
    base = q;
    stride = d3 ? (int64_t)P.Wm * P.Ny : (int64_t)P.Wm;
    kx = (int)(q % P.Wm);
    ky = d3 ? (int)(q / P.Wm) : -1;
  } else {
    q = (tile - tH) * Pn + p;
    active = q < (d3 ? P.Ny : 1);
    base = P.nspec_main + q;
    stride = d3 ? P.Ny : 1;
    kx = P.Nx / 2;
    ky = d3 ? (int)q : -1;
  }
  if (!active) {
    kx = 0;
    ky = d3 ? 0 : -1;
  }
  LdsPencil la{c.lds() + p, Pn};
  LdsPencil lb{c.lds() + (int64_t)Pn * N + p, Pn};
  auto addr = [&](int n) -> int64_t { return base + n * stride; };
  auto to_a = [&](int n, cf v, int) { la.set(n, v); };

  if (A.mode == SM_GRAD) {
    cf* S0 = P.S[0];
    auto ld = [&](int n) -> cf { return KW_LD(S0 + addr(n)); };
    fft<false>(c, pl, g, G, active, ld, to_a, la);
    // C = ifft(kappa .* S); x (and in 3D y) derivatives are per-pencil
    // constants times C
    const cf dx = active ? KW_LD(P.dpos[0] + kx) : cf{0.f, 0.f};
    const cf dy = (active && d3) ? KW_LD(P.dpos[1] + ky) : cf{0.f, 0.f};
    auto src1 = [&](int n) -> cf { return cscale(la.get(n), KW_LD(P.kappa + addr(n))); };
    auto st1 = [&](int n, cf v, int) {
      KW_ST(P.S[0] + addr(n), cmul(dx, v));
      if (d3) KW_ST(P.S[1] + addr(n), cmul(dy, v));
    };
    fft<true>(c, pl, g, G, active, src1, st1, lb);
    // derivative along the pencil axis
    const cf* dl = P.dpos[last];
    auto src2 = [&](int n) -> cf {
      return cmul(KW_LD(dl + n), cscale(la.get(n), KW_LD(P.kappa + addr(n))));
    };
    cf* So = P.S[last];
    auto st2 = [&](int n, cf v, int) { KW_ST(So + addr(n), v); };
    fft<true>(c, pl, g, G, active, src2, st2, lb);
  } else if (A.mode == SM_DIV) {
    for (int d = 0; d < P.dim; ++d) {
      cf* S = P.S[d];
      auto ld = [&](int n) -> cf { return KW_LD(S + addr(n)); };
      fft<false>(c, pl, g, G, active, ld, to_a, la);
      const cf dc = (d == last || !active) ? cf{0.f, 0.f}
                                           : KW_LD(P.dneg[d] + (d == 0 ? kx : ky));
      const cf* dl = P.dneg[d];
      const bool along = d == last;
      auto src = [&](int n) -> cf {
        const cf s = cscale(la.get(n), KW_LD(P.kappa + addr(n)));
        return cmul(along ? KW_LD(dl + n) : dc, s);
      };
      auto st = [&](int n, cf v, int) { KW_ST(S + addr(n), v); };
      fft<true>(c, pl, g, G, active, src, st, la);
    }
  } else {
    // SM_ABS: S0 <- nabla1 .* S0, S1 <- nabla2 .* S1;  SM_FILTER: one table
    const int nb = A.mode == SM_ABS ? 2 : 1;
    for (int j = 0; j < nb; ++j) {
      cf* S = P.S[A.mode == SM_ABS ? j : A.buf[0]];
      const float* tab = A.mode == SM_ABS ? (j == 0 ? P.nabla1 : P.nabla2) : A.table;
      auto ld = [&](int n) -> cf { return KW_LD(S + addr(n)); };
      fft<false>(c, pl, g, G, active, ld, to_a, la);
      auto src = [&](int n) -> cf { return cscale(la.get(n), KW_LD(tab + addr(n))); };
      auto st = [&](int n, cf v, int) { KW_ST(S + addr(n), v); };
      fft<true>(c, pl, g, G, active, src, st, la);
    }
  }
}

// ---------------------------------------------------------------------------
// x passes
// ---------------------------------------------------------------------------

// real grid(s) -> half spectrum (PassArgs.real[b] -> S[buf[b]])
template <class Ctx>
KW_BODY void body_x_r2c(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const XTile T = x_tile(c, P, tile);
  const int N = P.Nx;
  for (int b = 0; b < A.nbuf; ++b) {
    const float* f = A.real[b] ? A.real[b] : P.p;
    auto src = [&](int n) -> cf {
      const float a = KW_LD(f + T.r0 * N + n);
      const float bb = T.has1 ? KW_LD(f + T.r1 * N + n) : 0.f;
      return cf{a, bb};
    };
    auto dst = [&](int n, cf v, int) { T.lds.set(n, v); };
    fft<false>(c, P.fx, T.g, T.G, T.active, src, dst, T.lds);
    x_split_store(c, P, T, P.S[A.buf[b]]);
  }
}

// half spectrum -> real grid (S[buf[b]] -> PassArgs.real[b]); XC_ACCUM adds
template <class Ctx>
KW_BODY void body_x_c2r(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const XTile T = x_tile(c, P, tile);
  const int N = P.Nx;
  const bool acc = A.flags & XC_ACCUM;
  for (int b = 0; b < A.nbuf; ++b) {
    float* f = A.real[b];
    HalfPairLoad src{&P, P.S[A.buf[b]], T.r0, T.r1, T.has1};
    auto dst = [&](int n, cf v, int) {
      float* f0 = f + T.r0 * N + n;
      KW_ST(f0, acc ? KW_LD(f0) + v.x : v.x);
      if (T.has1) {
        float* f1 = f + T.r1 * N + n;
        KW_ST(f1, acc ? KW_LD(f1) + v.y : v.y);
      }
    };
    fft<true>(c, P.fx, T.g, T.G, T.active, src, dst, T.lds);
  }
}

// end of the gradient: u_d = pml_sg_d (pml_sg_d u_d - dt rho0_sg_d^-1 g_d)
// (kspaceFirstOrder3D.m:792-803), velocity sources (:806-886), then
// S[d] = fft_x(u_d) for the divergence.  XG_INIT: p0 initial condition
// u_d = dt rho0_sg_d^-1 g_d / 2 (:1023-1037), no fft.
template <class Ctx>
KW_BODY void body_x_grad(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const XTile T = x_tile(c, P, tile);
  const int N = P.Nx;
  const int t = KW_LD(P.t_index);
  const bool init = A.flags & XG_INIT;
  const float dt = P.dt;
  for (int d = 0; d < P.dim; ++d) {
    cf* S = P.S[d];
    float* u = P.u[d];
    const float* pv = P.pml_sg[d];
    float pr0 = 1.f, pr1 = 1.f;
    if (d > 0 && T.active) {
      pr0 = pml_row(pv, P, d, T.r0);
      if (T.has1) pr1 = pml_row(pv, P, d, T.r1);
    }
    const TimeSource& s = P.srcu[d];
    const bool src_on = !init && s.len >= t;
    const Field ri = P.rinv[d];
    auto upd = [&](int64_t row, int n, float g, float pr) -> float {
      const int64_t cell = row * N + n;
      if (init) {
        const float un = dt * ri.at(cell) * g / 2.f;
        KW_ST(u + cell, un);
        return un;
      }
      const float pm = d == 0 ? KW_LD(pv + n) : pr;
      float un = pm * (pm * KW_LD(u + cell) - dt * ri.at(cell) * g);
      if (src_on) un = apply_src(s, cell, un, t);
      KW_ST(u + cell, un);
      return un;
    };
    HalfPairLoad src{&P, S, T.r0, T.r1, T.has1};
    auto dst = [&](int n, cf v, int) {
      const float a = upd(T.r0, n, v.x, pr0);
      const float b = T.has1 ? upd(T.r1, n, v.y, pr1) : 0.f;
      T.lds.set(n, cf{a, b});
    };
    fft<true>(c, P.fx, T.g, T.G, T.active, src, dst, T.lds);
    if (!init) x_fwd_from_lds(c, P, T, S);
  }
}

// equation of state for one cell (kspaceFirstOrder3D.m:965-1019); `ab` are
// the two absorbing terms (tau-term input A and eta-term input B) already
// transformed back; `sdu` the velocity divergence sum (stokes)
KW_BODY float eos_p(const ProbDev& P, int64_t cell, float rs, float sdu, float Aterm,
                    float Bterm) {
  const float c0 = P.c0.at(cell);
  float inner;
  if (P.eos == EOS_ABSORBING) {
    inner = (rs + P.tau.at(cell) * Aterm) - P.eta.at(cell) * Bterm;
  } else if (P.eos == EOS_STOKES) {
    inner = rs + P.tau.at(cell) * P.rho0.at(cell) * sdu;
  } else {
    inner = rs;
  }
  if (P.nonlinear) inner = inner + P.bona.at(cell) * (rs * rs) / (2.f * P.rho0.at(cell));
  return (c0 * c0) * inner;
}

// end of the divergence: rho_d update (:912-927), pressure source
// (:930-962); then either the lossless/stokes EOS with sensors and fft_x(p)
// into S[0], or (absorbing) S[0] = fft_x(rho0 * sum du), S[1] = fft_x(sum rho)
// and rsum = sum rho
template <class Ctx>
KW_BODY void body_x_div(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const XTile T = x_tile(c, P, tile);
  const int N = P.Nx;
  const int t = KW_LD(P.t_index);
  const float dt = P.dt;
  const bool psrc = P.srcp.len >= t;
  const bool store_p = A.flags & X_STORE_P;
  cf sdu[kMaxSlots], srho[kMaxSlots], rpr[kMaxSlots];
#pragma unroll
  for (int i = 0; i < kMaxSlots; ++i) {
    sdu[i] = cf{0.f, 0.f};
    srho[i] = cf{0.f, 0.f};
    rpr[i] = cf{0.f, 0.f};
  }
  for (int d = 0; d < P.dim; ++d) {
    cf* S = P.S[d];
    float* rho = P.rho[d];
    const float* pv = P.pml[d];
    float pr0 = 1.f, pr1 = 1.f;
    if (d > 0 && T.active) {
      pr0 = pml_row(pv, P, d, T.r0);
      if (T.has1) pr1 = pml_row(pv, P, d, T.r1);
    }
    auto upd = [&](int64_t row, int n, float g, float pr, float& acc_du, float& acc_r,
                   float& rp) -> void {
      const int64_t cell = row * N + n;
      if (d == 0 && P.nonlinear) {
        float s = KW_LD(P.rho[0] + cell) + KW_LD(P.rho[1] + cell);
        if (P.dim == 3) s = s + KW_LD(P.rho[2] + cell);
        rp = 2.f * s + P.rho0.at(cell);
      }
      const float pm = d == 0 ? KW_LD(pv + n) : pr;
      const float coef = P.nonlinear ? rp : P.rho0.at(cell);
      float rn = pm * (pm * KW_LD(rho + cell) - dt * coef * g);
      if (psrc) rn = apply_src(P.srcp, cell, rn, t);
      KW_ST(rho + cell, rn);
      acc_du = acc_du + g;
      acc_r = acc_r + rn;
    };
    HalfPairLoad src{&P, S, T.r0, T.r1, T.has1};
    auto dst = [&](int n, cf v, int slot) {
      upd(T.r0, n, v.x, pr0, sdu[slot].x, srho[slot].x, rpr[slot].x);
      if (T.has1) upd(T.r1, n, v.y, pr1, sdu[slot].y, srho[slot].y, rpr[slot].y);
    };
    fft<true>(c, P.fx, T.g, T.G, T.active, src, dst, T.lds);
  }
  if (P.eos == EOS_ABSORBING) {
    auto put_a = [&](int slot, int n) {
      const float a = P.rho0.at(T.r0 * N + n) * sdu[slot].x;
      const float b = T.has1 ? P.rho0.at(T.r1 * N + n) * sdu[slot].y : 0.f;
      T.lds.set(n, cf{a, b});
    };
    last_slots(P.fx, T.g, T.G, T.active, put_a);
    c.sync();
    x_fwd_from_lds(c, P, T, P.S[0]);
    auto put_b = [&](int slot, int n) {
      KW_ST(P.rsum + T.r0 * N + n, srho[slot].x);
      if (T.has1) KW_ST(P.rsum + T.r1 * N + n, srho[slot].y);
      T.lds.set(n, srho[slot]);
    };
    last_slots(P.fx, T.g, T.G, T.active, put_b);
    c.sync();
    x_fwd_from_lds(c, P, T, P.S[1]);
  } else {
    auto put_p = [&](int slot, int n) {
      const int64_t c0 = T.r0 * N + n;
      float a = eos_p(P, c0, srho[slot].x, sdu[slot].x, 0.f, 0.f);
      a = finish_p(P, c0, a, t, store_p);
      float b = 0.f;
      if (T.has1) {
        const int64_t c1 = T.r1 * N + n;
        b = eos_p(P, c1, srho[slot].y, sdu[slot].y, 0.f, 0.f);
        b = finish_p(P, c1, b, t, store_p);
      }
      T.lds.set(n, cf{a, b});
    };
    last_slots(P.fx, T.g, T.G, T.active, put_p);
    c.sync();
    x_fwd_from_lds(c, P, T, P.S[0]);
  }
}

// end of the absorbing EOS: p = c0^2 (rsum + tau A - eta B [+ BonA term])
// with A = ifft(nabla1 fft(rho0 sum du)), B = ifft(nabla2 fft(rsum)); sensors;
// S[0] = fft_x(p)
template <class Ctx>
KW_BODY void body_x_eos(Ctx& c, const Batch& B, const PassArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const XTile T = x_tile(c, P, tile);
  const int N = P.Nx;
  const int t = KW_LD(P.t_index);
  const bool store_p = A.flags & X_STORE_P;
  cf ta[kMaxSlots];
#pragma unroll
  for (int i = 0; i < kMaxSlots; ++i) ta[i] = cf{0.f, 0.f};
  {
    HalfPairLoad src{&P, P.S[0], T.r0, T.r1, T.has1};
    auto dst = [&](int n, cf v, int slot) { ta[slot] = v; };
    fft<true>(c, P.fx, T.g, T.G, T.active, src, dst, T.lds);
  }
  {
    HalfPairLoad src{&P, P.S[1], T.r0, T.r1, T.has1};
    auto dst = [&](int n, cf v, int slot) {
      const int64_t c0 = T.r0 * N + n;
      float a = eos_p(P, c0, KW_LD(P.rsum + c0), 0.f, ta[slot].x, v.x);
      a = finish_p(P, c0, a, t, store_p);
      float b = 0.f;
      if (T.has1) {
        const int64_t c1 = T.r1 * N + n;
        b = eos_p(P, c1, KW_LD(P.rsum + c1), 0.f, ta[slot].y, v.y);
        b = finish_p(P, c1, b, t, store_p);
      }
      T.lds.set(n, cf{a, b});
    };
    fft<true>(c, P.fx, T.g, T.G, T.active, src, dst, T.lds);
  }
  x_fwd_from_lds(c, P, T, P.S[0]);
}

// ---------------------------------------------------------------------------
// elementwise kernels (block of EW_T threads, EW_PER cells per thread)
// ---------------------------------------------------------------------------
constexpr int EW_T = 256;

enum EwKind : int32_t {
  EW_INIT_P0 = 0,    // p = p0, rho_d = p0 / (dim c0^2)
  EW_TR_PROLOGUE,    // time reversal: enforce the first data column
  EW_SRC_GRID,       // per-step scatter grid of a many-series additive source
  EW_RECORD_POINTS,  // sensor statistics at sensor points (general mask)
  EW_RECORD_GRID,    // whole-grid statistics (*_all) not fused elsewhere
  EW_ADVANCE,        // t_index += 1
  EW_FILL,           // zero a float buffer (graph-safe memset)
};

struct EwArgs {
  int32_t kind;
  int32_t which;         // EW_SRC_GRID: 0 = p source, 1..3 = u source d-1
  int64_t n;             // EW_FILL length (in floats)
  float* ptr;            // EW_FILL target
  const float* p0;       // EW_INIT_P0
  uint32_t rec;          // EW_RECORD_*: which statistics to update
};

template <class Ctx>
KW_BODY void body_ew(Ctx& c, const Batch& B, const EwArgs& A) {
  int tile;
  const ProbDev& P = batch_prob(c, B, tile);
  const int64_t i = (int64_t)tile * EW_T + c.tid();
  const int t = KW_LD(P.t_index);
  switch (A.kind) {
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
        const float r = P.dim == 3 ? p / (3.f * c2) : (0.5f * p) / c2;
        for (int d = 0; d < P.dim; ++d) KW_ST(P.rho[d] + i, r);
      }
      KW_ST(P.p + i, p);
      return;
    }
    case EW_SRC_GRID: {
      if (i >= P.ncell) return;
      const TimeSource& s = A.which == 0 ? P.srcp : P.srcu[A.which - 1];
      const int32_t k = KW_LD(s.map + i);
      const float v = (k >= 0 && s.len >= t) ? s.value(k, t - 1) : 0.f;
      KW_ST((float*)s.grid + i, v);
      return;
    }
    case EW_RECORD_POINTS: {
      if (i >= P.nsens || t < P.record_start) return;
      const int fi = t - P.record_start + 1;
      const int64_t cell = P.sens_idx ? (int64_t)KW_LD(P.sens_idx + i) : i;
      const uint32_t rec = A.rec;
      const float pv = KW_LD(P.p + cell);
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
        const float uv = KW_LD(P.u[d] + cell);
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
      const uint32_t rec = A.rec;
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
    case EW_ADVANCE: {
      if (tile == 0 && c.tid() == 0) KW_ST(P.t_index, t + 1);
      return;
    }
    case EW_FILL: {
      if (i < A.n) KW_ST(A.ptr + i, 0.f);
      return;
    }
  }
}

}  // namespace kw
