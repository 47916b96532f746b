// Device-side description of one simulation ("problem") and the argument
// blocks of the pass kernels.
//
// Layout in HBM (all column-major, x fastest, as in the MATLAB reference):
//   real grid     : cell = x + Nx*(y + Ny*z), Nx*Ny*Nz floats
//   half spectrum : one allocation of nspec complex values.  The first
//                   Wm*Ny*Nz entries ("H") hold kx = 0 .. Wm-1 at
//                   kx + Wm*(y + Ny*z) with Wm = floor(Nx/2) for even Nx and
//                   (Nx+1)/2 for odd Nx.  For even Nx the Nyquist column
//                   kx = Nx/2 follows as a separate Ny*Nz plane ("Q") at
//                   Wm*Ny*Nz + (y + Ny*z), so that the strided passes work on
//                   a power-of-two-friendly width with no ragged tile.
//   spectral operator tables (kappa, nabla1, nabla2, filters) share the
//   half-spectrum layout and already include the 1/(Nx*Ny*Nz) of the
//   inverse transform.
#pragma once

#include "kw_fft.h"

namespace kw {

enum : uint32_t {
  REC_P = 1u << 0,
  REC_P_MAX = 1u << 1,
  REC_P_MIN = 1u << 2,
  REC_P_RMS = 1u << 3,
  REC_P_MAX_ALL = 1u << 4,
  REC_P_MIN_ALL = 1u << 5,
  REC_P_FINAL = 1u << 6,
  REC_U = 1u << 7,
  REC_U_MAX = 1u << 8,
  REC_U_MIN = 1u << 9,
  REC_U_RMS = 1u << 10,
  REC_U_MAX_ALL = 1u << 11,
  REC_U_MIN_ALL = 1u << 12,
  REC_U_FINAL = 1u << 13,
  REC_U_NON_STAGGERED = 1u << 14,
};

enum : int32_t { EOS_LOSSLESS = 0, EOS_ABSORBING = 1, EOS_STOKES = 2 };
enum : int32_t { SRC_DIRICHLET = 0, SRC_ADD_NO_CORR = 1, SRC_ADDITIVE = 2 };

// a grid quantity that is either a full array (mask = all ones) or a
// scalar held in a one-element array (mask = 0); the load is unconditional
struct Field {
  const float* a;
  uint32_t mask;
  KW_BODY float at(uint32_t i) const { return KW_LD32(a, i & mask); }
  KW_BODY bool grid() const { return mask != 0; }
};

template <bool B>
struct BoolC {
  static constexpr bool value = B;
};
// run f(BoolC<b>) with the flag as a compile-time constant
template <class F>
KW_BODY void with_flag(bool b, F&& f) {
  if (b)
    f(BoolC<true>{});
  else
    f(BoolC<false>{});
}
template <int V>
struct IntC {
  static constexpr int value = V;
};

// a time-varying source on a set of grid points
struct TimeSource {
  int32_t len;        // number of time points (the reference's *_source_flag); 0 = off
  int32_t mode;       // SRC_*
  int32_t nseries;    // rows of the signal matrix
  int32_t npts;       // source points
  const float* sig;   // [nseries x len], (i, t) at i + t*nseries, already scaled
  const int32_t* row; // per point: signal row; one entry 0 with rowmask 0 for
                      // a single series
  int32_t rowmask;
  const int32_t* map; // per cell: source point index or -1 (dirichlet / no-correction)
  const float* km;    // additive, single series: real(ifftn(source_kappa.*fftn(mask)))
  const float* grid;  // additive, several series: per-step filtered source grid
  KW_BODY float value(int32_t pt, int32_t t0) const {
    const int32_t r = KW_LD32(row, (uint32_t)(pt & rowmask));
    return KW_LD32(sig, (uint32_t)r + (uint32_t)t0 * (uint32_t)nseries);
  }
};

struct ProbDev {
  // geometry
  int32_t dim, Nx, Ny, Nz;
  int32_t Wm, nyq;  // main half width, has Nyquist plane
  int64_t ncell, nrows, nspec_main, nspec;
  int32_t Nt;
  // FFT plans per axis and tiling (threads per pencil G, pencils per block P)
  FftPlan fx, fy, fz;
  int32_t Gx, Px, Gy, Py, Gl, Pl;  // x pass, strided y pass (3D), last-axis pass
  // scalars
  float dt;
  int32_t eos, nonlinear;
  // medium
  Field c0, rho0, rinv[3], bona, tau, eta;
  // 1-D PML profiles (regular and staggered grid), length N of their axis
  const float* pml[3];
  const float* pml_sg[3];
  // 1-D shift-derivative operators ifftshift(i k exp(+-i k d/2)); the x
  // operators hold floor(Nx/2)+1 entries, y and z the full length
  const cf* dpos[3];
  const cf* dneg[3];
  // spectral tables in half-spectrum layout (include 1/N)
  const float* kappa;
  const float* nabla1;
  const float* nabla2;
  // fields
  float* p;
  float* u[3];
  float* rho[3];
  float* rsum;  // sum of the split densities (absorbing EOS input)
  cf* S[4];     // spectral work buffers; S[0] carries fft(p) between steps;
                // S[1..dim] are distinct when the velocity split is recorded
  // sources
  TimeSource srcp;
  TimeSource srcu[3];
  // sensor
  uint32_t record;
  int32_t record_start;  // 1-based
  int32_t n_rec;
  int32_t full_mask;     // sensor mask covers every cell in natural order
  int32_t nsens;
  const int32_t* sens_idx;
  // Cartesian points interpolated linearly: per point the dim+1 simplex
  // vertices (0-based cells) and weights, point-major ([nsens x (dim+1)]);
  // nullptr for binary masks
  const int32_t* sens_tri;
  const float* sens_bc;
  // 2-D directivity: the unique-angle number of each sensor point and the
  // pressure filtered for one angle (scratch grid)
  const int32_t* dir_id;
  float* dirp;
  float* p_ts;           // [nsens x n_rec]
  double *p_max, *p_min, *p_rms2;  // [nsens]; p_rms2 accumulates p^2
  double *p_max_all, *p_min_all;   // [ncell]
  float* u_ts[3];
  double *u_max[3], *u_min[3], *u_rms2[3];
  double *u_max_all[3], *u_min_all[3];
  // velocity on the non-staggered grid: 1-D shift operators per axis,
  // ifftshift(exp(-i k d/2)) with the 1/(Nx Ny Nz) of the round trip and the
  // real part at the Nyquist bin (what real(ifft(.)) keeps of it); the x
  // operator holds floor(Nx/2)+1 entries, y and z the full length
  const cf* ushift[3];
  float* uns;           // [ncell] shifted component (scratch)
  float* uns_ts[3];     // [nsens x n_rec]
  // velocity split into compressional and shear parts: tables of
  // k_d / |k| (0 at k = 0) in the last-axis pass's order, the shift
  // operators with their value at the Nyquist bin, and 1/(Nx Ny Nz)
  const float* khat[3];
  const cf* sshift[3];
  float* usp;           // [ncell] compressional part of one component (scratch)
  float* usplit_p[3];   // [nsens x n_rec]
  float* usplit_s[3];
  float* rho_ts;        // [nsens x n_rec] summed split density (KW_REC_RHO)
  // time reversal
  int32_t time_rev;
  const int32_t* tr_map;  // per cell: sensor point or -1
  const float* tr_data;   // [nsens x Nt], original (not reversed) column order
  // step number on the device: the number of steps done until the first
  // pass of a step advances it to the step being computed (1-based), so
  // kernels after that pass read the current step and the kernels before it
  // (the per-step source grids) read the step before
  int32_t* t_index;
  // per-step decisions taken on the device from t_index, so that one launch
  // can serve problems that are at different steps: p is stored to HBM at
  // steps t >= p_store_from (a sparse sensor reads it) and at t == p_store_at
  // (p_final); rec_points / rec_grid are the statistics the sensor kernels
  // update after the first step
  int32_t p_store_from;
  int32_t p_store_at;
  uint32_t rec_points;
  uint32_t rec_grid;
  // first step at which the pressure held a non-finite value (0: none),
  // written by the kernels that form p; device memory, or pinned host memory
  // the engine reads after each round without a copy
  int32_t* bad;
};

// per-launch mapping of blocks to problems: problem i owns blocks
// [tile_start[i], tile_start[i+1])
struct Batch {
  const ProbDev* probs;
  const int32_t* tile_start;
  int32_t nprob;
};

template <class Ctx>
KW_BODY int find_prob(const Batch& B, int blk) {
  int lo = 0, hi = B.nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (KW_LD(B.tile_start + mid) <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Pass kinds (launch granularity).  Every kind takes a Batch plus a small
// argument block; the kernels read everything else from ProbDev.
enum PassKind : int32_t {
  PK_YFWD = 0,      // strided forward FFT along y (3D), in place on S[]
  PK_YINV = 1,      // strided inverse FFT along y (3D)
  PK_SPEC = 2,      // last axis: forward, spectral multiply, inverse
  PK_X_R2C = 3,     // real field -> half spectrum
  PK_X_C2R = 4,     // half spectrum -> real field
  PK_X_GRAD = 5,    // end of the gradient: u update (+source) and fft(u)
  PK_X_DIV = 6,     // end of the divergence: rho update (+source), EOS or fft(a,b)
  PK_X_EOS = 7,     // end of the absorbing EOS: p, sensors, fft(p)
};

enum SpecMode : int32_t {
  SM_GRAD = 0,   // S[0] -> S[0..dim-1] = D+_d kappa S[0]
  SM_DIV = 1,    // S[d] -> D-_d kappa S[d], d < dim
  SM_ABS = 2,    // S[0] -> nabla1 S[0], S[1] -> nabla2 S[1]
  SM_FILTER = 3, // S[buf] -> table .* S[buf]
  SM_SHIFT = 4,  // S[buf] -> ushift_d .* S[buf], d = PassArgs.flags
  SM_SPLIT = 5,  // S[buf[a]] -> khat_a' .* sum_b khat_b sshift_b S[buf[b]]
};

struct PassArgs {
  int32_t mode;     // SpecMode for PK_SPEC; sub-mode for x passes
  int32_t nbuf;     // number of spectral buffers touched (y passes, r2c/c2r)
  int32_t buf[3];   // spectral buffer indices
  float* real[3];   // real grids for r2c/c2r (per problem when nprob == 1)
  const float* table;  // filter table for SM_FILTER (and SM_SHIFT), a row
                       // factor for x_r2c (nprob == 1)
  int32_t flags;
  // set on the first pass of every time step (the gradient's y pass in 3D,
  // its last-axis pass in 2D), which reads no step number: one thread per
  // problem advances t_index to the step being computed
  int32_t advance;
  // y passes over several buffers: when > 0, the launch has nbuf times this
  // many blocks and block k transforms buffer k / ysplit only
  int32_t ysplit;
};

enum : int32_t {
  XG_INIT = 1,     // p0 initial condition: u = dt/2 * rinv * grad p, no fft(u)
  XC_ACCUM = 1,    // c2r: add into the real grid instead of overwriting
  X_STORE_P = 2,   // producer of p must store p to HBM this step
};

}  // namespace kw
