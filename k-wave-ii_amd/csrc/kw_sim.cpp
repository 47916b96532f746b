// Host side of kwhip: validates a kw_problem, builds the operators in double
// precision and rounds them once to single (the reference's
// 'DataCast','single' policy, kspaceFirstOrder_dataCast.m:31-268), lays the
// problem out in device memory, sequences the passes of each time step and
// implements the single-simulation C ABI declared in include/kwhip.h.
#include "kw_sim.h"

#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace kw {

// ---------------------------------------------------------------------------
// problem validation and set-up
// ---------------------------------------------------------------------------

static std::vector<float> to_float(const double* a, int64_t n) {
  std::vector<float> v(n);
  parallel_for(n, [&](int64_t i) { v[i] = (float)a[i]; });
  return v;
}

static Field scalar_field(Sim& S, float v) {
  return Field{S.upload(std::vector<float>{v}), 0u};
}

static Field make_field(Sim& S, const double* a, double s, int64_t n) {
  if (a) return Field{S.upload(to_float(a, n)), ~0u};
  return scalar_field(S, (float)s);
}

static void check_index(const uint64_t* idx, int64_t n, int64_t ncell, const char* what) {
  if (n > 0 && !idx) fail(KW_EINVAL, std::string(what) + ": index list is NULL");
  for (int64_t i = 0; i < n; ++i)
    if (idx[i] < 1 || (int64_t)idx[i] > ncell)
      fail(KW_EINVAL, std::string(what) + ": index out of range (indices are 1-based)");
}

static void setup_source(Sim& S, const kw_time_source& src, TimeSource& ts, const char* what) {
  ts = TimeSource{};
  if (src.flag <= 0) return;
  const int64_t ncell = S.G.ncell;
  if (src.mode < 0 || src.mode > 2) fail(KW_EINVAL, std::string(what) + ": unknown mode");
  if (src.num_points <= 0) fail(KW_EINVAL, std::string(what) + ": no source points");
  check_index(src.index, src.num_points, ncell, what);
  if (src.num_series < 1 || !src.input) fail(KW_EINVAL, std::string(what) + ": no signal");
  const bool labelled = src.sig_index != nullptr;
  if (!labelled && src.num_series != 1 && src.num_series != src.num_points)
    fail(KW_EINVAL, std::string(what) +
                        ": the number of time series must be 1 or the number of source points");
  if (labelled)
    for (int64_t i = 0; i < src.num_points; ++i)
      if (src.sig_index[i] < 1 || (int64_t)src.sig_index[i] > src.num_series)
        fail(KW_EINVAL, std::string(what) + ": signal label out of range");
  const int64_t len = std::min<int64_t>(src.flag, S.Nt);
  ts.len = (int32_t)src.flag;
  ts.mode = src.mode;
  ts.nseries = (int32_t)src.num_series;
  ts.npts = (int32_t)src.num_points;
  // signal matrix: only the first min(flag, Nt) columns are ever used
  std::vector<float> sig((size_t)src.num_series * len);
  for (int64_t k = 0; k < (int64_t)sig.size(); ++k) sig[k] = (float)src.input[k];
  ts.sig = S.upload(sig);
  if (src.num_series > 1) {
    std::vector<int32_t> row(src.num_points);
    for (int64_t i = 0; i < src.num_points; ++i)
      row[i] = labelled ? (int32_t)(src.sig_index[i] - 1) : (int32_t)i;
    ts.row = S.upload(row);
    ts.rowmask = -1;
  } else {
    ts.row = S.upload(std::vector<int32_t>{0});
    ts.rowmask = 0;
  }
  const bool need_map = src.mode != SRC_ADDITIVE || src.num_series > 1;
  if (need_map) {
    std::vector<int32_t> map(ncell, -1);
    for (int64_t i = 0; i < src.num_points; ++i) map[src.index[i] - 1] = (int32_t)i;
    ts.map = S.upload(map);
  }
}

// K_M = real(ifftn(source_kappa .* fftn(M))) for the source mask M: the
// additive correction of a single-series source is s(t) K_M
// (kspaceFirstOrder3D.m:930-962 is linear in the signal)
static float* source_km(Sim& S, const kw_time_source& src) {
  std::vector<float> mask(S.G.ncell, 0.f);
  for (int64_t i = 0; i < src.num_points; ++i) mask[src.index[i] - 1] = 1.f;
  float* km = S.upload(mask);
  PassArgs a{};
  a.nbuf = 1;
  a.buf[0] = 1;
  a.real[0] = km;
  S.pass(PK_X_R2C, a);
  if (S.G.dim == 3) S.pass(PK_YFWD, a);
  PassArgs f = a;
  f.mode = SM_FILTER;
  f.table = S.skappa;
  S.pass(PK_SPEC, f);
  if (S.G.dim == 3) S.pass(PK_YINV, a);
  S.pass(PK_X_C2R, a);
  return km;
}

std::unique_ptr<Sim> create_sim(const kw_problem* pr_in, void* stream, Arena* arena) {
  if (!pr_in) fail(KW_EINVAL, "problem is NULL");
  const int pdim = pr_in->dim;
  if (pdim < 1 || pdim > 3) fail(KW_EINVAL, "dim must be 1, 2 or 3");
  // A 1-D problem (kspaceFirstOrder1D.m) runs as a 2-D grid of one row:
  // with Ny = 1 every y wavenumber is 0, so the y derivatives vanish, uy and
  // rho_y stay zero, and the 2-D update is the 1-D one; the kernels see
  // P.dim = 1 where the density enters as a single component (p0, time
  // reversal) and in the sensor records.
  kw_problem row = *pr_in;
  const kw_problem* pr = pr_in;
  // cuboid corners become the index list of their points
  std::vector<uint64_t> cuboid_idx;
  if (pr_in->sensor_num_cuboids < 0) fail(KW_EINVAL, "negative number of sensor cuboids");
  if (pr_in->sensor_mask_corners && pr_in->sensor_num_cuboids > 0) {
    if (pr_in->sensor_mask_index) fail(KW_EINVAL, "give sensor_mask_index or sensor_mask_corners");
    // the reference takes cuboid corners for forward simulations with
    // uninterpolated records only (kspaceFirstOrder_inputChecking.m:516-576)
    if (pr_in->time_reversal_boundary_data)
      fail(KW_EINVAL, "time reversal takes a binary sensor mask, not cuboid corners");
    if (pr_in->sensor_tri || pr_in->sensor_bc)
      fail(KW_EINVAL, "cuboid corners cannot be combined with interpolated Cartesian points");
    if (pr_in->directivity_angle)
      fail(KW_EINVAL, "sensor directivity needs a binary sensor mask");
    const int64_t Nd[3] = {pr_in->Nx, pdim >= 2 ? pr_in->Ny : 1, pdim == 3 ? pr_in->Nz : 1};
    const uint64_t* c = pr_in->sensor_mask_corners;
    // check every cuboid and count its points before enumerating any, so a
    // corner list that does not match sensor_num_points fails without a
    // large host allocation
    int64_t total = 0;
    for (int64_t q = 0; q < pr_in->sensor_num_cuboids; ++q) {
      int64_t cnt = 1;
      for (int a = 0; a < pdim; ++a) {
        const int64_t lo = (int64_t)c[q * 2 * pdim + a];
        const int64_t hi = (int64_t)c[q * 2 * pdim + pdim + a];
        if (hi < lo) fail(KW_EINVAL, "sensor.mask cuboid corners must be defined as [x1, ..., x2, ...].' where x2 => x1, etc.");
        if (lo < 1 || hi > Nd[a]) fail(KW_EINVAL, "sensor.mask cuboid corners must be within the grid.");
        cnt *= hi - lo + 1;
      }
      total += cnt;
      if (total > pr_in->sensor_num_points)
        fail(KW_EINVAL, "sensor_num_points must equal the number of points of the cuboids");
    }
    if (total != pr_in->sensor_num_points)
      fail(KW_EINVAL, "sensor_num_points must equal the number of points of the cuboids");
    cuboid_idx.reserve((size_t)total);
    for (int64_t q = 0; q < pr_in->sensor_num_cuboids; ++q) {
      int64_t lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
      for (int a = 0; a < pdim; ++a) {
        lo[a] = (int64_t)c[q * 2 * pdim + a];
        hi[a] = (int64_t)c[q * 2 * pdim + pdim + a];
      }
      for (int64_t z = lo[2]; z <= hi[2]; ++z)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
          for (int64_t x = lo[0]; x <= hi[0]; ++x)
            cuboid_idx.push_back((uint64_t)(x + Nd[0] * ((y - 1) + Nd[1] * (z - 1))));
    }
    row.sensor_mask_index = cuboid_idx.data();
    pr = &row;
  }
  if (pdim == 1) {
    if (pr_in->Ny != 1 || pr_in->Nz != 1) fail(KW_EINVAL, "Ny and Nz must be 1 in 1D");
    if (pr_in->u_source[1].flag > 0 || pr_in->u_source[2].flag > 0)
      fail(KW_EINVAL, "a 1D simulation takes the velocity source ux only");
    row.dim = 2;
    row.dy = pr_in->dx;
    row.pml_size[1] = row.pml_size[2] = 0;
    row.pml_alpha[1] = row.pml_alpha[2] = 0.0;
    row.rho0_sg[1] = row.rho0_sg[2] = nullptr;
    pr = &row;
  }
  if (pr->Nx < 2 || (pdim >= 2 && pr->Ny < 2) || (pdim == 3 && pr->Nz < 2))
    fail(KW_EINVAL, "grid sizes must be at least 2");
  if (pdim == 2 && pr->Nz != 1) fail(KW_EINVAL, "Nz must be 1 in 2D");
  if (pr->Nt < 1) fail(KW_EINVAL, "Nt must be positive");
  if (!(pr->dt > 0) || !(pr->dx > 0) || !(pr->dy > 0) || (pr->dim == 3 && !(pr->dz > 0)))
    fail(KW_EINVAL, "dt and grid spacings must be positive");
  if (!(pr->c_ref > 0)) fail(KW_EINVAL, "c_ref must be positive");

  {
    // kernels address every array with 32-bit byte offsets
    const Geometry g = make_geometry(pr->dim, pr->Nx, pr->Ny, pr->Nz);
    if ((double)g.nspec * 8.0 >= 4294967296.0 || (double)g.ncell * 8.0 >= 4294967296.0)
      fail(KW_EUNSUP, "grid too large: every array must stay below 4 GiB (at most 2^29 cells)");
  }
  auto S = std::make_unique<Sim>();
  // a simulation owns its stream unless the caller hands one in (the engine
  // runs every simulation on its own stream)
  S->arena = arena;
  S->own_stream = stream == nullptr && arena == nullptr;
  S->stream = (stream || arena) ? stream : dev_stream_create();
  S->G = make_geometry(pr->dim, pr->Nx, pr->Ny, pr->Nz);
  const Geometry& G = S->G;
  ProbDev& P = S->P;
  P.dim = pdim;
  P.Nx = G.Nx;
  P.Ny = G.Ny;
  P.Nz = G.Nz;
  P.Wm = G.Wm;
  P.nyq = G.nyq;
  P.ncell = G.ncell;
  P.nrows = G.nrows;
  P.nspec_main = G.nspec_main;
  P.nspec = G.nspec;
  S->Nt = pr->Nt;
  P.Nt = (int32_t)pr->Nt;
  P.dt = (float)pr->dt;
  const int dim = G.dim;
  const double d[3] = {pr->dx, pr->dy, pr->dz};
  for (int a = 0; a < 3; ++a) {
    S->op.d[a] = d[a];
    S->op.pml_size[a] = (int)pr->pml_size[a];
    S->op.pml_alpha[a] = pr->pml_alpha[a];
  }
  S->op.c_ref = pr->c_ref;
  S->op.use_sg = pr->use_sg != 0;
  S->op.use_k = pr->use_kspace != 0;
  const int Nd[3] = {G.Nx, G.Ny, G.Nz};

  // sensor / time reversal / record bookkeeping
  S->time_rev = pr->time_reversal_boundary_data != nullptr;
  S->has_p0 = pr->p0 != nullptr;
  if (S->time_rev && S->has_p0) fail(KW_EINVAL, "time reversal with source.p0 is not valid");
  if (S->has_p0 && pr->p_source.flag > 0)
    fail(KW_EINVAL, "source.p0 and source.p can't be defined in the same simulation.");
  S->nsens = pr->sensor_num_points;
  if (S->nsens < 0) fail(KW_EINVAL, "negative number of sensor points");
  S->record = pr->record_flags;
  if (S->time_rev) {
    if (S->nsens <= 0) fail(KW_EINVAL, "time reversal needs sensor points");
    if (pr->Nt < 2) fail(KW_EINVAL, "time reversal needs Nt >= 2");
    S->record = KW_REC_P_FINAL;
  }
  S->record_start = pr->record_start_index > 0 ? pr->record_start_index : 1;
  if (!S->time_rev && S->nsens > 0 && S->record_start > pr->Nt)
    fail(KW_EINVAL, "sensor.record_start_index must be less than or equal to kgrid.Nt");
  S->Nt_eff = S->time_rev ? pr->Nt - 1 : pr->Nt;
  S->n_rec = (S->nsens > 0 && !S->time_rev) ? pr->Nt - S->record_start + 1 : 0;
  if (S->nsens == 0) S->record &= (KW_REC_P_FINAL | KW_REC_U_FINAL);
  P.record = S->record;
  P.record_start = (int32_t)S->record_start;
  P.n_rec = (int32_t)S->n_rec;
  P.nsens = (int32_t)S->nsens;
  P.time_rev = S->time_rev ? 1 : 0;
  const bool interp = pr->sensor_tri != nullptr || pr->sensor_bc != nullptr;
  // 2-D sensor directivity (ignored by time reversal, as the reference does)
  const bool directivity = pr->directivity_angle != nullptr && S->nsens > 0 && !S->time_rev;
  if (directivity) {
    if (pdim != 2) fail(KW_EINVAL, "sensor directivity is only supported in 2D");
    if (interp) fail(KW_EINVAL, "sensor directivity needs a binary sensor mask");
    if (pr->directivity_pattern != KW_DIR_PRESSURE && pr->directivity_pattern != KW_DIR_GRADIENT)
      fail(KW_EINVAL, "Unsupported directivity pattern.");
  }
  if (interp) {
    if (!pr->sensor_tri || !pr->sensor_bc)
      fail(KW_EINVAL, "sensor_tri and sensor_bc must be given together");
    if (pr->sensor_mask_index)
      fail(KW_EINVAL, "sensor_mask_index must be NULL for interpolated Cartesian sensor points");
    if (S->time_rev) fail(KW_EINVAL, "time reversal takes a binary sensor mask");
    if (S->record & KW_REC_U_SPLIT_FIELD)
      fail(KW_EINVAL,
           "The option sensor.record = {'u_split_field'} is only compatible with a binary "
           "sensor mask.");
    check_index(pr->sensor_tri, S->nsens * (pdim + 1), G.ncell, "sensor_tri");
    S->full_mask = false;
  } else if (S->nsens > 0) {
    if (pr->sensor_mask_index) {
      check_index(pr->sensor_mask_index, S->nsens, G.ncell, "sensor.mask");
      bool identity = S->nsens == G.ncell;
      for (int64_t i = 0; identity && i < S->nsens; ++i)
        identity = pr->sensor_mask_index[i] == (uint64_t)(i + 1);
      S->full_mask = identity;
    } else {
      if (S->nsens != G.ncell)
        fail(KW_EINVAL, "sensor_mask_index is NULL but num_points is not the grid size");
      S->full_mask = true;
    }
    // the directional p is gathered per angle, so the p record of a full
    // mask is not fused into the p producer
    if (directivity) S->full_mask = false;
  }
  P.full_mask = S->full_mask ? 1 : 0;
  {
    // device-side per-step decisions (see ProbDev): the same rules the host
    // applies to a single simulation in step_once / record_step
    const uint32_t pts_p = KW_REC_P | KW_REC_P_MAX | KW_REC_P_MIN | KW_REC_P_RMS;
    const uint32_t pts_u = KW_REC_U | KW_REC_U_MAX | KW_REC_U_MIN | KW_REC_U_RMS;
    const bool sparse_p = S->nsens > 0 && !S->full_mask && !S->time_rev && (S->record & pts_p);
    P.p_store_from = sparse_p ? (int32_t)S->record_start : INT32_MAX;
    P.p_store_at = (S->record & KW_REC_P_FINAL) ? (int32_t)S->Nt_eff : -1;
    P.rec_points = S->time_rev ? 0u
                               : (S->record & (pts_p | pts_u)) & ~(S->full_mask ? pts_p : 0u) &
                                     ~(directivity ? (uint32_t)KW_REC_P : 0u);
    S->directivity = directivity;
    P.rec_grid = S->time_rev ? 0u : S->record & (KW_REC_U_MAX_ALL | KW_REC_U_MIN_ALL);
  }

  S->tiles = choose_tiles(*pr, S->stream, arena == nullptr);
  S->plan_launches();
  P.fx = S->make_plan(G.Nx);
  P.fy = S->make_plan(G.Ny);
  P.fz = dim == 3 ? S->make_plan(G.Nz) : P.fy;

  // medium (rounded once from double)
  const int64_t n = G.ncell;
  P.c0 = make_field(*S, pr->c0, pr->c0_scalar, n);
  P.rho0 = make_field(*S, pr->rho0, pr->rho0_scalar, n);
  if (!pr->c0 && !(pr->c0_scalar > 0)) fail(KW_EINVAL, "sound speed must be positive");
  if (!pr->rho0 && !(pr->rho0_scalar > 0)) fail(KW_EINVAL, "density must be positive");
  for (int a = 0; a < dim; ++a) {
    const double* r = pr->rho0_sg[a] ? pr->rho0_sg[a] : pr->rho0;
    if (r) {
      std::vector<float> inv(n);
      parallel_for(n, [&](int64_t i) { inv[i] = (float)(1.0 / r[i]); });
      P.rinv[a] = Field{S->upload(inv), ~0u};
    } else {
      P.rinv[a] = scalar_field(*S, (float)(1.0 / pr->rho0_scalar));
    }
  }
  for (int a = dim; a < 3; ++a) P.rinv[a] = P.rinv[0];
  P.nonlinear = pr->nonlinear_flag ? 1 : 0;
  P.bona = P.nonlinear ? make_field(*S, pr->BonA, pr->BonA_scalar, n) : scalar_field(*S, 0.f);
  P.tau = scalar_field(*S, 0.f);
  P.eta = P.tau;
  P.eos = pr->absorbing_flag;
  // kspaceFirstOrderAS on its radially expanded grid (see kw_problem)
  if (pr->axisymmetric == 2)
    fail(KW_EINVAL, "axisymmetric = 2 (the radial grid of the file contract) is written by "
                    "kw_save_problem and run from the file by kw_run_file");
  if (pr->axisymmetric < 0 || pr->axisymmetric > 3) fail(KW_EINVAL, "axisymmetric must be 0, 1 or 3");
  S->axisym = pr->axisymmetric == 1 || pr->axisymmetric == 3;
  S->as_ws = pr->axisymmetric == 3;
  if (S->axisym) {
    if (pdim != 2) fail(KW_EINVAL, "an axisymmetric problem is a 2D grid");
    if (!S->as_ws && G.Ny % 4 != 0) fail(KW_EINVAL, "the expanded radial grid of 'WSWA-FFT' has 4 x Ny points");
    if (S->as_ws && (G.Ny % 2 != 0 || G.Ny < 4))
      fail(KW_EINVAL, "the expanded radial grid of 'WSWS-FFT' has 2 x Ny - 2 points");
    if (!pr->use_sg) fail(KW_EINVAL, "Optional input 'UseSG' is not supported for axisymmetric simulations.");
    if (pr->absorbing_flag == 1)
      fail(KW_EUNSUP, "the axisymmetric code takes lossless or Stokes media");
    // sources and time-reversal data come laid out on the expanded grid with
    // the signs of each copy, so the 2-D source and time-reversal code
    // applies them as given.  A radial velocity source of 'WSWS-FFT' would
    // also drive the mirror point that stands in for the last radial
    // column, which the reference updates outside its layout.
    if (S->as_ws && pr->u_source[1].flag > 0 && pr->u_source[1].mode != SRC_ADDITIVE) {
      // a point source on that mirror point would set it apart from the
      // update the last column's value is recovered from
      const kw_time_source& us = pr->u_source[1];
      const int64_t row_last = G.Ny / 2;
      for (int64_t k = 0; k < us.num_points; ++k)
        if ((((int64_t)us.index[k] - 1) / G.Nx) % G.Ny == row_last)
          fail(KW_EUNSUP, "radial velocity sources (source.uy) with 'WSWS-FFT' in the 'dirichlet' "
                          "and 'additive-no-correction' modes are taken away from the last radial column");
    }
    // sensor points are grid points or simplices of the first copy (the
    // radial grid); directivity and the velocity split are 2-D options
    if (pr->directivity_angle || (pr->record_flags & KW_REC_U_SPLIT_FIELD))
      fail(KW_EUNSUP, "axisymmetric simulations take no sensor directivity or u_split_field");
  }
  if (P.eos < 0 || P.eos > 2) fail(KW_EINVAL, "absorbing_flag must be 0, 1 or 2");
  const double y = P.eos == EOS_STOKES ? 2.0 : pr->alpha_power;
  if (P.eos == EOS_ABSORBING) {
    if (!(y >= 0 && y < 3)) fail(KW_EINVAL, "medium.alpha_power must be a real number between 0 and 3.");
    if (pr->alpha_mode != 2 && y == 1.0)
      fail(KW_EINVAL, "The power law dispersion term in the equation of state is not valid for medium.alpha_power = 1.");
  }
  if (P.eos != EOS_LOSSLESS) {
    // createAbsorptionVariables.m:29-100 with db2neper.m
    // MATLAB sign(): -1, 0 or +1 (a zero sign removes the term)
    auto msign = [](double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); };
    const double s_tau = pr->alpha_sign_set ? msign(pr->alpha_sign[0]) : 1.0;
    const double s_eta = pr->alpha_sign_set ? msign(pr->alpha_sign[1]) : 1.0;
    const bool no_abs = P.eos == EOS_ABSORBING && pr->alpha_mode == 1;
    const bool no_disp = P.eos == EOS_ABSORBING && pr->alpha_mode == 2;
    const double conv = 100.0 * std::pow(1e-6 / (2.0 * kPi), y) / (20.0 * std::log10(std::exp(1.0)));
    auto tau_of = [&](double a, double c) {
      if (no_abs) return 0.0;
      const double an = a * conv;
      return P.eos == EOS_STOKES ? s_tau * (-2.0 * an * c) : s_tau * (-2.0 * an * std::pow(c, y - 1));
    };
    auto eta_of = [&](double a, double c) {
      if (no_disp || P.eos == EOS_STOKES) return 0.0;
      const double an = a * conv;
      return s_eta * (2.0 * an * std::pow(c, y) * std::tan(kPi * y / 2));
    };
    if (pr->alpha_coeff || pr->c0) {
      std::vector<float> tv(n), ev(n);
      parallel_for(n, [&](int64_t i) {
        const double a = pr->alpha_coeff ? pr->alpha_coeff[i] : pr->alpha_coeff_scalar;
        const double c = pr->c0 ? pr->c0[i] : pr->c0_scalar;
        tv[i] = (float)tau_of(a, c);
        ev[i] = (float)eta_of(a, c);
      });
      P.tau = Field{S->upload(tv), ~0u};
      if (P.eos == EOS_ABSORBING) P.eta = Field{S->upload(ev), ~0u};
    } else {
      P.tau = scalar_field(*S, (float)tau_of(pr->alpha_coeff_scalar, pr->c0_scalar));
      P.eta = scalar_field(*S, (float)eta_of(pr->alpha_coeff_scalar, pr->c0_scalar));
    }
    if (pr->alpha_coeff)
      for (int64_t i = 0; i < n; ++i)
        if (!(pr->alpha_coeff[i] >= 0)) fail(KW_EINVAL, "medium.alpha_coeff must be non-negative and real.");
  }

  // PML and shift-derivative operators (kspaceFirstOrder3D.m:619-650)
  const bool use_sg = pr->use_sg != 0;
  for (int a = 0; a < dim; ++a) {
    if (pr->pml_size[a] < 0 || 2 * pr->pml_size[a] > Nd[a]) fail(KW_EINVAL, "PML size larger than the grid");
    auto pr_ = pml_profile(Nd[a], d[a], pr->dt, pr->c_ref, (int)pr->pml_size[a], pr->pml_alpha[a], false);
    auto ps_ = pml_profile(Nd[a], d[a], pr->dt, pr->c_ref, (int)pr->pml_size[a], pr->pml_alpha[a], use_sg);
    if (S->axisym && a == 1) {
      // getPML.m with axisymmetric true on the radial grid of No points (the
      // outer PML only), laid out on the four copies of the expanded axis as
      // the fields are: p-like [f, f(No), flip(f(2:No)), f, f(No),
      // flip(f(2:No))] (the mirror columns hold p = 0), y-staggered [f,
      // flip(f), f, flip(f)]
      // 'WSWS-FFT': [f, flip(f(2:No-1))] for p-like grids (every copy
      // symmetric) and [f(1:No-1), flip(f(1:No-1))] for y-staggered ones
      // (the last staggered point, which the reference updates but never
      // feeds back, is the mirror of the one before)
      const int No = S->as_ws ? Nd[1] / 2 + 1 : Nd[1] / 4;
      if (pr->pml_size[1] > No) fail(KW_EINVAL, "PML size larger than the grid");
      auto rp = pml_profile(No, d[1], pr->dt, pr->c_ref, (int)pr->pml_size[1], pr->pml_alpha[1], false);
      auto rs = pml_profile(No, d[1], pr->dt, pr->c_ref, (int)pr->pml_size[1], pr->pml_alpha[1], true);
      const int ps = (int)pr->pml_size[1];
      if (2 * ps <= No)
        for (int i = 0; i < ps; ++i) rp[i] = rs[i] = 1.0;
      else
        fail(KW_EUNSUP, "an axisymmetric radial PML wider than half the radial grid");
      if (S->as_ws) {
        S->as_a = rs[No - 1];
        S->as_ap = rs[No - 2];
      }
      for (int j = 0; j < Nd[1]; ++j) {
        if (S->as_ws) {
          pr_[j] = j < No ? rp[j] : rp[2 * No - 2 - j];
          ps_[j] = j < No - 1 ? rs[j] : rs[2 * No - 3 - j];
          continue;
        }
        const int b = j / No, r = j % No;
        pr_[j] = (b & 1) == 0 ? rp[r] : (r == 0 ? rp[No - 1] : rp[No - r]);
        ps_[j] = (b & 1) == 0 ? rs[r] : rs[No - 1 - r];
      }
    }
    std::vector<float> f1(Nd[a]), f2(Nd[a]);
    for (int i = 0; i < Nd[a]; ++i) {
      f1[i] = (float)pr_[i];
      f2[i] = (float)ps_[i];
    }
    P.pml[a] = S->upload(f1);
    P.pml_sg[a] = S->upload(f2);
    const int len = a == 0 ? G.Nx / 2 + 1 : Nd[a];
    std::vector<cf> dp(len), dn(len);
    for (int i = 0; i < len; ++i) {
      const double k = kvec(i, Nd[a], d[a]);
      std::complex<double> ik(0.0, k);
      std::complex<double> p = ik, m = ik;
      if (use_sg) {
        p = ik * std::exp(std::complex<double>(0.0, k * d[a] / 2));
        m = ik * std::exp(std::complex<double>(0.0, -k * d[a] / 2));
      }
      if (Nd[a] % 2 == 0 && i == Nd[a] / 2) {
        // the Nyquist bin is its own mirror: real(ifftn(D .* F)) keeps only
        // the real part of D there (the Hermitian part of the operator), and
        // the half-spectrum transforms, which assume Hermitian spectra, are
        // exact with it.  Staggered operators are real there already; the
        // unstaggered i k ('UseSG' false) becomes 0.
        p = std::complex<double>(p.real(), 0.0);
        m = std::complex<double>(m.real(), 0.0);
      }
      dp[i] = cf{(float)p.real(), (float)p.imag()};
      dn[i] = cf{(float)m.real(), (float)m.imag()};
    }
    P.dpos[a] = S->upload(dp);
    P.dneg[a] = S->upload(dn);
  }

  // spectral tables (kspaceFirstOrder3D.m:653-661; createAbsorptionVariables.m)
  const double Ntot = (double)G.ncell;
  const double cref = pr->c_ref, dt = pr->dt;
  const bool use_k = pr->use_kspace != 0;
  auto kmag = [&](int ix, int iy, int iz) {
    const double kx = kvec(ix, G.Nx, d[0]), ky = kvec(iy, G.Ny, d[1]);
    double acc = 0.0 + kx * kx;
    acc = acc + ky * ky;
    if (dim == 3) {
      const double kz = kvec(iz, G.Nz, d[2]);
      acc = acc + kz * kz;
    }
    return std::sqrt(acc);
  };
  P.kappa = S->spectral_table([&](int ix, int iy, int iz) {
    return (use_k ? sinc(cref * kmag(ix, iy, iz) * dt / 2) : 1.0) / Ntot;
  });
  if (S->axisym) {
    // the radial divergence adds shift_y^- κ fft(u_y / r) (kspaceFirstOrderAS.m:
    // 954-969): κ times the y shift (which carries the 1/(Nx Ny) of the
    // round trip) in one last-axis pass
    S->as_kappa = S->spectral_table([&](int ix, int iy, int iz) {
      return use_k ? sinc(cref * kmag(ix, iy, iz) * dt / 2) : 1.0;
    });
    // ir: + on the copies that hold u_r, - on the mirrored ones, so that
    // ir .* u_r has the even layout of u_r / r (:966-977)
    const int No = S->as_ws ? (int)G.Ny / 2 + 1 : (int)G.Ny / 4;
    std::vector<float> ir(G.Ny);
    for (int j = 0; j < G.Ny; ++j) {
      int r;
      bool first;
      if (S->as_ws) {
        first = j < No - 1;
        r = first ? j : 2 * No - 3 - j;
      } else {
        first = ((j / No) & 1) == 0;
        r = first ? j % No : No - 1 - j % No;
      }
      ir[j] = (float)((first ? 1.0 : -1.0) / (r * d[1] + d[1] / 2));
    }
    S->as_ir = S->upload(ir);
    if (S->as_ws) {
      // kspaceFirstOrderAS.m:440-456 leaves rho0_sgy of the last radial
      // column at rho0 there (the interpolation past the grid is NaN); the
      // expanded grid holds rho0_sgy of the column before in its place
      const int64_t rl = (int64_t)(No - 1) * G.Nx, rp = (int64_t)(No - 2) * G.Nx;
      std::vector<float> k(G.Nx);
      for (int64_t x = 0; x < G.Nx; ++x) {
        const double rho_last = pr->rho0 ? pr->rho0[x + rl] : pr->rho0_scalar;
        const double sg_prev = pr->rho0_sg[1] ? pr->rho0_sg[1][x + rp]
                               : pr->rho0 ? pr->rho0[x + rp] : pr->rho0_scalar;
        k[x] = (float)(sg_prev / rho_last);
      }
      S->as_k = S->upload(k);
      S->as_col = S->alloc_zero<float>(2 * (size_t)G.Nx);
    }
  }
  if (P.eos == EOS_ABSORBING) {
    const bool no_abs = pr->alpha_mode == 1, no_disp = pr->alpha_mode == 2;
    // medium.alpha_filter (createAbsorptionVariables.m:56-76): both
    // operators times the centred filter, FFT bin i of an axis of length N
    // at centred index (i + floor(N/2)) mod N; the even part, which is what
    // real(ifftn(.)) keeps of the product
    const double* af = pr->alpha_filter;
    auto filt = [&](int ix, int iy, int iz) {
      if (!af) return 1.0;
      auto at = [&](int jx, int jy, int jz) {
        const int64_t cx = (jx + G.Nx / 2) % G.Nx, cy = (jy + G.Ny / 2) % G.Ny;
        const int64_t cz = dim == 3 ? (jz + G.Nz / 2) % G.Nz : 0;
        return af[cx + (int64_t)G.Nx * (cy + (int64_t)G.Ny * cz)];
      };
      const int mz = dim == 3 ? (G.Nz - iz) % G.Nz : 0;
      return 0.5 * (at(ix, iy, iz) + at((G.Nx - ix) % G.Nx, (G.Ny - iy) % G.Ny, mz));
    };
    P.nabla1 = S->spectral_table([&](int ix, int iy, int iz) {
      if (no_abs) return 0.0;
      const double v = std::pow(kmag(ix, iy, iz), y - 2);
      return std::isinf(v) ? 0.0 : v * filt(ix, iy, iz) / Ntot;
    });
    P.nabla2 = S->spectral_table([&](int ix, int iy, int iz) {
      if (no_disp) return 0.0;
      const double v = std::pow(kmag(ix, iy, iz), y - 1);
      return std::isinf(v) ? 0.0 : v * filt(ix, iy, iz) / Ntot;
    });
  }

  // fields
  P.p = S->alloc_zero<float>(n);
  for (int a = 0; a < dim; ++a) {
    P.u[a] = S->alloc_zero<float>(n);
    P.rho[a] = S->alloc_zero<float>(n);
  }
  if (P.eos == EOS_ABSORBING) P.rsum = S->alloc_zero<float>(n);
  // S[1..dim] distinct for the velocity split (S[0] holds the next step's
  // pressure spectrum when the sensors record); unused entries alias
  const bool split = pr->sensor_num_points > 0 && pr->time_reversal_boundary_data == nullptr &&
                     (pr->record_flags & KW_REC_U_SPLIT_FIELD);
  if (split && pdim == 1) fail(KW_EINVAL, "u_split_field needs a 2D or 3D simulation");
  const int nS = split ? dim + 1 : (S->axisym ? 3 : std::max(dim, 2));
  for (int b = 0; b < nS; ++b) P.S[b] = S->alloc_zero<cf>(G.nspec);
  for (int b = nS; b < 4; ++b) P.S[b] = P.S[nS - 1];

  // step counter: steps done (the first pass of each step advances it)
  int32_t* dt_index = S->alloc<int32_t>(1);
  const std::vector<int32_t> zero{0};
  dev_h2d(dt_index, zero.data(), sizeof(int32_t), S->stream);
  P.t_index = dt_index;
  P.bad = S->alloc_zero<int32_t>(1);
  S->d_tile0 = S->alloc<int32_t>(1);
  dev_h2d(S->d_tile0, zero.data(), sizeof(int32_t), S->stream);

  // sources
  setup_source(*S, pr->p_source, P.srcp, "source.p");
  for (int a = 0; a < dim; ++a) setup_source(*S, pr->u_source[a], P.srcu[a], "source.u");
  const bool need_sk = (P.srcp.len > 0 && P.srcp.mode == SRC_ADDITIVE) ||
                       [&]() {
                         for (int a = 0; a < dim; ++a)
                           if (P.srcu[a].len > 0 && P.srcu[a].mode == SRC_ADDITIVE) return true;
                         return false;
                       }();
  if (need_sk)
    S->skappa = S->spectral_table([&](int ix, int iy, int iz) {
      return (use_k ? std::cos(cref * kmag(ix, iy, iz) * dt / 2) : 1.0) / Ntot;
    });
  if (S->has_p0) S->p0 = S->upload(to_float(pr->p0, n));

  // time reversal data and map
  if (S->time_rev) {
    std::vector<int32_t> map(n, -1);
    for (int64_t i = 0; i < S->nsens; ++i)
      map[S->full_mask ? i : (int64_t)pr->sensor_mask_index[i] - 1] = (int32_t)i;
    P.tr_map = S->upload(map);
    P.tr_data = S->upload(to_float(pr->time_reversal_boundary_data, S->nsens * pr->Nt));
  }
  // sensor outputs
  if (S->nsens > 0 && !S->time_rev) {
    if (interp) {
      // record.tri / record.bc, column-major [points x (dim+1)], stored
      // point-major so that a point's vertices are adjacent
      const int K = pdim + 1;
      std::vector<int32_t> tri((size_t)S->nsens * K);
      std::vector<float> bc((size_t)S->nsens * K);
      for (int64_t i = 0; i < S->nsens; ++i)
        for (int k = 0; k < K; ++k) {
          tri[(size_t)i * K + k] = (int32_t)(pr->sensor_tri[i + k * S->nsens] - 1);
          bc[(size_t)i * K + k] = (float)pr->sensor_bc[i + k * S->nsens];
        }
      P.sens_tri = S->upload(tri);
      P.sens_bc = S->upload(bc);
    } else if (!S->full_mask) {
      std::vector<int32_t> idx(S->nsens);
      for (int64_t i = 0; i < S->nsens; ++i)
        idx[i] = pr->sensor_mask_index ? (int32_t)(pr->sensor_mask_index[i] - 1) : (int32_t)i;
      P.sens_idx = S->upload(idx);
    }
    if (directivity && (S->record & KW_REC_P)) {
      // directionalResponse.m: one k-space filter per unique angle, built on
      // the centred wavenumber grids and moved to FFT order with fftshift
      // (for an odd length, one bin off ifftshift: FFT bin i reads the
      // centred index (i - floor(N/2)) mod N).  real(ifft2(.)) keeps the
      // even part of the filter, (D(k) + D(-k))/2, which the half-spectrum
      // transforms apply exactly; 1/(Nx Ny) of the round trip folded in.
      std::vector<double> ang(pr->directivity_angle, pr->directivity_angle + S->nsens);
      std::vector<double> uniq = ang;
      std::sort(uniq.begin(), uniq.end());
      uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
      std::vector<int32_t> id(S->nsens);
      for (int64_t i = 0; i < S->nsens; ++i)
        id[i] = (int32_t)(std::lower_bound(uniq.begin(), uniq.end(), ang[i]) - uniq.begin());
      P.dir_id = S->upload(id);
      P.dirp = S->alloc<float>(G.ncell);
      auto kq = [](int i, int N, double d) {
        const int j = ((i - N / 2) % N + N) % N;
        return (2.0 * kPi / (N * d)) * (double)(j - N / 2);
      };
      const double size = pr->directivity_size, Ntot = (double)G.ncell;
      const int pattern = pr->directivity_pattern;
      for (double th : uniq) {
        auto D = [&](int ix, int iy) {
          const double kx = kq(ix, G.Nx, pr->dx), ky = kq(iy, G.Ny, pr->dy);
          if (pattern == KW_DIR_PRESSURE) return sinc(std::fabs(std::cos(th) * ky - std::sin(th) * kx) * size / 2);
          const double k = std::sqrt(kx * kx + ky * ky);
          return k == 0.0 ? 0.0 : std::fabs(std::sin(th) * ky + std::cos(th) * kx) / k;
        };
        S->dir_tab.push_back(S->spectral_table([&](int ix, int iy, int) {
          const int mx = (G.Nx - ix) % G.Nx, my = (G.Ny - iy) % G.Ny;
          return 0.5 * (D(ix, iy) + D(mx, my)) / Ntot;
        }));
      }
    }
    const int64_t ns = S->nsens;
    const uint32_t r = S->record;
    if (r & KW_REC_P) P.p_ts = S->alloc_zero<float>((size_t)ns * S->n_rec);
    if (r & KW_REC_RHO) P.rho_ts = S->alloc_zero<float>((size_t)ns * S->n_rec);
    if (r & KW_REC_P_MAX) P.p_max = S->alloc_zero<double>(ns);
    if (r & KW_REC_P_MIN) P.p_min = S->alloc_zero<double>(ns);
    if (r & KW_REC_P_RMS) P.p_rms2 = S->alloc_zero<double>(ns);
    if (r & KW_REC_P_MAX_ALL) P.p_max_all = S->alloc_zero<double>(n);
    if (r & KW_REC_P_MIN_ALL) P.p_min_all = S->alloc_zero<double>(n);
    if (r & (KW_REC_U_NON_STAGGERED | KW_REC_U_SPLIT_FIELD)) {
      // kspaceFirstOrder_createStorageVariables.m:86-114: shift operators
      // ifftshift(exp(-i k d/2)) (1 without a staggered grid); here with the
      // 1/(Nx Ny Nz) of the round trip through the spectrum and the real
      // part at the Nyquist bin of even lengths, which is all that
      // real(ifft(.)) keeps of that bin
      P.uns = S->alloc<float>(n);
      const double Ntot = (double)n;
      for (int a = 0; a < dim; ++a) {
        if (r & KW_REC_U_NON_STAGGERED) P.uns_ts[a] = S->alloc_zero<float>((size_t)ns * S->n_rec);
        const int len = a == 0 ? G.Nx / 2 + 1 : Nd[a];
        std::vector<cf> sh(len);
        for (int i = 0; i < len; ++i) {
          std::complex<double> v(1.0, 0.0);
          if (pr->use_sg) v = std::exp(std::complex<double>(0.0, -kvec(i, Nd[a], d[a]) * d[a] / 2));
          if (Nd[a] % 2 == 0 && i == Nd[a] / 2) v = std::complex<double>(v.real(), 0.0);
          sh[i] = cf{(float)(v.real() / Ntot), (float)(v.imag() / Ntot)};
        }
        P.ushift[a] = S->upload(sh);
      }
    }
    if (r & KW_REC_U_SPLIT_FIELD) {
      // kspaceFirstOrder_createStorageVariables.m:118-137: k_a / |k| (0 at
      // k = 0); the shifts keep their Nyquist value here (the pass zeroes the
      // anti-Hermitian products instead) and carry the 1/(Nx Ny Nz)
      P.usp = S->alloc<float>(n);
      const double Ntot = (double)n;
      for (int a = 0; a < dim; ++a) {
        P.usplit_p[a] = S->alloc_zero<float>((size_t)ns * S->n_rec);
        P.usplit_s[a] = S->alloc_zero<float>((size_t)ns * S->n_rec);
        P.khat[a] = S->spectral_table([&, a](int ix, int iy, int iz) {
          const double km = kmag(ix, iy, iz);
          if (km == 0.0) return 0.0;
          const int idx[3] = {ix, iy, iz};
          return kvec(idx[a], Nd[a], d[a]) / km;
        });
        const int len = a == 0 ? G.Nx / 2 + 1 : Nd[a];
        std::vector<cf> sh(len);
        for (int i = 0; i < len; ++i) {
          std::complex<double> v(1.0, 0.0);
          if (pr->use_sg) v = std::exp(std::complex<double>(0.0, -kvec(i, Nd[a], d[a]) * d[a] / 2));
          sh[i] = cf{(float)(v.real() / Ntot), (float)(v.imag() / Ntot)};
        }
        P.sshift[a] = S->upload(sh);
      }
    }
    for (int a = 0; a < dim; ++a) {
      if (r & KW_REC_U) P.u_ts[a] = S->alloc_zero<float>((size_t)ns * S->n_rec);
      if (r & KW_REC_U_MAX) P.u_max[a] = S->alloc_zero<double>(ns);
      if (r & KW_REC_U_MIN) P.u_min[a] = S->alloc_zero<double>(ns);
      if (r & KW_REC_U_RMS) P.u_rms2[a] = S->alloc_zero<double>(ns);
      if (r & KW_REC_U_MAX_ALL) P.u_max_all[a] = S->alloc_zero<double>(n);
      if (r & KW_REC_U_MIN_ALL) P.u_min_all[a] = S->alloc_zero<double>(n);
    }
  }

  if (S->axisym && !P.ushift[1]) {
    // the y shift of the u_r / r term: ifftshift(exp(-i ky dy/2)) with the
    // 1/(Nx Ny) of the round trip and the real part at the Nyquist bin, as
    // the non-staggered velocity record builds it
    std::vector<cf> sh(G.Ny);
    for (int i = 0; i < G.Ny; ++i) {
      std::complex<double> v = std::exp(std::complex<double>(0.0, -kvec(i, G.Ny, d[1]) * d[1] / 2));
      if (G.Ny % 2 == 0 && i == G.Ny / 2) v = std::complex<double>(v.real(), 0.0);
      sh[i] = cf{(float)(v.real() / (double)n), (float)(v.imag() / (double)n)};
    }
    P.ushift[1] = S->upload(sh);
  }

  // device copy of the descriptor (needed before any pass runs)
  S->dP = S->alloc<ProbDev>(1);
  dev_h2d(S->dP, &S->P, sizeof(ProbDev), S->stream);

  // additive source correction: K_M for single-series sources, per-step
  // grids for several series
  bool changed = false;
  auto prep = [&](const kw_time_source& src, TimeSource& ts, bool& grid_flag) {
    if (ts.len <= 0 || ts.mode != SRC_ADDITIVE) return;
    if (ts.nseries == 1) {
      ts.km = source_km(*S, src);
    } else {
      ts.grid = S->alloc_zero<float>(n);
      grid_flag = true;
    }
    changed = true;
  };
  prep(pr->p_source, P.srcp, S->p_grid);
  for (int a = 0; a < dim; ++a) prep(pr->u_source[a], P.srcu[a], S->u_grid[a]);
  if (changed) dev_h2d(S->dP, &S->P, sizeof(ProbDev), S->stream);
  dev_sync(S->stream);
  return S;
}

// ---------------------------------------------------------------------------
// time stepping
// ---------------------------------------------------------------------------

PassArgs bufs(int nb, int b0, int b1, int b2) {
  PassArgs a{};
  a.nbuf = nb;
  a.buf[0] = b0;
  a.buf[1] = b1;
  a.buf[2] = b2;
  return a;
}

// the gradient opens every step: its first pass advances the step number
static void gradient(Sim& S, int xflags) {
  const int dim = S.G.dim;
  PassArgs y = bufs(1, 0);
  y.advance = 1;
  if (dim == 3) S.pass(PK_YFWD, y);
  PassArgs g{};
  g.mode = SM_GRAD;
  g.advance = dim == 3 ? 0 : 1;
  S.pass(PK_SPEC, g);
  if (dim == 3) S.pass(PK_YINV, bufs(3));
  PassArgs x{};
  x.flags = xflags;
  S.pass(PK_X_GRAD, x);
}

static void r2c_p(Sim& S) {
  PassArgs a = bufs(1, 0);
  a.real[0] = S.P.p;
  S.pass(PK_X_R2C, a);
}

void src_grid(Sim& S, int which, int t) {
  const TimeSource& ts = which == 0 ? S.P.srcp : S.P.srcu[which - 1];
  if (t > ts.len) return;
  EwArgs e{};
  e.kind = EW_SRC_GRID;
  e.which = which;
  S.ew(e, S.lcell.nb);
  PassArgs a = bufs(1, 1);
  a.real[0] = (float*)ts.grid;
  S.pass(PK_X_R2C, a);
  if (S.G.dim == 3) S.pass(PK_YFWD, a);
  PassArgs f = a;
  f.mode = SM_FILTER;
  f.table = S.skappa;
  S.pass(PK_SPEC, f);
  if (S.G.dim == 3) S.pass(PK_YINV, a);
  S.pass(PK_X_C2R, a);
}

static const uint32_t kPointP = KW_REC_P | KW_REC_P_MAX | KW_REC_P_MIN | KW_REC_P_RMS;
static const uint32_t kPointU = KW_REC_U | KW_REC_U_MAX | KW_REC_U_MIN | KW_REC_U_RMS;

// Component a of the velocity on the non-staggered grid into P.uns
// (kspaceFirstOrder_extractSensorData.m:36-57): a round trip through the
// spectrum in S[1] with the half-cell shift along axis a.
static void velocity_unstaggered(Sim& S, int a) {
  const int dim = S.G.dim;
  PassArgs r = bufs(1, 1);
  r.real[0] = S.P.u[a];
  S.pass(PK_X_R2C, r);
  if (dim == 3) S.pass(PK_YFWD, r);
  PassArgs sh = r;
  sh.mode = SM_SHIFT;
  sh.flags = a;
  S.pass(PK_SPEC, sh);
  if (dim == 3) S.pass(PK_YINV, r);
  PassArgs c2r = r;
  c2r.real[0] = S.P.uns;
  S.pass(PK_X_C2R, c2r);
}

// Velocity records that need transforms, at the sensor points: the
// non-staggered components (:111-122) and their split into compressional
// and shear parts (:124-292).  The split transforms all components into
// S[1..dim] (free at the end of a step, and distinct when it is recorded),
// forms the compressional part of each in one last-axis pass, and the shear
// part is the non-staggered component minus it.  The gather kernels check
// the step on the device, so the passes may run at every step (the
// engine's captured rounds do).
void record_velocity(Sim& S) {
  const int dim = S.P.dim;
  const bool split = S.record & KW_REC_U_SPLIT_FIELD;
  if (split) {
    PassArgs all = bufs(dim, 1, 2, 3);
    for (int b = 0; b < dim; ++b) {
      PassArgs r = bufs(1, 1 + b);
      r.real[0] = S.P.u[b];
      S.pass(PK_X_R2C, r);
    }
    if (dim == 3) S.pass(PK_YFWD, all);
    PassArgs sp = all;
    sp.mode = SM_SPLIT;
    S.pass(PK_SPEC, sp);
    if (dim == 3) S.pass(PK_YINV, all);
  }
  for (int a = 0; a < dim; ++a) {
    EwArgs e{};
    e.which = a;
    if (split) {
      // the compressional part first: S[1] is needed again below
      PassArgs c2r = bufs(1, 1 + a);
      c2r.real[0] = S.P.usp;
      S.pass(PK_X_C2R, c2r);
      e.kind = EW_RECORD_SPLIT;
    } else {
      e.kind = EW_RECORD_UNS;
    }
    velocity_unstaggered(S, a);
    S.ew(e, S.lsens.nb);
  }
}

// The p time series of a sensor with directivity
// (kspaceFirstOrder2D.m:945-947, directionalResponse.m): per unique angle,
// p filtered in k-space through S[1] (free at the end of a step) into
// P.dirp, then a gather at the points of that angle.
void record_directivity(Sim& S) {
  for (size_t a = 0; a < S.dir_tab.size(); ++a) {
    PassArgs r = bufs(1, 1);
    r.real[0] = S.P.p;
    S.pass(PK_X_R2C, r);
    PassArgs f = r;
    f.mode = SM_FILTER;
    f.table = S.dir_tab[a];
    S.pass(PK_SPEC, f);
    PassArgs c2r = r;
    c2r.real[0] = S.P.dirp;
    S.pass(PK_X_C2R, c2r);
    EwArgs e{};
    e.kind = EW_RECORD_DIR;
    e.which = (int32_t)a;
    S.ew(e, S.lsens.nb);
  }
}

// The velocity half of the new toolbox's transition step between time
// steps (executeTimeStep.m:143-158 and :168-176):
//   u_d = pmlm_d (pmlm_d u_d - dtm / rho_sg_d  D+_d κm p + dtm KS(pmlm_d u_d))
// with the mean step dtm, its PML pmlm, the two-step κm and the split
// correction KS (kappaSplitCorrection.m).  The correction is folded into
// u first (u_d += dtm KS(pmlm_d u_d) / pmlm_d, so that pmlm_d u_d gains
// dtm KS(.)), then the usual gradient pass runs with the transition's dt,
// κ and PML; the divergence and everything after use the new step's.
static void transition_gradient(Sim& S) {
  Sim::DtChange& c = S.dtc;
  const int dim = S.G.dim;
  ProbDev& P = S.P;
  if (c.ksplit) {
    if (!S.scratch_grid) S.scratch_grid = S.alloc<float>((size_t)S.G.ncell);
    const int nd = P.dim;  // 1-D problems carry one velocity component
    for (int a = 0; a < nd; ++a) {
      EwArgs m{};
      m.kind = EW_AXIS_MUL;
      m.which = a;
      m.p0 = P.u[a];
      m.ptr = S.scratch_grid;
      m.prof = c.pml_sg_mean[a];
      S.ew(m, S.lcell.nb);
      // KS filter on spectral buffer 1 (buffer 0 carries fft(p))
      PassArgs f = bufs(1, 1);
      f.real[0] = S.scratch_grid;
      S.pass(PK_X_R2C, f);
      if (dim == 3) S.pass(PK_YFWD, f);
      PassArgs fl = f;
      fl.mode = SM_FILTER;
      fl.table = c.ksplit;
      S.pass(PK_SPEC, fl);
      if (dim == 3) S.pass(PK_YINV, f);
      S.pass(PK_X_C2R, f);
      EwArgs fo{};
      fo.kind = EW_AXIS_FOLD;
      fo.which = a;
      fo.p0 = S.scratch_grid;
      fo.ptr = P.u[a];
      fo.prof = c.pml_sg_mean[a];
      fo.alpha = c.dt_mean;
      S.ew(fo, S.lcell.nb);
    }
  }
  // the velocity half with the mean step
  const float* old_kappa = P.kappa;
  const float* old_pml[3] = {P.pml[0], P.pml[1], P.pml[2]};
  const float* old_sg[3] = {P.pml_sg[0], P.pml_sg[1], P.pml_sg[2]};
  P.dt = c.dt_mean;
  P.kappa = c.kappa_mix;
  for (int a = 0; a < dim; ++a) P.pml_sg[a] = c.pml_sg_mean[a];
  dev_h2d(S.dP, &P, sizeof(ProbDev), S.stream);
  gradient(S, 0);
  // the new step from here on (the copy waits for the gradient's kernels,
  // so the old tables can go)
  P.dt = c.dt_new;
  P.kappa = c.kappa_new;
  for (int a = 0; a < dim; ++a) {
    P.pml[a] = c.pml_new[a];
    P.pml_sg[a] = c.pml_sg_new[a];
  }
  dev_h2d(S.dP, &P, sizeof(ProbDev), S.stream);
  S.release(old_kappa);
  S.release(c.kappa_mix);
  S.release(c.ksplit);
  for (int a = 0; a < dim; ++a) {
    S.release(old_pml[a]);
    S.release(old_sg[a]);
    S.release(c.pml_sg_mean[a]);
  }
  c = Sim::DtChange{};
}

// The u_r / r term of an axisymmetric divergence (kspaceFirstOrderAS.m:
// 954-969 with 'WSWA-FFT'): the radial divergence is
// shift_y^- κ (i ky fft2(u_y) + fft2(ir .* u_y)) on the expanded grid, ir the
// signed inverse radius of each copy.  The second part goes through the
// spare spectral buffer 2: the x transform of u_y with each row scaled by
// ir as it is read (radial_term), then, after the divergence pass, one
// last-axis pass with κ and the y shift (whose table carries the
// 1/(Nx Ny) of the round trip) that adds the result to the y component the
// divergence pass left (divergence_radial), so the density update, the
// nonlinear coefficient and the equation of state see the whole
// divergence.
static void radial_term(Sim& S) {
  PassArgs f = bufs(1, 2);
  f.real[0] = S.P.u[1];
  f.table = S.as_ir;
  S.pass(PK_X_R2C, f);
}

// Nonlinear axisymmetric steps: the nonlinear mass-conservation coefficient
// and equation of state are even in the density while the mirrored copies
// hold -rho, so copies 2-4 of rho_x, rho_y and p are rewritten from copy 1
// after the update (the reference mirrors its radial grid every step), and
// the spectrum of p that the next step's gradient starts from is formed
// again.  The velocities stay symmetric: their update is linear in p.
static void mirror_copies(Sim& S) {
  EwArgs e{};
  e.kind = EW_AS_MIRROR;
  e.which = 3;  // rho_x, rho_y and p in one launch
  S.ew(e, S.lcell.nb);
  r2c_p(S);
}

static void ws_column(Sim& S, int which) {
  EwArgs e{};
  e.kind = EW_AS_WS_COL;
  e.which = which;
  e.ptr = S.as_col;
  e.prof = S.as_k;
  e.alpha = (float)S.as_a;
  e.beta = (float)S.as_ap;
  S.ew(e, (int)((S.G.Nx + EW_T - 1) / EW_T));
}

static void divergence_radial(Sim& S) {
  PassArgs sh = bufs(2, 2, 1);
  sh.mode = SM_SHIFT;
  sh.flags = 1;
  sh.table = S.as_kappa;
  S.pass(PK_SPEC, sh);
}

static void record_step(Sim& S, int t, bool init) {
  if (S.time_rev || S.nsens == 0 || t < S.record_start) return;
  uint32_t pts = S.record & (kPointP | kPointU);
  if (S.full_mask && !init) pts &= ~kPointP;  // fused into the p producer
  if (S.directivity) pts &= ~(uint32_t)KW_REC_P;  // the directional p below
  if (pts) {
    EwArgs e{};
    e.kind = EW_RECORD_POINTS;
    e.rec = pts;
    S.ew(e, S.lsens.nb);
  }
  if (S.record & (KW_REC_U_NON_STAGGERED | KW_REC_U_SPLIT_FIELD)) record_velocity(S);
  if (S.record & KW_REC_RHO) {
    EwArgs e{};
    e.kind = EW_RECORD_RHO;
    S.ew(e, S.lsens.nb);
  }
  if (!S.dir_tab.empty()) record_directivity(S);
  uint32_t grid = S.record & (KW_REC_U_MAX_ALL | KW_REC_U_MIN_ALL);
  if (init) grid |= S.record & (KW_REC_P_MAX_ALL | KW_REC_P_MIN_ALL);
  if (grid) {
    EwArgs e{};
    e.kind = EW_RECORD_GRID;
    e.rec = grid;
    S.ew(e, S.lcell.nb);
  }
}

void step_once(Sim& S) {
  const int t = (int)(S.steps_done + 1);
  const int dim = S.G.dim;
  const bool last = S.steps_done + 1 == S.Nt_eff;
  if (S.as_ws && S.steps_done > 0) ws_column(S, 0);
  if (t == 1 && S.has_p0) {
    // kspaceFirstOrder3D.m:1023-1037: the first step's update is replaced
    // by the initial condition; the sensor records p0
    EwArgs e{};
    e.kind = EW_INIT_P0;
    e.p0 = S.p0;
    S.ew(e, S.lcell.nb);
    r2c_p(S);
    gradient(S, XG_INIT);
    if (S.as_ws) ws_column(S, 2);
    record_step(S, t, true);
    r2c_p(S);
  } else {
    if (t == 1 && S.time_rev) {
      EwArgs e{};
      e.kind = EW_TR_PROLOGUE;
      S.ew(e, S.lcell.nb);
      r2c_p(S);
    }
    const bool need_p = (S.nsens > 0 && !S.full_mask && !S.time_rev && t >= S.record_start &&
                         (S.record & kPointP)) ||
                        (last && (S.record & KW_REC_P_FINAL));
    // a nonlinear axisymmetric step rewrites the mirrored copies of p from
    // the stored field
    const int pflag = (need_p || (S.axisym && !S.as_ws && S.P.nonlinear)) ? X_STORE_P : 0;
    if (S.p_grid) src_grid(S, 0, t);
    for (int a = 0; a < dim; ++a)
      if (S.u_grid[a]) src_grid(S, a + 1, t);
    if (S.dtc.armed)
      transition_gradient(S);
    else
      gradient(S, 0);
    if (S.axisym) radial_term(S);
    if (dim == 3) S.pass(PK_YFWD, bufs(3));
    PassArgs dv{};
    dv.mode = SM_DIV;
    S.pass(PK_SPEC, dv);
    if (dim == 3) S.pass(PK_YINV, bufs(3));
    if (S.axisym) divergence_radial(S);
    PassArgs xd{};
    xd.flags = pflag;
    S.pass(PK_X_DIV, xd);
    if (S.P.eos == EOS_ABSORBING) {
      if (dim == 3) S.pass(PK_YFWD, bufs(2));
      PassArgs ab{};
      ab.mode = SM_ABS;
      S.pass(PK_SPEC, ab);
      if (dim == 3) S.pass(PK_YINV, bufs(2));
      PassArgs xe{};
      xe.flags = pflag;
      S.pass(PK_X_EOS, xe);
    }
    // 'WSWS-FFT' copies are all +rho, which the even nonlinear terms keep
    if (S.axisym && !S.as_ws && S.P.nonlinear) mirror_copies(S);
    if (S.as_ws) ws_column(S, 1);
    record_step(S, t, false);
  }
  S.steps_done++;
}

void step_n(Sim& S, int64_t nsteps) {
  if (nsteps < 0) fail(KW_EINVAL, "negative step count");
  if (S.steps_done + nsteps > S.Nt_eff)
    fail(KW_ESTATE, "requested steps exceed the number of time steps of the simulation");
  for (int64_t i = 0; i < nsteps; ++i) step_once(S);
}

// ---------------------------------------------------------------------------
// the new toolbox's repeated run(): more steps, a new time step, an initial
// particle velocity
// ---------------------------------------------------------------------------

void extend(Sim& S, int64_t Nt_total) {
  if (S.time_rev) fail(KW_EINVAL, "a time-reversal simulation cannot be extended");
  if (S.arena) fail(KW_ESTATE, "the no-allocation entry point cannot extend a simulation");
  if (Nt_total < S.Nt) fail(KW_EINVAL, "Nt_total is below the simulation's number of time steps");
  if (Nt_total >= INT32_MAX) fail(KW_EUNSUP, "too many time steps");
  if (Nt_total == S.Nt) return;
  dev_sync(S.stream);
  ProbDev& P = S.P;
  const int64_t ns = S.nsens;
  const int64_t n_old = S.n_rec;
  const int64_t n_new = ns > 0 ? Nt_total - S.record_start + 1 : 0;
  // time series [ns x n_rec], column-major: the old columns stay in front
  auto grow = [&](float*& buf) {
    if (!buf || n_new == n_old) return;
    float* nb = S.alloc_zero<float>((size_t)(ns * n_new));
    if (n_old > 0) {
      std::vector<float> h((size_t)(ns * n_old));
      dev_d2h(h.data(), buf, h.size() * sizeof(float), S.stream);
      dev_h2d(nb, h.data(), h.size() * sizeof(float), S.stream);
    }
    S.release(buf);
    buf = nb;
  };
  grow(P.p_ts);
  grow(P.rho_ts);
  for (int a = 0; a < 3; ++a) {
    grow(P.u_ts[a]);
    grow(P.uns_ts[a]);
    grow(P.usplit_p[a]);
    grow(P.usplit_s[a]);
  }
  S.Nt = S.Nt_eff = Nt_total;
  S.n_rec = n_new;
  P.Nt = (int32_t)Nt_total;
  P.n_rec = (int32_t)n_new;
  if (S.record & KW_REC_P_FINAL) P.p_store_at = (int32_t)S.Nt_eff;
  dev_h2d(S.dP, &P, sizeof(ProbDev), S.stream);
}

void change_dt(Sim& S, double dt_prev, double dt_new) {
  if (!(dt_new > 0) || !(dt_prev >= 0) || !std::isfinite(dt_new) || !std::isfinite(dt_prev))
    fail(KW_EINVAL, "time steps must be finite, dt_new positive and dt_prev non-negative");
  if (S.time_rev) fail(KW_EINVAL, "time reversal runs at one time step");
  if (S.dtc.armed) fail(KW_ESTATE, "a time-step change is already pending");
  const float f_prev = (float)dt_prev, f_new = (float)dt_new;
  if (f_prev == f_new) return;
  dev_sync(S.stream);
  const Geometry& G = S.G;
  const int dim = G.dim;
  const Sim::OpParams& o = S.op;
  // the reference's prevTimeStep is single: the mean is formed in single
  const float f_mean = (f_new + f_prev) / 2.0f;
  const double t1 = dt_prev, t2 = dt_new;
  const int Nd[3] = {G.Nx, G.Ny, G.Nz};
  const double Ntot = (double)G.ncell;
  auto kmag = [&](int ix, int iy, int iz) {
    const double kx = kvec(ix, G.Nx, o.d[0]), ky = kvec(iy, G.Ny, o.d[1]);
    double acc = kx * kx + ky * ky;
    if (dim == 3) {
      const double kz = kvec(iz, G.Nz, o.d[2]);
      acc += kz * kz;
    }
    return std::sqrt(acc);
  };
  Sim::DtChange c;
  c.dt_mean = f_mean;
  c.dt_new = f_new;
  // setkSpaceCorrection.m:45-53 (NaN, at k = 0, becomes 1)
  c.kappa_mix = S.spectral_table([&](int ix, int iy, int iz) {
    if (!o.use_k) return 1.0 / Ntot;
    const double ck = o.c_ref * kmag(ix, iy, iz);
    double v = 2.0 * (std::sin(ck * t2 / 2) + std::sin(ck * t1 / 2) * (std::cos(ck * t2 / 2) /
                                                                       std::cos(ck * t1 / 2))) /
               (ck * (t1 + t2));
    if (std::isnan(v)) v = 1.0;
    return v / Ntot;
  });
  if (o.use_k)
    c.ksplit = S.spectral_table([&](int ix, int iy, int iz) {
      const double ck = o.c_ref * kmag(ix, iy, iz);
      double v = 2.0 * (std::cos(ck * t2 / 2) / std::cos(ck * t1 / 2) - 1.0) / (t1 + t2);
      if (std::isnan(v)) v = 1.0;
      return v / Ntot;
    });
  c.kappa_new = S.spectral_table([&](int ix, int iy, int iz) {
    return (o.use_k ? sinc(o.c_ref * kmag(ix, iy, iz) * t2 / 2) : 1.0) / Ntot;
  });
  // setupQuarticPML.m with the mean step (velocity half) and the new step
  auto profile = [&](int a, double dt, bool staggered) {
    const auto v = pml_profile(Nd[a], o.d[a], dt, o.c_ref, o.pml_size[a], o.pml_alpha[a],
                               staggered);
    std::vector<float> f(Nd[a]);
    for (int i = 0; i < Nd[a]; ++i) f[i] = (float)v[i];
    return S.upload(f);
  };
  for (int a = 0; a < dim; ++a) {
    c.pml_sg_mean[a] = profile(a, (double)f_mean, o.use_sg);
    c.pml_new[a] = profile(a, t2, false);
    c.pml_sg_new[a] = profile(a, t2, o.use_sg);
  }
  c.armed = true;
  S.dtc = c;
}

void init_velocity(Sim& S, const double* const u[3]) {
  if (S.steps_done != 0) fail(KW_ESTATE, "the initial velocity is the first step");
  if (!S.has_p0) fail(KW_EINVAL, "an initial velocity needs the initial pressure p0");
  if (!u) fail(KW_EINVAL, "velocity is NULL");
  const int nd = S.P.dim;
  for (int a = 0; a < nd; ++a)
    if (!u[a]) fail(KW_EINVAL, "a velocity component is NULL");
  EwArgs e{};
  e.kind = EW_INIT_P0;
  e.p0 = S.p0;
  S.ew(e, S.lcell.nb);
  const int64_t n = S.G.ncell;
  std::vector<float> h((size_t)n);
  for (int a = 0; a < nd; ++a) {
    parallel_for(n, [&](int64_t i) { h[(size_t)i] = (float)u[a][i]; });
    dev_h2d(S.P.u[a], h.data(), (size_t)n * sizeof(float), S.stream);
  }
  // the step counter advances in the first pass of a step; the initial
  // condition has none, so it is set here
  const int32_t one = 1;
  dev_h2d(S.P.t_index, &one, sizeof(int32_t), S.stream);
  record_step(S, 1, true);
  r2c_p(S);
  S.steps_done = 1;
}

int64_t recorded(const Sim& S) {
  if (S.nsens == 0 || S.time_rev) return 0;
  return std::max<int64_t>(0, std::min<int64_t>(S.steps_done, S.Nt) - S.record_start + 1);
}

void read_outputs(Sim& S, kw_outputs* o) {
  if (!o) fail(KW_EINVAL, "outputs is NULL");
  dev_sync(S.stream);
  const ProbDev& P = S.P;
  const int64_t n = S.G.ncell, ns = S.nsens;
  const int64_t nr = recorded(S);
  auto get = [&](void* h, const void* dsrc, size_t bytes) {
    if (h && dsrc) dev_d2h(h, dsrc, bytes, S.stream);
  };
  auto rms = [&](double* h, const double* d2) {
    if (!h || !d2) return;
    dev_d2h(h, d2, ns * sizeof(double), S.stream);
    for (int64_t i = 0; i < ns; ++i) h[i] = nr > 0 ? std::sqrt(h[i] / (double)nr) : 0.0;
  };
  get(o->p, P.p_ts, (size_t)ns * S.n_rec * sizeof(float));
  get(o->rho, P.rho_ts, (size_t)ns * S.n_rec * sizeof(float));
  get(o->p_max, P.p_max, ns * sizeof(double));
  get(o->p_min, P.p_min, ns * sizeof(double));
  rms(o->p_rms, P.p_rms2);
  get(o->p_max_all, P.p_max_all, n * sizeof(double));
  get(o->p_min_all, P.p_min_all, n * sizeof(double));
  if (o->p_final) get(o->p_final, P.p, n * sizeof(float));
  for (int a = 0; a < S.G.dim; ++a) {
    get(o->u[a], P.u_ts[a], (size_t)ns * S.n_rec * sizeof(float));
    get(o->u_max[a], P.u_max[a], ns * sizeof(double));
    get(o->u_min[a], P.u_min[a], ns * sizeof(double));
    rms(o->u_rms[a], P.u_rms2[a]);
    get(o->u_max_all[a], P.u_max_all[a], n * sizeof(double));
    get(o->u_min_all[a], P.u_min_all[a], n * sizeof(double));
    if (o->u_final[a]) get(o->u_final[a], P.u[a], n * sizeof(float));
    get(o->u_non_staggered[a], P.uns_ts[a], (size_t)ns * S.n_rec * sizeof(float));
    get(o->u_split_p[a], P.usplit_p[a], (size_t)ns * S.n_rec * sizeof(float));
    get(o->u_split_s[a], P.usplit_s[a], (size_t)ns * S.n_rec * sizeof(float));
  }
}

// ---------------------------------------------------------------------------
// tile autotuner
// ---------------------------------------------------------------------------
// The pencils per block of each pass family are chosen by measurement the
// first time a grid shape and equation of state is seen: a scratch
// simulation of that shape (scalar medium, no sources or sensors, so every
// field stays zero and no record is written) steps under each candidate
// tile with every launch bracketed by device events, and each family takes
// the candidate with the least time of its own kernels (the families are
// independent knobs).  The built-in rule stays unless a candidate beats it
// by more than 2 %.  Choices are cached for the process and, with
// KW_TUNE_FILE set, in that text file (one line per grid:
// "dim Nx Ny Nz absorbing nonlinear x y l"), which later processes read
// instead of measuring again.  KW_TUNE=0 keeps the built-in rule; grids of
// 2^24 cells and more keep it too (their launches fill the device with the
// rule's tiles; it is the best of the sweeps at 256^3 and 384^3); the
// emulator measures only with KW_TUNE=1.

namespace {

struct TuneKey {
  int dim, Nx, Ny, Nz, absorbing, nonlinear;
  bool operator<(const TuneKey& o) const {
    return std::tie(dim, Nx, Ny, Nz, absorbing, nonlinear) <
           std::tie(o.dim, o.Nx, o.Ny, o.Nz, o.absorbing, o.nonlinear);
  }
};

std::mutex g_tune_mu;
std::map<TuneKey, Tiles> g_tune;
bool g_tune_file_read = false;
thread_local bool t_no_tune = false;        // inside a scratch simulation
thread_local const Tiles* t_force = nullptr;  // a candidate of the measurement
thread_local bool t_cache_only = false;       // read the cache, never measure

struct NoTune {
  bool prev;
  NoTune() : prev(t_no_tune) { t_no_tune = true; }
  ~NoTune() { t_no_tune = prev; }
};

void read_tune_file() {
  g_tune_file_read = true;
  const char* path = std::getenv("KW_TUNE_FILE");
  if (!path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  TuneKey k{};
  Tiles t;
  while (std::fscanf(f, "%d %d %d %d %d %d %d %d %d", &k.dim, &k.Nx, &k.Ny, &k.Nz, &k.absorbing,
                     &k.nonlinear, &t.x, &t.y, &t.l) == 9)
    if (t.x > 0 && t.y > 0 && t.l > 0) g_tune[k] = t;
  std::fclose(f);
}

// the measurement (g_tune_mu held).  Every candidate gets a simulation of
// its own: the spectral tables are laid out by the last-axis tile when a
// simulation is created, so a simulation never changes its tiles.
Tiles measure_tiles(const kw_problem& pr, void* stream) {
  NoTune guard;
  kw_problem tp{};
  tp.dim = pr.dim;
  tp.Nx = pr.Nx;
  tp.Ny = pr.Ny;
  tp.Nz = pr.Nz;
  tp.Nt = 16;
  tp.dx = pr.dx;
  tp.dy = pr.dy;
  tp.dz = pr.dz;
  tp.dt = pr.dt;
  tp.c_ref = pr.c_ref;
  tp.c0_scalar = pr.c_ref;
  tp.rho0_scalar = 1000.0;
  tp.nonlinear_flag = pr.nonlinear_flag;
  tp.BonA_scalar = 6.0;
  tp.absorbing_flag = pr.absorbing_flag;
  tp.alpha_coeff_scalar = 0.5;
  tp.alpha_power = pr.alpha_power;
  tp.alpha_mode = pr.alpha_mode;
  for (int a = 0; a < 3; ++a) {
    tp.pml_size[a] = pr.pml_size[a];
    tp.pml_alpha[a] = pr.pml_alpha[a];
  }
  tp.use_kspace = pr.use_kspace;
  tp.use_sg = pr.use_sg;
  // candidates: the built-in rule first, then powers of two (each family
  // clamps a candidate to what its threads and LDS allow)
  std::vector<Tiles> cand{Tiles{}};
  for (int c = 1; c <= 64; c *= 2) cand.push_back(Tiles{c, c, c});
  struct Eff {
    int x, y, l;
    double tx, ty, tl;
  };
  std::vector<Eff> eff;
  for (const Tiles& c : cand) {
    t_force = &c;
    std::unique_ptr<Sim> S;
    try {
      S = create_sim(&tp, stream);
    } catch (...) {
      t_force = nullptr;
      throw;
    }
    t_force = nullptr;
    const Eff e0{S->P.Px, S->P.Py, S->P.Pl, 0.0, 0.0, 0.0};
    bool seen = false;
    for (const Eff& e : eff) seen = seen || (e.x == e0.x && e.y == e0.y && e.l == e0.l);
    if (seen) continue;
    step_n(*S, 2);  // the first step takes its own path; one warm-up step
    dev_sync(S->stream);
    S->timing = true;
    step_n(*S, 8);
    S->timing = false;
    double t[KW_NKINDS] = {};
    for (size_t k = 0; k < S->ev_used.size(); ++k) {
      t[S->ev_kind[k]] += dev_event_elapsed_ms(S->ev_used[k].first, S->ev_used[k].second);
      S->ev_free.push_back(S->ev_used[k].first);
      S->ev_free.push_back(S->ev_used[k].second);
    }
    S->ev_used.clear();
    S->ev_kind.clear();
    S->ev_bytes.clear();
    Eff e = e0;
    e.ty = t[PK_YFWD] + t[PK_YINV];
    e.tl = t[PK_SPEC];
    e.tx = t[PK_X_R2C] + t[PK_X_C2R] + t[PK_X_GRAD] + t[PK_X_DIV] + t[PK_X_EOS];
    eff.push_back(e);
    dev_sync(S->stream);
  }
  // each family: the fastest candidate, the rule's tile unless beaten by 2 %
  auto pick = [&](auto time, auto tile) {
    const Eff& rule = eff[0];
    int best = tile(rule);
    double bt = time(rule);
    for (const Eff& e : eff)
      if (time(e) < bt && time(e) < 0.98 * time(rule)) {
        bt = time(e);
        best = tile(e);
      }
    return best;
  };
  Tiles out;
  out.x = pick([](const Eff& e) { return e.tx; }, [](const Eff& e) { return e.x; });
  out.y = pr.dim == 3 ? pick([](const Eff& e) { return e.ty; }, [](const Eff& e) { return e.y; })
                      : 0;
  out.l = pick([](const Eff& e) { return e.tl; }, [](const Eff& e) { return e.l; });
  return out;
}

}  // namespace

TuneCacheOnly::TuneCacheOnly() : prev(t_cache_only) { t_cache_only = true; }
TuneCacheOnly::~TuneCacheOnly() { t_cache_only = prev; }

Tiles choose_tiles(const kw_problem& pr, void* stream, bool may_tune) {
  if (t_force) return *t_force;
  if (t_no_tune) return Tiles{};
  const char* env = std::getenv("KW_TUNE");
  if (env && std::strcmp(env, "0") == 0) return Tiles{};
  const TuneKey key{pr.dim,          (int)pr.Nx,     (int)pr.Ny,
                    (int)pr.Nz,      pr.absorbing_flag != 0 ? 1 : 0,
                    pr.nonlinear_flag != 0 ? 1 : 0};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tune_file_read) read_tune_file();
  auto it = g_tune.find(key);
  if (it != g_tune.end()) return it->second;
  const bool emulator = std::strcmp(backend_name(), "emulator") == 0;
  const double cells = (double)pr.Nx * (double)pr.Ny * (double)(pr.dim == 3 ? pr.Nz : 1);
  const bool forced = env && std::strcmp(env, "1") == 0;
  if (!may_tune || t_cache_only || pr.Ny < 2 || (emulator && !forced) ||
      (!forced && cells >= 16777216.0))
    return Tiles{};
  const Tiles t = measure_tiles(pr, stream);
  g_tune[key] = t;
  if (const char* path = std::getenv("KW_TUNE_FILE")) {
    if (FILE* f = std::fopen(path, "a")) {
      std::fprintf(f, "%d %d %d %d %d %d %d %d %d\n", key.dim, key.Nx, key.Ny, key.Nz,
                   key.absorbing, key.nonlinear, t.x, t.y, t.l);
      std::fclose(f);
    }
  }
  return t;
}

// kw_smooth: one filter pass on a scratch problem
static void smooth(int dim, const int64_t N[3], const double* in, const double* win, double* out) {
  NoTune guard;
  kw_problem pr{};
  pr.dim = dim;
  pr.Nx = N[0];
  pr.Ny = dim >= 2 ? N[1] : 1;
  pr.Nz = dim == 3 ? N[2] : 1;
  pr.Nt = 1;
  pr.dx = pr.dy = pr.dz = 1.0;
  pr.dt = 1.0;
  pr.c_ref = 1.0;
  pr.c0_scalar = 1.0;
  pr.rho0_scalar = 1.0;
  pr.use_kspace = 1;
  pr.use_sg = 1;
  auto S = create_sim(&pr);
  const Geometry& G = S->G;
  const double Ntot = (double)G.ncell;
  const int Nd[3] = {G.Nx, G.Ny, G.Nz};
  // ifftshift(win)[i] = win[(i + floor(N/2)) mod N] on each axis
  float* table = S->spectral_table([&](int ix, int iy, int iz) {
    const int64_t wx = (ix + Nd[0] / 2) % Nd[0];
    const int64_t wy = (iy + Nd[1] / 2) % Nd[1];
    const int64_t wz = dim == 3 ? (iz + Nd[2] / 2) % Nd[2] : 0;
    return win[wx + (int64_t)Nd[0] * (wy + (int64_t)Nd[1] * wz)] / Ntot;
  });
  float* f = S->upload(to_float(in, G.ncell));
  PassArgs a = bufs(1, 0);
  a.real[0] = f;
  S->pass(PK_X_R2C, a);
  if (dim == 3) S->pass(PK_YFWD, a);
  PassArgs fl = a;
  fl.mode = SM_FILTER;
  fl.table = table;
  S->pass(PK_SPEC, fl);
  if (dim == 3) S->pass(PK_YINV, a);
  S->pass(PK_X_C2R, a);
  std::vector<float> h(G.ncell);
  dev_d2h(h.data(), f, G.ncell * sizeof(float), S->stream);
  for (int64_t i = 0; i < G.ncell; ++i) out[i] = h[i];
}

}  // namespace kw

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

struct kw_sim {
  std::unique_ptr<kw::Sim> s;
};

static const char* kKindNames[KW_NKINDS] = {"y_fwd", "y_inv", "spectral", "x_r2c", "x_c2r",
                                            "x_grad", "x_div", "x_eos", "elementwise"};

extern "C" {

int kw_abi_version(void) { return KW_ABI_VERSION; }

const char* kw_backend(void) { return kw::backend_name(); }

const char* kw_kind_name(int k) { return (k >= 0 && k < KW_NKINDS) ? kKindNames[k] : ""; }

int kw_output_sizes(const kw_problem* prob, kw_sizes* sizes, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!prob || !sizes) kw::fail(KW_EINVAL, "NULL argument");
    const bool tr = prob->time_reversal_boundary_data != nullptr;
    const int64_t rs = prob->record_start_index > 0 ? prob->record_start_index : 1;
    sizes->num_sensor_points = prob->sensor_num_points;
    sizes->num_recorded_time_points =
        (prob->sensor_num_points > 0 && !tr) ? std::max<int64_t>(0, prob->Nt - rs + 1) : 0;
    sizes->grid_cells = prob->Nx * prob->Ny * (prob->dim == 3 ? prob->Nz : 1);
  });
}

int kw_tiles(const kw_problem* prob, int32_t tiles[3], char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!prob || !tiles) kw::fail(KW_EINVAL, "NULL argument");
    if (prob->dim < 1 || prob->dim > 3) kw::fail(KW_EINVAL, "dim must be 1, 2 or 3");
    kw_problem row = *prob;
    if (row.dim == 1) {  // a 1-D problem runs as a 2-D grid of one row
      row.dim = 2;
      row.Ny = row.Nz = 1;
      row.dy = row.dx;
    }
    if (row.dim == 2) row.Nz = 1;
    if (row.Nx < 2 || (prob->dim >= 2 && row.Ny < 2) || (row.dim == 3 && row.Nz < 2))
      kw::fail(KW_EINVAL, "grid sizes must be at least 2");
    if (!(row.dt > 0) || !(row.dx > 0) || !(row.dy > 0) || (row.dim == 3 && !(row.dz > 0)) ||
        !(row.c_ref > 0))
      kw::fail(KW_EINVAL, "dt, grid spacings and c_ref must be positive");
    kw::Sim S;
    S.own_stream = false;
    S.G = kw::make_geometry(row.dim, row.Nx, row.Ny, row.Nz);
    S.tiles = kw::choose_tiles(row, nullptr, true);
    S.plan_launches();
    tiles[0] = S.P.Px;
    tiles[1] = row.dim == 3 ? S.P.Py : 0;
    tiles[2] = S.P.Pl;
  });
}

int kw_sim_create(const kw_problem* prob, kw_sim** sim, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle pointer");
    *sim = nullptr;
    auto s = kw::create_sim(prob);
    *sim = new kw_sim{std::move(s)};
  });
}

int kw_sim_step(kw_sim* sim, int64_t nsteps, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::step_n(*sim->s, nsteps);
  });
}

int kw_sim_sync(kw_sim* sim, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::dev_sync(sim->s->stream);
  });
}

int kw_sim_read(kw_sim* sim, kw_outputs* out, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::read_outputs(*sim->s, out);
  });
}

int64_t kw_sim_steps_done(const kw_sim* sim) { return sim ? sim->s->steps_done : -1; }

int64_t kw_sim_device_bytes(const kw_sim* sim) { return sim ? sim->s->bytes : -1; }

int kw_sim_extend(kw_sim* sim, int64_t Nt_total, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::extend(*sim->s, Nt_total);
  });
}

int kw_sim_change_dt(kw_sim* sim, double dt_prev, double dt_new, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::change_dt(*sim->s, dt_prev, dt_new);
  });
}

int kw_sim_init_velocity(kw_sim* sim, const double* const u[3], char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::init_velocity(*sim->s, u);
  });
}

int kw_sim_get_fields(kw_sim* sim, float* p, float* const u[3], float* const rho[3], char* err,
                      size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::Sim& S = *sim->s;
    kw::dev_sync(S.stream);
    const size_t nb = (size_t)S.G.ncell * sizeof(float);
    if (p) kw::dev_d2h(p, S.P.p, nb, S.stream);
    for (int a = 0; a < S.P.dim; ++a) {
      if (u && u[a]) kw::dev_d2h(u[a], S.P.u[a], nb, S.stream);
      if (rho && rho[a]) kw::dev_d2h(rho[a], S.P.rho[a], nb, S.stream);
    }
  });
}

int kw_sim_destroy(kw_sim* sim) {
  if (!sim) return KW_OK;
  try {
    kw::dev_sync(sim->s->stream);
  } catch (...) {
  }
  delete sim;
  return KW_OK;
}

int kw_sim_step_timed(kw_sim* sim, int64_t nsteps, double* ms, kw_kernel_stats* stats,
                      char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!sim) kw::fail(KW_EINVAL, "NULL handle");
    kw::Sim& S = *sim->s;
    void* a = kw::dev_event_create();
    void* b = kw::dev_event_create();
    S.timing = stats != nullptr;
    try {
      kw::dev_sync(S.stream);
      kw::dev_event_record(a, S.stream);
      kw::step_n(S, nsteps);
      kw::dev_event_record(b, S.stream);
      const double t = kw::dev_event_elapsed_ms(a, b);
      if (ms) *ms = t;
      if (stats) {
        for (int k = 0; k < KW_NKINDS; ++k) {
          stats->ms[k] = 0.0;
          stats->launches[k] = 0;
          stats->bytes[k] = 0.0;
        }
        for (size_t i = 0; i < S.ev_used.size(); ++i) {
          const int k = S.ev_kind[i];
          stats->ms[k] += kw::dev_event_elapsed_ms(S.ev_used[i].first, S.ev_used[i].second);
          stats->launches[k] += 1;
          stats->bytes[k] += S.ev_bytes[i];
          S.ev_free.push_back(S.ev_used[i].first);
          S.ev_free.push_back(S.ev_used[i].second);
        }
      }
    } catch (...) {
      S.timing = false;
      kw::dev_event_destroy(a);
      kw::dev_event_destroy(b);
      throw;
    }
    S.ev_used.clear();
    S.ev_kind.clear();
    S.ev_bytes.clear();
    S.timing = false;
    kw::dev_event_destroy(a);
    kw::dev_event_destroy(b);
  });
}

int kw_run(const kw_problem* prob, kw_outputs* out, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    auto s = kw::create_sim(prob);
    kw::step_n(*s, s->Nt_eff);
    kw::read_outputs(*s, out);
  });
}

// One call at a time per process: every call carves its buffers from the
// same scratch array of the current device.
static std::mutex g_noalloc_mu;

int kw_run_noalloc(const kw_problem* prob, kw_outputs* out, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    kw::Arena arena;
    arena.base = (char*)kw::dev_scratch(&arena.size);
    auto s = kw::create_sim(prob, nullptr, &arena);
    kw::step_n(*s, s->Nt_eff);
    kw::read_outputs(*s, out);
  });
}

int kw_smooth(int32_t dim, const int64_t N[3], const double* in, const double* win, double* out,
              char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!N || !in || !win || !out) kw::fail(KW_EINVAL, "NULL argument");
    if (dim < 1 || dim > 3) kw::fail(KW_EINVAL, "dim must be 1, 2 or 3");
    kw::smooth(dim, N, in, win, out);
  });
}

}  // extern "C"
