// The native engines' file contract: kw_save_problem writes the HDF5 input
// file of a problem (the variable set and names of
// +kwave/+legacy/private/kspaceFirstOrder_saveToDisk.m, matrices as
// +kwave/+legacy/writeMatrix.m stores them, file attributes of
// writeAttributes.m and private/getH5Literals.m), and kw_run_file runs such
// a file and writes the output file that kspaceFirstOrder3DC.m:571-777
// reads back (/p, /p_max, ..., /p_final, per-cuboid groups, the sizes).
#include <ctime>
#include <mutex>

#include "kw_h5.h"
#include "kw_sim.h"

namespace {

using kw::fail;
using U64 = std::vector<uint64_t>;

// MATLAB array size [a b c] as HDF5 dims (slowest first)
U64 mdims(uint64_t a, uint64_t b = 1, uint64_t c = 1) { return U64{c, b, a}; }

std::string now_string() {
  char buf[64];
  const std::time_t t = std::time(nullptr);
  std::tm tm{};
  localtime_r(&t, &tm);
  std::strftime(buf, sizeof buf, "%d-%b-%Y-%H-%M-%S", &tm);
  return buf;
}

// root attributes of a contract file (writeAttributes.m, getH5Literals.m)
void file_attrs(kwh5::Writer& w, const char* type, const std::string& descr) {
  w.attr("", "major_version", "1");
  w.attr("", "minor_version", "2");
  w.attr("", "created_by", "kwhip (MI355X) native engine");
  w.attr("", "file_description", descr);
  w.attr("", "file_type", type);
  w.attr("", "creation_date", now_string());
}

// a matrix of the contract with its data-type attributes (writeMatrix.m)
struct Store {
  kwh5::Writer& w;
  std::vector<std::unique_ptr<std::vector<float>>> f32;
  std::vector<std::unique_ptr<U64>> u64;
  void real(const std::string& name, const U64& dims, std::vector<float> v) {
    f32.push_back(std::make_unique<std::vector<float>>(std::move(v)));
    w.add_f32(name, dims, f32.back()->data());
    w.attr(name, "domain_type", "real");
    w.attr(name, "data_type", "float");
  }
  void real(const std::string& name, const U64& dims, const double* a, int64_t n) {
    std::vector<float> v((size_t)n);
    for (int64_t i = 0; i < n; ++i) v[i] = (float)a[i];
    real(name, dims, std::move(v));
  }
  void scalar(const std::string& name, double v) { real(name, mdims(1), std::vector<float>{(float)v}); }
  void integer(const std::string& name, const U64& dims, U64 v) {
    u64.push_back(std::make_unique<U64>(std::move(v)));
    w.add_u64(name, dims, u64.back()->data());
    w.attr(name, "domain_type", "real");
    w.attr(name, "data_type", "long");
  }
  void integer(const std::string& name, uint64_t v) { integer(name, mdims(1), U64{v}); }
};

// -------------------------------------------------------------------------
// problem -> input file
// -------------------------------------------------------------------------
void save_problem(const kw_problem* pr, const char* path) {
  if (!pr || !path) fail(KW_EINVAL, "NULL argument");
  const int dim = pr->dim;
  if (dim != 2 && dim != 3) fail(KW_EUNSUP, "the file contract holds 2-D and 3-D problems only");
  if (pr->time_reversal_boundary_data)
    fail(KW_EUNSUP, "the file contract takes forward simulations (time reversal goes over as a "
                    "dirichlet pressure source, kspaceFirstOrder3DC.m:441-449)");
  if (pr->sensor_tri || pr->directivity_angle)
    fail(KW_EUNSUP, "the file contract takes binary or cuboid-corner sensor masks only");
  if (pr->alpha_filter || pr->alpha_mode != 0 || pr->alpha_sign_set)
    fail(KW_EUNSUP, "alpha_filter, alpha_mode and alpha_sign are not part of the file contract");
  if (!pr->use_kspace || !pr->use_sg)
    fail(KW_EUNSUP, "'UsekSpace' and 'UseSG' false are not part of the file contract");
  // an axisymmetric problem goes over on its radial grid (2), as
  // kspaceFirstOrderAS.m writes it with 'SaveToDisk'; the expanded grid that
  // kw_run steps (1) is not a file layout
  if (pr->axisymmetric == 1 || pr->axisymmetric == 3)
    fail(KW_EINVAL, "an axisymmetric problem is written on its radial grid (axisymmetric = 2)");
  if (pr->axisymmetric != 0 && pr->axisymmetric != 2) fail(KW_EINVAL, "axisymmetric must be 0, 1, 2 or 3");
  if (pr->axisymmetric == 2 && dim != 2) fail(KW_EINVAL, "an axisymmetric problem is a 2D grid");
  const int64_t Nx = pr->Nx, Ny = pr->Ny, Nz = dim == 3 ? pr->Nz : 1;
  const int64_t n = Nx * Ny * Nz;
  kwh5::Writer w;
  Store s{w, {}, {}};
  const U64 grid = mdims((uint64_t)Nx, (uint64_t)Ny, (uint64_t)Nz);
  auto field = [&](const std::string& name, const double* a, double sc) {
    if (a)
      s.real(name, grid, a, n);
    else
      s.scalar(name, sc);
  };
  // integer variables
  s.integer("Nx", (uint64_t)Nx);
  s.integer("Ny", (uint64_t)Ny);
  s.integer("Nz", (uint64_t)Nz);
  s.integer("Nt", (uint64_t)pr->Nt);
  const kw_time_source& ps = pr->p_source;
  s.integer("p_source_flag", (uint64_t)std::max<int64_t>(0, ps.flag));
  s.integer("p0_source_flag", pr->p0 ? 1 : 0);
  const char* un[3] = {"ux", "uy", "uz"};
  for (int d = 0; d < 3; ++d)
    s.integer(std::string(un[d]) + "_source_flag",
              d < dim ? (uint64_t)std::max<int64_t>(0, pr->u_source[d].flag) : 0);
  for (const char* nm : {"sxx", "syy", "szz", "sxy", "sxz", "syz"})
    s.integer(std::string(nm) + "_source_flag", 0);
  s.integer("transducer_source_flag", 0);
  s.integer("nonuniform_grid_flag", 0);
  s.integer("nonlinear_flag", pr->nonlinear_flag ? 1 : 0);
  s.integer("absorbing_flag", (uint64_t)pr->absorbing_flag);
  s.integer("elastic_flag", 0);
  s.integer("axisymmetric_flag", pr->axisymmetric == 2 ? 1 : 0);
  const bool corners = pr->sensor_mask_corners && pr->sensor_num_cuboids > 0;
  s.integer("sensor_mask_type", corners ? 1 : 0);
  s.integer("pml_x_size", (uint64_t)pr->pml_size[0]);
  s.integer("pml_y_size", (uint64_t)pr->pml_size[1]);
  s.integer("pml_z_size", dim == 3 ? (uint64_t)pr->pml_size[2] : 0);
  // the velocity sources share one mode, series count and index list
  int ud = -1;
  for (int d = 0; d < dim; ++d)
    if (pr->u_source[d].flag > 0) {
      const kw_time_source& a = pr->u_source[d];
      if (a.sig_index) fail(KW_EUNSUP, "labelled source masks are not part of the file contract");
      if (ud >= 0) {
        const kw_time_source& b = pr->u_source[ud];
        if (a.mode != b.mode || a.num_points != b.num_points ||
            (a.num_series > 1) != (b.num_series > 1) ||
            !std::equal(a.index, a.index + a.num_points, b.index))
          fail(KW_EUNSUP, "the file contract has one velocity source mask and mode");
      }
      ud = d;
    }
  if (ud >= 0) {
    const kw_time_source& u = pr->u_source[ud];
    s.integer("u_source_mode", (uint64_t)u.mode);
    s.integer("u_source_many", u.num_series > 1 ? 1 : 0);
    s.integer("u_source_index", mdims((uint64_t)u.num_points), U64(u.index, u.index + u.num_points));
  }
  if (ps.flag > 0) {
    if (ps.sig_index) fail(KW_EUNSUP, "labelled source masks are not part of the file contract");
    s.integer("p_source_mode", (uint64_t)ps.mode);
    s.integer("p_source_many", ps.num_series > 1 ? 1 : 0);
    s.integer("p_source_index", mdims((uint64_t)ps.num_points), U64(ps.index, ps.index + ps.num_points));
  }
  if (corners) {
    // [x1 y1 z1 x2 y2 z2] per cuboid; a 2-D rectangle has z1 = z2 = 1
    U64 c((size_t)(6 * pr->sensor_num_cuboids), 1);
    for (int64_t q = 0; q < pr->sensor_num_cuboids; ++q)
      for (int a = 0; a < dim; ++a) {
        c[6 * q + a] = pr->sensor_mask_corners[q * 2 * dim + a];
        c[6 * q + 3 + a] = pr->sensor_mask_corners[q * 2 * dim + dim + a];
      }
    s.integer("sensor_mask_corners", mdims(6, (uint64_t)pr->sensor_num_cuboids), std::move(c));
  } else {
    U64 idx;
    if (pr->sensor_mask_index) {
      idx.assign(pr->sensor_mask_index, pr->sensor_mask_index + pr->sensor_num_points);
    } else {
      idx.resize((size_t)pr->sensor_num_points);
      for (int64_t i = 0; i < pr->sensor_num_points; ++i) idx[i] = (uint64_t)i + 1;
    }
    if (idx.empty()) fail(KW_EINVAL, "the file contract needs a non-empty sensor mask");
    const U64 dims = mdims(idx.size());
    s.integer("sensor_mask_index", dims, std::move(idx));
  }
  // real variables (names with a z are not written for 2-D problems)
  s.scalar("dx", pr->dx);
  s.scalar("dy", pr->dy);
  if (dim == 3) s.scalar("dz", pr->dz);
  s.scalar("pml_x_alpha", pr->pml_alpha[0]);
  s.scalar("pml_y_alpha", pr->pml_alpha[1]);
  if (dim == 3) s.scalar("pml_z_alpha", pr->pml_alpha[2]);
  s.scalar("dt", pr->dt);
  field("c0", pr->c0, pr->c0_scalar);
  s.scalar("c_ref", pr->c_ref);
  field("rho0", pr->rho0, pr->rho0_scalar);
  const char* sg[3] = {"rho0_sgx", "rho0_sgy", "rho0_sgz"};
  for (int d = 0; d < dim; ++d)
    field(sg[d], pr->rho0_sg[d] ? pr->rho0_sg[d] : pr->rho0, pr->rho0_scalar);
  if (pr->nonlinear_flag) field("BonA", pr->BonA, pr->BonA_scalar);
  if (pr->absorbing_flag) {
    field("alpha_coeff", pr->alpha_coeff, pr->alpha_coeff_scalar);
    s.scalar("alpha_power", pr->alpha_power);
  }
  for (int d = 0; d < dim; ++d) {
    const kw_time_source& u = pr->u_source[d];
    if (u.flag <= 0) continue;
    s.real(std::string(un[d]) + "_source_input", mdims((uint64_t)u.num_series, (uint64_t)u.flag),
           u.input, u.num_series * u.flag);
  }
  if (ps.flag > 0)
    s.real("p_source_input", mdims((uint64_t)ps.num_series, (uint64_t)ps.flag), ps.input,
           ps.num_series * ps.flag);
  if (pr->p0) s.real("p0_source_input", grid, pr->p0, n);
  file_attrs(w, "input", "Input data created by kwhip (kw_save_problem)");
  w.write(path);
}

// -------------------------------------------------------------------------
// input file -> problem
// -------------------------------------------------------------------------
struct Loaded {
  kw_problem pb{};
  std::vector<double> c0, rho0, sg[3], bona, alpha, p0, psig, usig[3];
  U64 pidx, uidx, sens, corners;
  bool axisym = false;  // axisymmetric_flag: the radial grid of kspaceFirstOrderAS
  int64_t Ny_radial = 0;
};

void load_problem(const kwh5::Reader& r, Loaded& L) {
  kw_problem& p = L.pb;
  if (r.has_attr("", "file_type") && r.attr_string("", "file_type") != "input")
    fail(KW_EINVAL, "the file is not a k-Wave input file (file_type '" +
                        r.attr_string("", "file_type") + "')");
  auto u = [&](const char* nm) -> uint64_t {
    if (!r.has(nm)) fail(KW_EINVAL, std::string("the input file has no '") + nm + "'");
    const U64 v = r.read_u64(nm);
    if (v.size() != 1) fail(KW_EINVAL, std::string("'") + nm + "' is not a scalar");
    return v[0];
  };
  auto uopt = [&](const char* nm) -> uint64_t { return r.has(nm) ? u(nm) : 0; };
  auto real = [&](const char* nm) -> double {
    if (!r.has(nm)) fail(KW_EINVAL, std::string("the input file has no '") + nm + "'");
    return r.scalar(nm);
  };
  L.axisym = uopt("axisymmetric_flag") != 0;
  for (const char* nm : {"elastic_flag", "transducer_source_flag", "nonuniform_grid_flag", "sxx_source_flag", "syy_source_flag",
                         "szz_source_flag", "sxy_source_flag", "sxz_source_flag",
                         "syz_source_flag"})
    if (uopt(nm) != 0)
      fail(KW_EUNSUP, std::string(nm) + " is set: only the fluid, uniform-grid, Cartesian "
                                        "simulations are supported");
  const int64_t Nx = (int64_t)u("Nx"), Ny = (int64_t)u("Ny"), Nz = (int64_t)u("Nz");
  const int dim = Nz > 1 ? 3 : 2;
  if (L.axisym && dim != 2) fail(KW_EINVAL, "an axisymmetric input file holds a 2D grid");
  const int64_t n = Nx * Ny * Nz;
  p.dim = dim;
  p.Nx = Nx;
  p.Ny = Ny;
  p.Nz = Nz;
  p.Nt = (int64_t)u("Nt");
  p.dx = real("dx");
  p.dy = real("dy");
  p.dz = dim == 3 ? real("dz") : 0.0;
  p.dt = real("dt");
  p.c_ref = real("c_ref");
  // a medium quantity: one value or one per grid point
  auto field = [&](const char* nm, std::vector<double>& store, const double*& ptr, double& sc) {
    store = r.read_double(nm);
    if (store.size() == 1) {
      sc = store[0];
      ptr = nullptr;
    } else if ((int64_t)store.size() == n) {
      ptr = store.data();
    } else {
      fail(KW_EINVAL, std::string("'") + nm + "' must hold 1 or Nx*Ny*Nz values");
    }
  };
  field("c0", L.c0, p.c0, p.c0_scalar);
  field("rho0", L.rho0, p.rho0, p.rho0_scalar);
  const char* sgn[3] = {"rho0_sgx", "rho0_sgy", "rho0_sgz"};
  for (int d = 0; d < dim; ++d) {
    double sc = 0.0;
    const double* ptr = nullptr;
    field(sgn[d], L.sg[d], ptr, sc);
    if (!ptr && !(p.rho0 == nullptr && sc == p.rho0_scalar)) {
      // a single staggered density other than the (single) density: a grid
      // of that value
      L.sg[d].assign((size_t)n, sc);
      ptr = L.sg[d].data();
    }
    p.rho0_sg[d] = ptr;  // NULL: the density itself
  }
  p.pml_size[0] = (int64_t)u("pml_x_size");
  p.pml_size[1] = (int64_t)u("pml_y_size");
  p.pml_size[2] = dim == 3 ? (int64_t)u("pml_z_size") : 0;
  p.pml_alpha[0] = real("pml_x_alpha");
  p.pml_alpha[1] = real("pml_y_alpha");
  p.pml_alpha[2] = dim == 3 ? real("pml_z_alpha") : 0.0;
  p.use_kspace = 1;
  p.use_sg = 1;
  p.nonlinear_flag = uopt("nonlinear_flag") ? 1 : 0;
  if (p.nonlinear_flag) field("BonA", L.bona, p.BonA, p.BonA_scalar);
  p.absorbing_flag = (int32_t)uopt("absorbing_flag");
  if (p.absorbing_flag < 0 || p.absorbing_flag > 2) fail(KW_EINVAL, "unknown absorbing_flag");
  if (p.absorbing_flag) {
    field("alpha_coeff", L.alpha, p.alpha_coeff, p.alpha_coeff_scalar);
    p.alpha_power = real("alpha_power");
  }
  // sources: one signal (one row) or one per source point
  auto source = [&](kw_time_source& ts, uint64_t flag, const char* mode, const char* many,
                    const U64& idx, const char* input, std::vector<double>& sig) {
    ts.flag = (int64_t)flag;
    if (flag == 0) return;
    ts.mode = (int32_t)u(mode);
    ts.num_points = (int64_t)idx.size();
    ts.index = idx.data();
    sig = r.read_double(input);
    const int64_t rows = u(many) ? ts.num_points : 1;
    if ((int64_t)sig.size() != rows * (int64_t)flag)
      fail(KW_EINVAL, std::string("'") + input + "' does not hold " + std::to_string(rows) +
                          " series of " + std::to_string(flag) + " samples");
    ts.num_series = rows;
    ts.input = sig.data();
  };
  const uint64_t pf = uopt("p_source_flag");
  if (pf) L.pidx = r.read_u64("p_source_index");
  source(p.p_source, pf, "p_source_mode", "p_source_many", L.pidx, "p_source_input", L.psig);
  const char* un[3] = {"ux", "uy", "uz"};
  for (int d = 0; d < dim; ++d) {
    const uint64_t f = uopt((std::string(un[d]) + "_source_flag").c_str());
    if (f && L.uidx.empty()) L.uidx = r.read_u64("u_source_index");
    source(p.u_source[d], f, "u_source_mode", "u_source_many", L.uidx,
           (std::string(un[d]) + "_source_input").c_str(), L.usig[d]);
  }
  if (uopt("p0_source_flag")) {
    L.p0 = r.read_double("p0_source_input");
    if ((int64_t)L.p0.size() != n) fail(KW_EINVAL, "'p0_source_input' must hold Nx*Ny*Nz values");
    p.p0 = L.p0.data();
  }
  const uint64_t smt = uopt("sensor_mask_type");
  if (smt == 0) {
    L.sens = r.read_u64("sensor_mask_index");
    p.sensor_num_points = (int64_t)L.sens.size();
    p.sensor_mask_index = L.sens.data();
  } else if (smt == 1) {
    const U64 c = r.read_u64("sensor_mask_corners");
    if (c.empty() || c.size() % 6 != 0) fail(KW_EINVAL, "'sensor_mask_corners' must be 6 x N");
    const int64_t nc = (int64_t)c.size() / 6;
    L.corners.resize((size_t)(2 * dim * nc));
    int64_t total = 0;
    for (int64_t q = 0; q < nc; ++q) {
      int64_t cnt = 1;
      for (int a = 0; a < dim; ++a) {
        const uint64_t lo = c[6 * q + a], hi = c[6 * q + 3 + a];
        L.corners[2 * dim * q + a] = lo;
        L.corners[2 * dim * q + dim + a] = hi;
        if (hi < lo) fail(KW_EINVAL, "sensor_mask_corners: x2 < x1");
        cnt *= (int64_t)(hi - lo + 1);
      }
      total += cnt;
    }
    p.sensor_mask_corners = L.corners.data();
    p.sensor_num_cuboids = nc;
    p.sensor_num_points = total;
  } else {
    fail(KW_EINVAL, "unknown sensor_mask_type");
  }
}

// -------------------------------------------------------------------------
// axisymmetric input files
// -------------------------------------------------------------------------
// kspaceFirstOrderAS.m writes its radial grid (x axial, y radial from the
// axis) with the 'WSWA' symmetry that the reference's native code steps
// with discrete trigonometric transforms.  The library steps the expanded
// grid of 'WSWA-FFT' (the same field, kw_problem.axisymmetric = 1), four
// times the radial size, so the radial inputs are laid out there:
// p-like grids as [f, 0, -flip(f(2:end)), -f, 0, flip(f(2:end))] along y
// (kspaceFirstOrderAS.m:703-707), the y-staggered density as [f, flip(f),
// f, flip(f)], material grids unsigned with each mirror column repeating
// its neighbour, and every pressure-source and axial-velocity-source point
// (i, j) as the points (i, j), (i, 2Ny - j), (i, 2Ny + j), (i, 4Ny - j)
// with signs +, -, -, + (the second and fourth left out for j = 0), a
// radial-velocity-source point as (i, j), (i, 2Ny - 1 - j), (i, 2Ny + j),
// (i, 4Ny - 1 - j) with +, +, -, -, each with a series of its own.
// Sensor points keep their indices (the first copy is the radial grid).
struct AxisymExpanded {
  std::vector<double> c0, rho0, sg[2], bona, alpha, p0, psig, usig[2];
  U64 pidx, uidx[2];
};

// p-like layout of an Nx x Ny grid (signed: copies +, -, -, +, mirror
// columns 0; unsigned: every copy +, mirror columns repeat their neighbour)
std::vector<double> layout_p(const double* f, int64_t Nx, int64_t Ny, bool sign) {
  std::vector<double> e((size_t)(Nx * 4 * Ny), 0.0);
  const double sg[4] = {1.0, -1.0, -1.0, 1.0};
  for (int b = 0; b < 4; ++b)
    for (int64_t jj = 0; jj < Ny; ++jj) {
      int64_t src;
      if (b % 2 == 0)
        src = jj;
      else if (jj > 0)
        src = Ny - jj;
      else if (sign)
        continue;  // the mirror column of a signed grid stays 0
      else
        src = Ny - 1;
      const double k = sign ? sg[b] : 1.0;
      double* o = e.data() + (b * Ny + jj) * Nx;
      const double* a = f + src * Nx;
      for (int64_t i = 0; i < Nx; ++i) o[i] = k * a[i];
    }
  return e;
}

// y-staggered layout, unsigned: [f, flip(f), f, flip(f)]
std::vector<double> layout_u(const double* f, int64_t Nx, int64_t Ny) {
  std::vector<double> e((size_t)(Nx * 4 * Ny));
  for (int b = 0; b < 4; ++b)
    for (int64_t jj = 0; jj < Ny; ++jj) {
      const int64_t src = b % 2 == 0 ? jj : Ny - 1 - jj;
      std::copy(f + src * Nx, f + (src + 1) * Nx, e.begin() + (b * Ny + jj) * Nx);
    }
  return e;
}

void expand_axisym(Loaded& L, AxisymExpanded& E) {
  kw_problem& p = L.pb;
  const int64_t Nx = p.Nx, Ny = p.Ny;
  L.Ny_radial = Ny;
  auto grid_p = [&](const double*& ptr, std::vector<double>& store, bool sign) {
    if (ptr) {
      store = layout_p(ptr, Nx, Ny, sign);
      ptr = store.data();
    }
  };
  grid_p(p.c0, E.c0, false);
  grid_p(p.rho0, E.rho0, false);
  grid_p(p.rho0_sg[0], E.sg[0], false);
  if (p.rho0_sg[1]) {
    E.sg[1] = layout_u(p.rho0_sg[1], Nx, Ny);
    p.rho0_sg[1] = E.sg[1].data();
  }
  grid_p(p.BonA, E.bona, false);
  grid_p(p.alpha_coeff, E.alpha, false);
  grid_p(p.p0, E.p0, true);
  // a source on the expanded grid: one signed series per expanded point;
  // p-like points (p, ux) at columns j, 2Ny - j, 2Ny + j, 4Ny - j with
  // signs +, -, -, + (the second and fourth left out for j = 0), y-staggered
  // ones (uy) at j, 2Ny - 1 - j, 2Ny + j, 4Ny - 1 - j with +, +, -, -
  auto expand_source = [&](kw_time_source& ts, bool staggered, U64& idx, std::vector<double>& sig) {
    if (ts.flag <= 0) return;
    struct Pt {
      uint64_t lin;
      int64_t row;
      double sign;
    };
    std::vector<Pt> pts;
    pts.reserve((size_t)(4 * ts.num_points));
    for (int64_t k = 0; k < ts.num_points; ++k) {
      const int64_t lin = (int64_t)ts.index[k] - 1;
      if (lin < 0 || lin >= Nx * Ny) fail(KW_EINVAL, "source index outside the grid");
      const int64_t i = lin % Nx, j = lin / Nx;
      const int64_t row = ts.num_series > 1 ? k : 0;
      if (staggered) {
        pts.push_back({(uint64_t)(i + Nx * j), row, 1.0});
        pts.push_back({(uint64_t)(i + Nx * (2 * Ny - 1 - j)), row, 1.0});
        pts.push_back({(uint64_t)(i + Nx * (2 * Ny + j)), row, -1.0});
        pts.push_back({(uint64_t)(i + Nx * (4 * Ny - 1 - j)), row, -1.0});
        continue;
      }
      pts.push_back({(uint64_t)(i + Nx * j), row, 1.0});
      pts.push_back({(uint64_t)(i + Nx * (2 * Ny + j)), row, -1.0});
      if (j > 0) {
        pts.push_back({(uint64_t)(i + Nx * (2 * Ny - j)), row, -1.0});
        pts.push_back({(uint64_t)(i + Nx * (4 * Ny - j)), row, 1.0});
      }
    }
    std::sort(pts.begin(), pts.end(), [](const Pt& a, const Pt& b) { return a.lin < b.lin; });
    const int64_t ne = (int64_t)pts.size(), len = ts.flag, rows = ts.num_series;
    idx.resize((size_t)ne);
    sig.resize((size_t)(ne * len));
    for (int64_t e = 0; e < ne; ++e) {
      idx[e] = pts[e].lin + 1;
      for (int64_t t = 0; t < len; ++t) sig[e + ne * t] = pts[e].sign * ts.input[pts[e].row + rows * t];
    }
    ts.num_points = ne;
    ts.index = idx.data();
    ts.num_series = ne;
    ts.input = sig.data();
  };
  expand_source(p.p_source, false, E.pidx, E.psig);
  expand_source(p.u_source[0], false, E.uidx[0], E.usig[0]);
  expand_source(p.u_source[1], true, E.uidx[1], E.usig[1]);
  p.Ny = 4 * Ny;
  p.axisymmetric = 1;
}

// -------------------------------------------------------------------------
// outputs -> output file
// -------------------------------------------------------------------------
struct OutBufs {
  std::vector<float> f[64];
  std::vector<double> d[64];
  int nf = 0, nd = 0;
  float* fl(size_t n) {
    f[nf].assign(n, 0.f);
    return f[nf++].data();
  }
  double* db(size_t n) {
    d[nd].assign(n, 0.0);
    return d[nd++].data();
  }
};

void run_file(const char* in, const char* out, uint32_t rec, int64_t start) {
  if (!in || !out) fail(KW_EINVAL, "NULL file name");
  kwh5::Reader r(in);
  Loaded L;
  load_problem(r, L);
  AxisymExpanded E;
  if (L.axisym) expand_axisym(L, E);
  kw_problem& p = L.pb;
  p.record_flags = rec;
  p.record_start_index = start > 0 ? start : 1;
  char err[1024];
  kw_sizes sz{};
  int rc = kw_output_sizes(&p, &sz, err, sizeof err);
  if (rc != KW_OK) fail(rc, err);
  const int dim = p.dim;
  const size_t ns = (size_t)sz.num_sensor_points, nr = (size_t)sz.num_recorded_time_points,
               nc = (size_t)sz.grid_cells;
  OutBufs B;
  kw_outputs o{};
  if (rec & KW_REC_P) o.p = B.fl(ns * nr);
  if (rec & KW_REC_P_MAX) o.p_max = B.db(ns);
  if (rec & KW_REC_P_MIN) o.p_min = B.db(ns);
  if (rec & KW_REC_P_RMS) o.p_rms = B.db(ns);
  if (rec & KW_REC_P_MAX_ALL) o.p_max_all = B.db(nc);
  if (rec & KW_REC_P_MIN_ALL) o.p_min_all = B.db(nc);
  if (rec & KW_REC_P_FINAL) o.p_final = B.fl(nc);
  for (int d = 0; d < dim; ++d) {
    if (rec & KW_REC_U) o.u[d] = B.fl(ns * nr);
    if (rec & KW_REC_U_MAX) o.u_max[d] = B.db(ns);
    if (rec & KW_REC_U_MIN) o.u_min[d] = B.db(ns);
    if (rec & KW_REC_U_RMS) o.u_rms[d] = B.db(ns);
    if (rec & KW_REC_U_MAX_ALL) o.u_max_all[d] = B.db(nc);
    if (rec & KW_REC_U_MIN_ALL) o.u_min_all[d] = B.db(nc);
    if (rec & KW_REC_U_FINAL) o.u_final[d] = B.fl(nc);
    if (rec & KW_REC_U_NON_STAGGERED) o.u_non_staggered[d] = B.fl(ns * nr);
  }
  rc = kw_run(&p, &o, err, sizeof err);
  if (rc != KW_OK) fail(rc, err);

  // grid outputs of an axisymmetric file are the radial grid: the first
  // copy, i.e. the first Nx * Ny cells in column-major order
  const int64_t Ny_out = L.axisym ? L.Ny_radial : p.Ny;
  const size_t nco = L.axisym ? (size_t)(p.Nx * Ny_out) : nc;
  kwh5::Writer w;
  Store s{w, {}, {}};
  s.integer("Nx", (uint64_t)p.Nx);
  s.integer("Ny", (uint64_t)Ny_out);
  s.integer("Nz", (uint64_t)(dim == 3 ? p.Nz : 1));
  s.integer("Nt", (uint64_t)p.Nt);
  s.integer("pml_x_size", (uint64_t)p.pml_size[0]);
  s.integer("pml_y_size", (uint64_t)p.pml_size[1]);
  s.integer("pml_z_size", (uint64_t)p.pml_size[2]);
  s.scalar("dt", p.dt);
  const U64 grid = mdims((uint64_t)p.Nx, (uint64_t)Ny_out, (uint64_t)(dim == 3 ? p.Nz : 1));
  auto to_f = [](const double* a, size_t n) {
    std::vector<float> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = (float)a[i];
    return v;
  };
  // sensor records: one dataset per record, or one per cuboid in a group
  // named after the record (datasets "1", "2", ...)
  const bool cub = p.sensor_mask_corners && p.sensor_num_cuboids > 0;
  auto series = [&](const std::string& nm, const float* a) {
    if (!a) return;
    if (!cub) {
      s.real(nm, mdims(ns, nr), std::vector<float>(a, a + ns * nr));
      return;
    }
    w.add_group(nm);
    size_t off = 0;
    for (int64_t q = 0; q < p.sensor_num_cuboids; ++q) {
      uint64_t e[3] = {1, 1, 1};
      for (int k = 0; k < dim; ++k)
        e[k] = p.sensor_mask_corners[q * 2 * dim + dim + k] - p.sensor_mask_corners[q * 2 * dim + k] + 1;
      const size_t m = (size_t)(e[0] * e[1] * e[2]);
      std::vector<float> v(m * nr);
      for (size_t t = 0; t < nr; ++t)
        std::copy(a + t * ns + off, a + t * ns + off + m, v.begin() + t * m);
      s.real(nm + "/" + std::to_string(q + 1), U64{nr, e[2], e[1], e[0]}, std::move(v));
      off += m;
    }
  };
  auto stat = [&](const std::string& nm, const double* a) {
    if (!a) return;
    if (!cub) {
      s.real(nm, mdims(ns), to_f(a, ns));
      return;
    }
    w.add_group(nm);
    size_t off = 0;
    for (int64_t q = 0; q < p.sensor_num_cuboids; ++q) {
      uint64_t e[3] = {1, 1, 1};
      for (int k = 0; k < dim; ++k)
        e[k] = p.sensor_mask_corners[q * 2 * dim + dim + k] - p.sensor_mask_corners[q * 2 * dim + k] + 1;
      const size_t m = (size_t)(e[0] * e[1] * e[2]);
      s.real(nm + "/" + std::to_string(q + 1), mdims(e[0], e[1], e[2]), to_f(a + off, m));
      off += m;
    }
  };
  auto gridf = [&](const std::string& nm, const float* a) {
    if (a) s.real(nm, grid, std::vector<float>(a, a + nco));
  };
  auto gridd = [&](const std::string& nm, const double* a) {
    if (a) s.real(nm, grid, to_f(a, nco));
  };
  series("p", o.p);
  stat("p_max", o.p_max);
  stat("p_min", o.p_min);
  stat("p_rms", o.p_rms);
  gridd("p_max_all", o.p_max_all);
  gridd("p_min_all", o.p_min_all);
  gridf("p_final", o.p_final);
  const char* un[3] = {"ux", "uy", "uz"};
  for (int d = 0; d < dim; ++d) {
    const std::string c = un[d];
    series(c, o.u[d]);
    stat(c + "_max", o.u_max[d]);
    stat(c + "_min", o.u_min[d]);
    stat(c + "_rms", o.u_rms[d]);
    gridd(c + "_max_all", o.u_max_all[d]);
    gridd(c + "_min_all", o.u_min_all[d]);
    gridf(c + "_final", o.u_final[d]);
    series(c + "_non_staggered", o.u_non_staggered[d]);
  }
  file_attrs(w, "output", "Output data created by kwhip (kspaceFirstOrder-HIP)");
  w.write(out);
}

template <class F>
int guarded(char* err, size_t errlen, F f) {
  try {
    f();
    kw_set_err(err, errlen, "");
    return KW_OK;
  } catch (const kwh5::Error& e) {
    kw_set_err(err, errlen, e.what());
    return KW_EINVAL;
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

}  // namespace

extern "C" {

int kw_save_problem(const kw_problem* prob, const char* path, char* err, size_t errlen) {
  return guarded(err, errlen, [&]() { save_problem(prob, path); });
}

int kw_run_file(const char* input_file, const char* output_file, uint32_t record_flags,
                int64_t record_start_index, char* err, size_t errlen) {
  return guarded(err, errlen,
                 [&]() { run_file(input_file, output_file, record_flags, record_start_index); });
}

}  // extern "C"
